"""MI355X-native filtered vector search query path (drop-in for PAVES's GPU engine).

The compute lives in ``libpaves_amd.so`` (HIP, gfx950) behind the C ABI of
``include/paves_gpu.h``; ``paves_amd.engine`` is the host-side glue that
builds the ABI's batch arrays.
"""
__all__ = ["engine", "synth"]
