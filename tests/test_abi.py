"""The C-ABI library loads and exports every symbol its header declares; the
header's types have the reference byte layouts; the host glue builds the
batch arrays the way engine.cpp does. No GPU needed (no compute calls)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from paves_amd import _lib as L
from paves_amd import engine as E

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "paves_gpu.h")


def declared_functions():
    src = open(HDR).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fv_(?:gpu|hnsw)_[a-z0-9_]+|fv_l2sq)\s*\(", src)))


def exported_symbols(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if ln.strip()}


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert set(L.REF_SYMBOLS) <= set(names)
    assert set(L.EXT_SYMBOLS) <= set(names)
    exp = exported_symbols(L.LIB_PATH)
    missing = [n for n in names if n not in exp]
    assert not missing, missing
    lib = L.load()
    for n in names:
        assert hasattr(lib, n)


def test_reference_symbols_resolve_from_engine_cpp_build():
    """engine.cpp built with -DFV_GPU leaves exactly the reference's fv_gpu_*
    entry points undefined, and this library defines all of them."""
    ref = os.path.join(ROOT, "oracle", "_ref", "libpaves_ref_gpu.so")
    if not os.path.exists(ref):
        pytest.skip("oracle/_ref not built")
    out = subprocess.run(["nm", "-D", "--undefined-only", ref], capture_output=True, text=True,
                         check=True).stdout
    undef = {ln.split()[-1] for ln in out.splitlines() if "fv_gpu" in ln}
    assert undef, "engine.cpp -DFV_GPU should import the fv_gpu_* ABI"
    assert undef <= set(L.REF_SYMBOLS)
    assert undef <= exported_symbols(L.LIB_PATH)


def test_header_layouts_compile(tmp_path):
    c = tmp_path / "t.c"
    c.write_text('#include "paves_gpu.h"\n#include <stddef.h>\n'
                 "_Static_assert(offsetof(FvGpuQueryDesc, entry) == 20, \"entry\");\n"
                 "_Static_assert(offsetof(FvGpuQueryDesc, total) == 24, \"total\");\n"
                 "_Static_assert(offsetof(FvGpuPredNode, value) == 32, \"value\");\n"
                 "_Static_assert(offsetof(FvHit, dist) == 8, \"dist\");\n"
                 "int main(void){return 0;}\n")
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), "-c", str(c), "-o",
                    str(tmp_path / "t.o")], check=True)
    cxx = tmp_path / "t.cpp"
    cxx.write_text('#include "paves_gpu.h"\nint main(){return 0;}\n')
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-c", str(cxx),
                    "-o", str(tmp_path / "t2.o")], check=True)


def test_numpy_dtypes_match_header():
    assert L.PRED_DT.itemsize == 40 and L.PRED_DT.fields["value"][1] == 32
    assert L.QD_DT.fields["entry"][1] == 20 and L.QD_DT.fields["total"][1] == 24
    assert L.HIT_DT.itemsize == 16 and L.SPAN_DT.itemsize == 12


def test_no_device_paths_fail_cleanly():
    """Without a visible device the probes report none and upload returns
    NULL with a message (the caller's CPU-fallback contract)."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is visible")
    except ImportError:
        pass
    lib = L.load()
    assert lib.fv_gpu_available() == 0
    free = C.c_uint64()
    tot = C.c_uint64()
    assert lib.fv_gpu_mem_probe(C.byref(free), C.byref(tot)) == -1
    x = np.zeros((64, 16), np.float32)
    eb = L.errbuf()
    h = lib.fv_gpu_index_upload(x.ctypes.data, None, None, None, 0, 64, 16, 0, -1, 0, eb, len(eb))
    assert not h and b"gpu" in eb.value
    assert lib.fv_gpu_tile_variant_qt(128) == 64


def test_concat_rebases_like_gpuconcat():
    p1 = E.and_(E.cmp(0, E.CMP_LT, 5.0), E.isin(1, [1.0, 2.0]))
    p2 = E.or_(E.isin(1, [7.0]), E.not_(E.cmp(0, E.CMP_GE, 3.0)))
    s1, s2 = E.serialize(p1, 100), E.serialize(p2, 100)
    b = E.concat([s1, E.serialize(None, 100), s2], 100)
    assert list(b.qd["proot"]) == [0, -1, 3]
    assert list(b.qd["total"]) == [100, 100, 100]
    # children of query 2 are absolute node indices
    n2 = b.nodes[3]
    kids = b.children[n2["child0"]: n2["child0"] + n2["nchildren"]]
    assert list(kids) == [4, 5]
    assert b.nodes[4]["kind"] == E.NODE_IN and b.invals[b.nodes[4]["val0"]] == 7.0
    notn = b.nodes[5]
    assert b.children[notn["child0"]] == 6
