/*
 * pv_hnsw.h — the host builder's hooks for the device graph builder
 * (csrc/host/pv_hnsw.cpp <-> pv_build.hip). Host-only declarations.
 */
#ifndef PV_HNSW_H
#define PV_HNSW_H

#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "../../include/paves_gpu.h"

namespace pv {

FvHnswGraph *hnsw_seed(const float *vectors, uint64_t count, int dim, int M,
                       int ef_construction, int threads, std::vector<uint32_t> &rest,
                       std::vector<uint32_t> &entries, char *errbuf, size_t errlen,
                       double *t_entries_s = nullptr);
uint32_t *hnsw_l0(FvHnswGraph *h, uint32_t *stride);
void hnsw_detach(FvHnswGraph *h);

} // namespace pv

#endif
