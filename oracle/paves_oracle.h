/*
 * paves_oracle.h — CPU restatement of PAVES's query arithmetic.
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP engine; never linked by
 * the product). Types mirror engine_api.h:102-148 byte for byte.
 */
#ifndef PAVES_ORACLE_H
#define PAVES_ORACLE_H

#include <stdint.h>

enum { OR_CMP_LT, OR_CMP_LE, OR_CMP_EQ, OR_CMP_GE, OR_CMP_GT, OR_CMP_NE };
enum { OR_NODE_CMP, OR_NODE_IN, OR_NODE_AND, OR_NODE_OR, OR_NODE_NOT };

typedef struct OrPredNode { /* = FvGpuPredNode, engine_api.h:102-112 */
    int32_t kind, col, op, nchildren, child0, nvalues, val0, pad;
    double value;
} OrPredNode;

typedef struct OrHit { /* = FvHit, engine_api.h:145-148 */
    uint64_t node;
    float dist;
} OrHit;

float or_l2sq(const float *a, const float *b, int dim);
double or_l2sq_f64(const float *a, const float *b, int dim);
uint16_t or_bf16_bits(float x);
int or_pred_eval(const OrPredNode *nodes, const int32_t *children,
                 const double *invals, const double *cols, uint64_t count,
                 int32_t ni, uint32_t row);
void or_sort_hits(OrHit *h, uint32_t n);
uint64_t or_brute_topb(const float *vectors, uint64_t count, int dim,
                       const double *cols, const float *query,
                       const OrPredNode *nodes, const int32_t *children,
                       const double *invals, int32_t proot, uint32_t topb,
                       OrHit *out, uint32_t *out_cnt, uint8_t *truncated);
int or_brute_batch(const float *vectors, uint64_t count, int dim,
                   const double *cols, int nq, const float *queries,
                   const int32_t *proots, const OrPredNode *nodes,
                   const int32_t *children, const double *invals,
                   uint32_t topb, OrHit *out, uint32_t *out_cnt,
                   uint8_t *truncated, uint64_t *passing);
void or_sq8_params(const float *vectors, uint64_t count, int dim, float *mid, float *inv);
void or_sq8_codes(const float *x, uint64_t n, int dim, const float *mid, float inv,
                  int8_t *codes);
int or_sq8_brute_batch(const float *vectors, uint64_t count, int dim, const double *cols,
                       int nq, const float *queries, const int32_t *proots,
                       const OrPredNode *nodes, const int32_t *children,
                       const double *invals, uint32_t topb, uint32_t over, OrHit *out,
                       uint32_t *out_cnt, uint8_t *truncated);
uint8_t or_fp8_encode(float v);
float or_fp8_decode(uint8_t c);
void or_fp8_params(const float *vectors, uint64_t count, int dim, float *mid, float *scale);
void or_fp8_codes(const float *x, uint64_t n, int dim, const float *mid, float scale,
                  uint8_t *codes);
int or_fp8_brute_batch(const float *vectors, uint64_t count, int dim, const double *cols,
                       int nq, const float *queries, const int32_t *proots,
                       const OrPredNode *nodes, const int32_t *children,
                       const double *invals, uint32_t topb, uint32_t over, OrHit *out,
                       uint32_t *out_cnt, uint8_t *truncated);
uint32_t or_hnsw_search(const float *vectors, uint64_t count, int dim,
                        const uint32_t *l0, uint32_t l0_stride,
                        const double *cols, const OrPredNode *nodes,
                        const int32_t *children, const double *invals,
                        int32_t proot, const float *query, uint32_t entry,
                        uint32_t ef, uint64_t max_visited, OrHit *out,
                        uint8_t *exhausted, uint64_t *visited_out);
int or_hnsw_batch(const float *vectors, uint64_t count, int dim,
                  const uint32_t *l0, uint32_t l0_stride, const double *cols,
                  const OrPredNode *nodes, const int32_t *children,
                  const double *invals, int nq, const int32_t *proots,
                  const float *queries, const uint32_t *entries, uint32_t ef,
                  uint64_t max_visited, OrHit *out, uint32_t *out_cnt,
                  uint8_t *exhausted);
uint32_t or_graph_entry(const float *vectors, int dim,
                        const uint64_t *upper_offs, const uint32_t *upper,
                        uint32_t M, uint64_t entry_node, int max_level,
                        const float *query);
int or_hnsw_batch_insert(const float *X, uint64_t count, int dim, uint32_t M, uint32_t efc,
                         uint32_t *l0, const uint32_t *rest, const uint32_t *entries,
                         const int32_t *sizes, int nbatches);
double or_recall(const uint64_t *got, uint32_t ngot, const uint64_t *truth,
                 uint32_t ntruth);

#endif
