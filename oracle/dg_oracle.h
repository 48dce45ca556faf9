/*
 * dg_oracle.h - CPU oracle for the DistGER embedding hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithms (cited per function, paths relative to the reference repository)
 * used to check the HIP library.  Only the test suite, the smoke check and the
 * CPU-baseline leg of the benchmark load it; the product library never links,
 * calls or falls back to it.
 *
 * Build variants (see oracle/Makefile):
 *   libdgoracle.so       elementary functions from dg_math.h (the same bits as
 *                        the GPU kernels);
 *   libdgoracle_libm.so  the same source with glibc pow/log2/sqrt/tanh/log, as
 *                        the reference calls them; used to count decisions that
 *                        differ between the two math libraries.
 */
#ifndef DG_ORACLE_H
#define DG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 12-byte binary edge record, include/type.hpp:64-70 + include/storage.hpp:15-38
 * (Edge<float>: u32 src, u32 dst, f32 data). */
typedef struct {
  uint32_t src;
  uint32_t dst;
  float data;
} dgo_edge;

typedef struct {
  uint32_t v_num;
  uint64_t m;              /* directed CSR entries */
  const uint64_t* row_ptr; /* v_num + 1 */
  const uint32_t* nbr;     /* m */
  const uint32_t* cn;      /* m, common-neighbour count per entry */
  const float* w;          /* m edge weights, or NULL: unweighted acceptance */
} dgo_graph;

typedef struct {
  uint64_t seed;
  uint32_t min_length; /* --min_L, include/walk.hpp:404,493-496 */
  uint32_t max_length; /* 0 = no cap (reference); otherwise cap on vertices per walk */
  double mu;           /* R threshold, include/walk.hpp:24 (0.999) */
} dgo_walk_cfg;

/* Graph build (W1/W2).  Returns 0, or -1 on an out-of-range vertex id.
 * row_ptr has v_num+1 entries; nbr/data have e_num*(undirected?2:1) entries
 * (either may be NULL: the edge data alone, in CSR order, is then the
 * per-entry weight array of a CSR built elsewhere in the same order). */
int dgo_build_csr(uint32_t v_num, const dgo_edge* edges, uint64_t e_num, int undirected,
                  uint64_t* row_ptr, uint32_t* nbr, float* data);

/* Common-neighbour count per CSR entry (W3): |(N(u) & N(v)) \ {u, v}| over
 * distinct vertices.  Returns 0. */
int dgo_common_neighbors(uint32_t v_num, const uint64_t* row_ptr, const uint32_t* nbr,
                         uint32_t* cn_out);

/* One walk (W5-W7).  Writes the vertex sequence to path (capacity cap) and, if
 * state_out is non-NULL, the final 9-double walker state
 * (H, S, EH, ES, EHS, EHH, ESS, R, N).  Returns the length, or -1 if cap is
 * too small. */
int64_t dgo_walk_one(const dgo_graph* g, const dgo_walk_cfg* cfg, uint32_t round,
                     uint32_t walk_id, uint32_t* path, uint64_t cap, double* state_out);

/* All walks of one round, ids walker_begin .. walker_begin+walker_num-1, in id
 * order, on `threads` OpenMP threads.  offsets has walker_num+1 entries.
 * Returns the token count, or -1 if tok_cap is too small. */
int64_t dgo_walk_round(const dgo_graph* g, const dgo_walk_cfg* cfg, uint32_t round,
                       uint32_t walker_begin, uint32_t walker_num, uint32_t* tokens,
                       uint64_t tok_cap, uint64_t* offsets, int threads);

/* Relative entropy of the degree distribution against the cumulative visit
 * frequencies (W8, include/walk.hpp:693-725), reduced in the fixed order
 * documented in the implementation.  *status = 0, or -1 if some p_i or q_i is
 * zero (the reference asserts there). */
double dgo_relative_entropy(uint32_t v_num, const uint32_t* deg, const uint32_t* freq,
                            int* status);

/* Walk text file (W9, include/path.hpp:67-119): for each walk, " %u" per vertex
 * then "\n"; walks ordered in `threads` segments (segment t holds ids
 * t, t+T, t+2T, ...; threads = 1 gives id order).  Returns 0 or -1 on I/O error. */
int dgo_write_walks(const char* path, const uint32_t* tokens, const uint64_t* offsets,
                    uint64_t nwalks, uint32_t threads);

/* Elementary functions as the oracle build uses them (exported for tests). */
double dgo_pow(double x, double y);
double dgo_log2(double x);
double dgo_log(double x);
double dgo_sqrt(double x);
double dgo_tanh(double x);

/* Philox stream (exported for tests). */
void dgo_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                uint32_t out[4]);
void dgo_stream_draws(uint64_t seed, uint32_t c1, uint32_t c2, uint32_t tag, uint32_t n,
                      uint32_t bound, uint32_t* out);

#ifdef __cplusplus
}
#endif

#endif /* DG_ORACLE_H */
