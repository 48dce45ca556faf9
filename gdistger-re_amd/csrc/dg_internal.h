// dg_internal.h - context, error and timing plumbing shared by the library's
// translation units.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/distger_hip.h"

// Device-side walk queue and round statistics (one per context).
struct dg_walk_queue {
  unsigned long long next;    // next walk index to hand out (relative to walker_begin)
  unsigned long long cursor;  // tokens of walks longer than a slot, in the overflow region
  unsigned int flags;         // DG_QFLAG_*
  unsigned int max_len;       // longest walk
  unsigned int zero_terms;    // RE terms with p_i == 0 or q_i == 0
  unsigned int pad;
  double re;                  // relative entropy of the round
  unsigned long long draws;   // candidate draws (adjacency units read)
  unsigned long long moves;   // row-offset pairs read (walk starts + moves)
  unsigned long long tokens;  // tokens written
};

// Hot-vertex set for visit counting: at most kHotMax vertices.
constexpr uint32_t kHotMax = 4096;
constexpr uint16_t kNotHot = 0xffffu;

#define DG_QFLAG_TOKEN_OVERFLOW 1u
#define DG_QFLAG_LENGTH_LIMIT 2u

struct dg_timing_entry {
  uint64_t launches = 0;
  double total_ms = 0.0;
};

struct dg_pending_event {
  std::string name;
  hipEvent_t start;
  hipEvent_t stop;
};

struct dg_sgns_state;   // defined by the trainer translation unit
struct dg_walk_plan;    // defined by the walker translation unit
struct dg_pipe_state;   // defined by the pipeline translation unit
struct dg_tuple_state;  // defined by the tuple-interface translation unit
struct dg_dsgl_state;   // defined by the DSGL translation unit

struct dg_ctx {
  int device = 0;
  int num_cus = 0;
  hipStream_t stream = nullptr;
  std::string err;

  // graph (device-resident CSR)
  uint32_t v_num = 0;
  uint64_t m = 0;
  uint64_t* d_row_ptr = nullptr;
  uint32_t* d_nbr = nullptr;  // neighbour ids, reference order
  uint32_t* d_cn = nullptr;   // common-neighbour counts
  uint2* d_adj = nullptr;     // walker view {neighbour, bounce threshold}
  uint32_t* d_deg = nullptr;
  // the highest-degree vertices, counted per workgroup in LDS when visit
  // counts are taken (their global counters would be contended)
  uint16_t* d_hot_slot = nullptr;  // V: slot in [0, n_hot) or kNotHot
  uint32_t* d_hot_id = nullptr;    // n_hot: vertex of each slot
  uint32_t n_hot = 0;

  // walker buffers
  dg_walk_queue* d_queue = nullptr;
  dg_walk_queue* h_queue = nullptr;  // pinned mirror
  uint32_t* d_freq = nullptr;        // cumulative visit counts
  uint64_t freq_total = 0;           // sum of d_freq
  uint32_t* d_freq_round = nullptr;  // visits of the round being walked
  size_t freq_round_cap = 0;
  uint32_t* d_hot_part = nullptr;    // per pack workgroup: hot-vertex counts
  size_t hot_part_cap = 0;
  uint32_t* d_tok_raw = nullptr;     // tokens in completion order
  size_t tok_raw_cap = 0;
  uint64_t* d_walk_off = nullptr;    // per-walk offset into d_tok_raw
  uint32_t* d_walk_len = nullptr;
  size_t walk_cap = 0;
  size_t walk_len_cap = 0;
  uint32_t* d_tokens = nullptr;      // current round corpus, walk-id order
  size_t tokens_cap = 0;
  uint64_t* d_offsets = nullptr;     // n_walks + 1
  size_t offsets_cap = 0;
  uint32_t* d_spill = nullptr;       // per-lane overflow of the LDS path buffer
  size_t spill_cap = 0;
  uint64_t* d_scan_tmp = nullptr;    // block sums for scans
  size_t scan_tmp_cap = 0;
  double* d_term_table = nullptr;    // tabulated entropy terms for short walks
  double* d_re_terms = nullptr;      // v_num
  double* d_re_chunks = nullptr;
  size_t re_cap = 0;
  dg_walk_plan* wplan = nullptr;     // the round planned for the captured step
  uint64_t cur_walks = 0;
  uint64_t cur_tokens = 0;
  double prev_re = 0.0;
  uint32_t rounds_done = 0;

  // training corpus (concatenated rounds)
  uint32_t* d_train_tokens = nullptr;
  size_t train_tokens_cap = 0;
  uint64_t* d_train_offsets = nullptr;
  size_t train_offsets_cap = 0;
  uint64_t train_walks = 0;
  uint64_t train_tokens = 0;

  dg_sgns_state* sgns = nullptr;
  dg_tuple_state* tuples = nullptr;
  dg_pipe_state* pipe = nullptr;
  dg_dsgl_state* dsgl = nullptr;
  // this device's copy of the allocation-free entry point's scratch array
  // (its address from hipGetSymbolAddress, set when the context opens)
  float* noalloc_scratch = nullptr;

  // timing
  std::map<std::string, dg_timing_entry> timing;
  // inside a captured graph (events recorded there cannot be timed): device
  // wall-clock stamps, two per timed interval, named in stamp_names
  bool stamping = false;
  unsigned long long* d_stamps = nullptr;
  uint32_t stamps_cap = 0;  // intervals
  std::vector<std::string> stamp_names;
  double wall_khz = 0.0;
  std::vector<dg_pending_event> pending;
  std::vector<hipEvent_t> event_pool;
};

// Record an error message on the context and return `code`.
int dg_fail(dg_ctx* ctx, int code, const char* fmt, ...);

#define DG_HIP(ctx, expr)                                                            \
  do {                                                                               \
    hipError_t _e = (expr);                                                          \
    if (_e != hipSuccess)                                                            \
      return dg_fail((ctx), DG_ERR_HIP, "%s failed: %s (%s:%d)", #expr,              \
                     hipGetErrorString(_e), __FILE__, __LINE__);                     \
  } while (0)

#define DG_CHECK(expr)            \
  do {                            \
    int _rc = (expr);             \
    if (_rc != DG_OK) return _rc; \
  } while (0)

// Grow a device buffer to hold at least `need` elements (contents discarded).
template <typename T>
int dg_reserve(dg_ctx* ctx, T** ptr, size_t* cap, size_t need) {
  if (*cap >= need && *ptr) return DG_OK;
  if (*ptr) {
    hipError_t e = hipFree(*ptr);
    *ptr = nullptr;
    *cap = 0;
    if (e != hipSuccess) return dg_fail(ctx, DG_ERR_HIP, "hipFree: %s", hipGetErrorString(e));
  }
  size_t n = need ? need : 1;
  hipError_t e = hipMalloc((void**)ptr, n * sizeof(T));
  if (e != hipSuccess) {
    *ptr = nullptr;
    return dg_fail(ctx, DG_ERR_NOMEM, "hipMalloc(%zu bytes): %s", n * sizeof(T),
                   hipGetErrorString(e));
  }
  *cap = n;
  return DG_OK;
}

// Grow keeping the first `keep` elements.
template <typename T>
int dg_grow_keep(dg_ctx* ctx, T** ptr, size_t* cap, size_t need, size_t keep) {
  if (*cap >= need && *ptr) return DG_OK;
  size_t n = need > 2 * *cap ? need : 2 * *cap;
  T* np = nullptr;
  hipError_t e = hipMalloc((void**)&np, n * sizeof(T));
  if (e != hipSuccess)
    return dg_fail(ctx, DG_ERR_NOMEM, "hipMalloc(%zu bytes): %s", n * sizeof(T),
                   hipGetErrorString(e));
  if (*ptr && keep) {
    e = hipMemcpyAsync(np, *ptr, keep * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream);
    if (e != hipSuccess) return dg_fail(ctx, DG_ERR_HIP, "hipMemcpyAsync: %s", hipGetErrorString(e));
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return dg_fail(ctx, DG_ERR_HIP, "sync: %s", hipGetErrorString(e));
  }
  if (*ptr) hipFree(*ptr);
  *ptr = np;
  *cap = n;
  return DG_OK;
}

// Kernel timing with HIP events on the context stream.
int dg_time_begin(dg_ctx* ctx, const char* name, size_t* token);
int dg_time_end(dg_ctx* ctx, size_t token);
int dg_time_resolve(dg_ctx* ctx);
// Device-stamp timing for work that is being captured: begin (allocates; call
// before the capture), end (after it), resolve (after the graph has run: adds
// the intervals to the per-family totals and returns the walk and trainer sums).
int dg_stamps_begin(dg_ctx* ctx);
void dg_stamps_end(dg_ctx* ctx);
int dg_stamps_resolve(dg_ctx* ctx, double* walk_ms, double* sgns_ms);

// Graph teardown (graph.hip), trainer teardown (sgns.hip), tuple-interface
// teardown (tuples.hip).
void dg_graph_free(dg_ctx* ctx);
void dg_walk_free(dg_ctx* ctx);
void dg_sgns_free(dg_ctx* ctx);
void dg_tuples_free(dg_ctx* ctx);
// Resolve the device address of the allocation-free entry point's scratch
// array for ctx's device (once per device; the device must be current).
int dg_tuples_scratch_setup(dg_ctx* ctx);
// Forget the gradients of the last tuple backward (the tables they belong to
// are being replaced).
void dg_tuples_invalidate(dg_ctx* ctx);

// The trainer's FP32 tables as the tuple interface sees them.  `version`
// changes whenever the tables are written (init, set, train, tuple updates),
// so a derived copy (the BF16 shadow) knows when it is stale.
struct dg_table_view {
  float* syn0;
  float* syn1;
  uint32_t rows;
  uint32_t dim;
  uint64_t version;
};
int dg_sgns_view(dg_ctx* ctx, dg_table_view* v);
// Record a write to the tables; returns the new version.
uint64_t dg_sgns_written(dg_ctx* ctx);

// The walk round and the trainer epoch as the captured pipeline step runs
// them: plan (validation, launch shapes, buffers sized for any outcome; may
// allocate), enqueue (kernels only, no host waits), collect (after the
// stream has finished).  A collect that sets *retry found the round too large
// for its token buffer: the buffer has grown and nothing else changed.
int dg_walk_graph_plan(dg_ctx* ctx, const dg_walk_params* p, uint32_t round,
                       uint32_t walker_begin, uint32_t walker_num, uint64_t* tok_bound,
                       uint64_t* slot_tokens, uint64_t* overflow_tokens);
int dg_walk_graph_enqueue(dg_ctx* ctx, bool timed);
int dg_walk_graph_collect(dg_ctx* ctx, double walk_ms, double total_ms, dg_round_stats* out,
                          bool* retry);
// Whether the trainer can run as the captured step with these parameters
// (state, form, dim); checked before the step plans anything.
int dg_sgns_graph_check(dg_ctx* ctx, const dg_sgns_params* p);
int dg_sgns_graph_plan(dg_ctx* ctx, const dg_sgns_params* p, float alpha, uint64_t n_walks,
                       uint64_t tok_bound, uint64_t slot_tokens, uint64_t overflow_tokens);
int dg_sgns_graph_enqueue(dg_ctx* ctx, const uint32_t* d_tok, const uint64_t* d_off, bool timed);
int dg_sgns_graph_collect(dg_ctx* ctx, bool commit, double ms, dg_sgns_stats* out);
void dg_pipe_free(dg_ctx* ctx);

// One call of the DSGL trainer (dsgl.hip) over a device corpus, on the
// trainer's tables.
struct dg_dsgl_call {
  const uint32_t* tok;
  const uint64_t* off;
  uint64_t n_walks, n_tok;
  uint64_t max_walk;        // longest walk of the corpus
  const uint32_t* rank;
  uint32_t rank_size;       // entries of rank (vertex ids at or above it have no rank)
  const uint32_t* keep;     // DSGL keep limits per rank
  const int32_t* table;
  uint32_t table_size, vocab;
  const float* expt;
  float* syn0;
  float* syn1;
  uint32_t dim, window, K, iter, workers;
  float alpha0;
  uint64_t train_words, seed;
  uint32_t epoch0;
  int exact, deterministic;
};
int dg_dsgl_train(dg_ctx* ctx, const dg_dsgl_call& c, uint64_t* kept, double* ms, float* alpha_end);
void dg_dsgl_free(dg_ctx* ctx);

// Exclusive scan of n uint32 lengths into n+1 uint64 offsets (walker.hip).
int dg_scan_lengths(dg_ctx* ctx, const uint32_t* d_len, uint64_t n, uint64_t* d_off);
