/*
 * distger_hip.h - C ABI of the MI355X (gfx950) DistGER embedding hot path:
 * the information-oriented random walker and the skip-gram negative-sampling
 * trainer, with the CSR graph, the walk corpus and both embedding tables
 * resident in one GPU's HBM.
 *
 * Conventions
 *   - Every entry point returns DG_OK (0) or a negative DG_ERR_* code; the
 *     message of the last failure on a context is dg_last_error(ctx).  Nothing
 *     here calls exit() or throws across the boundary (the reference exits on
 *     CUDA errors, src/word2vec.cu:36-42, and asserts elsewhere).
 *   - The caller owns every host buffer; the library owns device memory behind
 *     the opaque dg_ctx.  A context is bound to one device; calls on one
 *     context must not overlap (the reference trainer keeps all state in
 *     globals, src/word2vec.cu:53-75, so it is not re-entrant at all).
 *   - Vertex ids are uint32 (include/type.hpp:15), CSR offsets uint64
 *     (edge_id_t, include/type.hpp:17).
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   GraphEngine::load_graph          include/graph.hpp:642      -> dg_graph_*
 *   WalkEngine::random_walk          include/walk.hpp:511-524   -> dg_walk_*
 *   PathSet::dump                    include/path.hpp:67-119    -> dg_corpus_write_text
 *   train_corpus_cuda                src/word2vec.cu:1429       -> dg_sgns_*
 *   dsgl (CPU trainer)               include/dsgl.hpp:1464      -> dg_sgns_* (batched forms)
 */
#ifndef DISTGER_HIP_H
#define DISTGER_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define DG_OK 0
#define DG_ERR_ARG -1          /* invalid argument */
#define DG_ERR_HIP -2          /* HIP runtime failure (message has the HIP error) */
#define DG_ERR_NOMEM -3        /* device or host allocation failed */
#define DG_ERR_IO -4           /* file missing, unreadable or truncated */
#define DG_ERR_STATE -5        /* call out of order (e.g. walk before a graph) */
#define DG_ERR_WALK_TOO_LONG -6 /* a walk reached the safety length limit */
#define DG_ERR_RE_UNDEFINED -7 /* relative entropy undefined: a vertex with degree or visits 0 */
#define DG_ERR_UNSUPPORTED -8  /* configuration not supported (e.g. embedding dim) */

typedef struct dg_ctx dg_ctx;

/* 12-byte binary edge record of the reference (Edge<float>,
 * include/type.hpp:64-70; file layout include/storage.hpp:15-38). */
typedef struct {
  uint32_t src;
  uint32_t dst;
  float data;
} dg_edge;

/* ---- context ----------------------------------------------------------- */

/* Open a context on HIP device `device`.  No reference counterpart (the
 * reference binds the device implicitly in train_corpus_cuda). */
int dg_open(int device, dg_ctx** out);
void dg_close(dg_ctx* ctx);
const char* dg_last_error(const dg_ctx* ctx);
/* Wait for all work the context queued.  hipDeviceSynchronize analogue. */
int dg_sync(dg_ctx* ctx);

/* ---- graph (W1-W3) ------------------------------------------------------ */

/* CN source for the walker's acceptance test. */
#define DG_CN_FROM_DATA 0  /* edge data holds the common-neighbour count, as the
                              reference's binary input does (gcn -> gconverter,
                              walk.hpp:920 reads candidate->data) */
#define DG_CN_COMPUTE 1    /* compute |N(u) & N(v) \ {u,v}| on the device */
#define DG_CN_COMPUTE_WEIGHTED 2 /* compute the counts on the device; the edge data
                                    is the weight w(u,v) and the walker accepts a
                                    move with tanh(pa * w(u,v)), the paper's weighted
                                    form (DistGER_Full_Version.pdf p.3, sec. 2.1; the
                                    reference code has no weighted walk) */

/* Build the device CSR from `e_num` edges.  Replaces GraphEngine::load_graph
 * (include/graph.hpp:642-824) for one partition: optional undirected doubling
 * [e0, rev e0, e1, rev e1, ...] (:671-686), degrees (:694-699), CSR in that
 * order, stable by source (:509-540). */
int dg_graph_from_edges(dg_ctx* ctx, uint32_t v_num, const dg_edge* edges, uint64_t e_num,
                        int undirected, int cn_mode);

/* Same, reading the reference binary edge file (include/storage.hpp:15-38). */
int dg_graph_load_bin(dg_ctx* ctx, const char* path, uint32_t v_num, int undirected,
                      int cn_mode);

/* Graph size: vertices and directed CSR entries. */
int dg_graph_info(dg_ctx* ctx, uint32_t* v_num, uint64_t* m);

/* Copy the device CSR back: row_ptr[v_num+1], nbr[m], cn[m] (any may be NULL). */
int dg_graph_get_csr(dg_ctx* ctx, uint64_t* row_ptr, uint32_t* nbr, uint32_t* cn);

/* Per-entry bounce thresholds thr[m]: the walker moving u -> v bounces back
 * (walk.hpp:938, tanh(pa) <= rand()/RAND_MAX) exactly when its 31-bit draw is
 * >= thr.  Exported for verification. */
int dg_graph_get_thresholds(dg_ctx* ctx, uint32_t* thr);

/* ---- walker (W4-W9) ---------------------------------------------------- */

/* Safety limit on vertices per walk when dg_walk_params.length_limit is 0.
 * The reference has no limit (walk.hpp:979-985); a walk that reaches it ends
 * the round with DG_ERR_WALK_TOO_LONG.  The walker keeps (limit - 23)
 * vertices of spill space per resident lane. */
#define DG_WALK_LENGTH_LIMIT_DEFAULT 1024

#define DG_ROUNDS_REFERENCE 0 /* stop after max_rounds rounds (walk.hpp:740-744) */
#define DG_ROUNDS_CONVERGE 1  /* also stop once |delta RE| <= delta (paper; walk.hpp:745-749) */

typedef struct {
  uint64_t seed;             /* Philox key for every walk draw */
  uint32_t walkers_per_round; /* -w (option_helper.hpp:142-180); rounds >= 2 run this many */
  uint32_t init_round;       /* --min_R: round 1 runs walkers_per_round * init_round (walk.hpp:614) */
  uint32_t min_length;       /* --min_L, default 20 (walk.hpp:404) */
  uint32_t max_length;       /* cap on vertices per walk, 0 = none (reference has none) */
  double mu;                 /* termination threshold on R^2, 0.999 (walk.hpp:24) */
  double delta;              /* RE convergence threshold, 0.001 (walk.hpp:26) */
  uint32_t max_rounds;       /* 30 (walk.hpp:740) */
  uint32_t rounds_mode;      /* DG_ROUNDS_* */
  uint32_t length_limit;     /* safety limit on vertices per walk; 0 = DG_WALK_LENGTH_LIMIT_DEFAULT */
  uint32_t reserved;
} dg_walk_params;

/* Fill `p` with the reference defaults (mu 0.999, delta 0.001, min_length 20,
 * init_round 5, max_rounds 30, no cap) for a graph of v_num vertices. */
void dg_walk_params_default(dg_walk_params* p, uint32_t v_num);

typedef struct {
  uint32_t round;        /* 1-based, the reference's `iter` (walk.hpp:645) */
  uint32_t walkers;      /* walks in this round */
  uint64_t tokens;       /* vertices written (footprints incl. step 0) */
  uint64_t steps;        /* appended vertices (footprints with step >= 1), the
                            reference's p_step (walk.hpp:397) */
  double re;             /* relative entropy after this round (walk.hpp:724) */
  double delta_re;       /* h_k - h_{k-1} (h_1 for round 1) (walk.hpp:728) */
  double walk_ms;        /* device time of the walker kernel */
  double total_ms;       /* device time of the whole round */
  uint32_t max_walk_len; /* longest walk of the round */
  uint32_t stop;         /* 1 if the round driver stops after this round */
  uint64_t draws;        /* candidate draws (8-byte adjacency units read) */
  uint64_t moves;        /* row-offset pairs read (walk starts + moves) */
} dg_round_stats;

/* Run one round: walks walker_begin .. walker_begin+walker_num-1 (start vertex
 * id mod V, walk.hpp:501-509), then pack the corpus in walk-id order, add the
 * visits to the cumulative per-vertex counts and compute the relative entropy.
 * `round` is the 1-based round number (part of every walk's Philox counter).
 * The round's corpus replaces the previous round's (dg_corpus_*). */
int dg_walk_round(dg_ctx* ctx, const dg_walk_params* p, uint32_t round, uint32_t walker_begin,
                  uint32_t walker_num, dg_round_stats* out);

/* Run the whole round loop of internal_random_walk (walk.hpp:608-782):
 * round 1 with walkers_per_round*init_round walks, later rounds with
 * walkers_per_round, until the stop rule.  `per_round` receives up to
 * max_stats entries; *n_rounds the number of rounds run.  If keep_corpus is
 * non-zero every round's corpus is appended to the device training corpus
 * (dg_corpus_append). */
int dg_walk_run(dg_ctx* ctx, const dg_walk_params* p, dg_round_stats* per_round,
                uint32_t max_stats, uint32_t* n_rounds, int keep_corpus);

/* Reset the cumulative per-vertex visit counts (start of a new walk run). */
int dg_walk_reset(dg_ctx* ctx);

/* Current round corpus: number of walks and tokens. */
int dg_corpus_info(dg_ctx* ctx, uint64_t* n_walks, uint64_t* n_tokens);

/* Copy the current round corpus: tokens[n_tokens], offsets[n_walks+1]. */
int dg_corpus_copy(dg_ctx* ctx, uint32_t* tokens, uint64_t* offsets);

/* Write the current round corpus as the reference walk file
 * (include/path.hpp:67-119): " %u" per vertex, "\n" per walk, walks in
 * `threads` segments (segment t = walks t, t+T, ...; 1 = id order). */
int dg_corpus_write_text(dg_ctx* ctx, const char* path, uint32_t threads);

/* Cumulative per-vertex visit counts (vertex_cn / context_map_freq,
 * include/path.hpp:91-95). */
int dg_vertex_freq(dg_ctx* ctx, uint32_t* freq);

/* Append the current round corpus to the device training corpus. */
int dg_corpus_append(dg_ctx* ctx);
/* Training corpus size (walks, tokens); and clear it. */
int dg_train_corpus_info(dg_ctx* ctx, uint64_t* n_walks, uint64_t* n_tokens);
int dg_train_corpus_clear(dg_ctx* ctx);

/* ---- skip-gram negative-sampling trainer (S1-S9) ------------------------ */

#define DG_SGNS_PLAIN_FP32 0   /* skip_gram_kernel semantics (word2vec.cu:305-399), FP32 */
#define DG_SGNS_BATCHED_FP32 1 /* per-position shared negatives, sentence pools on FP32 MFMA */
#define DG_SGNS_BATCHED_BF16 2 /* same with rows rounded to BF16 for BF16 MFMA, FP32 rows in HBM */
#define DG_SGNS_DSGL_FP32 3    /* DSGL, the reference's CPU trainer (dsgl.hpp:783-1427): two sentences
                                  in lockstep, a shared negative batch, [target, negatives with
                                  multiplicity, target2] from one snapshot per step; one wavefront
                                  per compute thread; DSGL's unigram table and subsampling rule */
#define DG_SGNS_NEGREUSE_FP32 4 /* __sgNegReuse (word2vec.cu:231-303): K negatives per sentence shared
                                   by all its positions, the full window, FP32 */

#define DG_VOCAB_DEGREE 0 /* word counts = in + out degree (word2vec.cu:791-809, huge_walk.cpp:43-46) */
#define DG_VOCAB_VISITS 1 /* word counts = walk visit counts (dsgl.hpp:314-382) */

#define DG_ALPHA_LINEAR 0 /* alpha0 * max(1e-4, 1 - words_done / (iter * words + 1)) (dsgl.hpp:886-897) */
#define DG_ALPHA_FIXED 1  /* alpha for the whole call (word2vec.cu:1290 sets it per walk file) */

typedef struct {
  uint32_t dim;            /* -size (multiple of 64, <= 512) */
  uint32_t window;         /* -window */
  uint32_t negative;       /* -negative (<= 16) */
  uint32_t min_count;      /* -min-count */
  uint32_t iter;           /* -iter (epochs over the corpus) */
  uint32_t batch_size;     /* -batch-size: DSGL's flag, forced to 2*window+1 (dsgl.hpp:832); recorded */
  float sample;            /* -sample */
  float alpha;             /* -alpha (starting learning rate) */
  uint32_t form;           /* DG_SGNS_* */
  uint32_t vocab_source;   /* DG_VOCAB_* */
  uint32_t alpha_schedule; /* DG_ALPHA_* */
  uint32_t deterministic;  /* 1: sentences one after another in one wavefront (a fixed
                              order of every update); 0: concurrent sentences with atomic
                              row updates (Hogwild, as the reference kernels) */
  uint32_t exact_sigmoid;  /* 0: the reference's 1000-entry sigmoid table; 1: 1/(1+e^-f) */
  uint32_t minibatch;      /* batched: sentences per launch (0 = 131072); sets how often the
                              linear learning rate is updated */
  uint32_t table_size;     /* unigram table entries (1e8 in the reference) */
  uint32_t init_float_ops; /* 0: word2vec.cu init rounding (double), 1: dsgl.hpp (float) */
  uint32_t negative_draws; /* batched: negatives drawn per position, >= negative (0 =
                              negative); each one's gradient is scaled by negative /
                              negative_draws, so the expected update is that of `negative` */
  uint32_t block_waves;    /* batched launch shape: sentences (wavefronts) per workgroup,
                              1, 2 or 4 (0 = 2; 4 runs as 2 where four pools exceed a CU's
                              LDS); does not change results */
  uint32_t pool_split;     /* batched launch shape: 16 = sentences of <= 16 distinct
                              vertices on a 16-row pool, the rest on 32 rows; 32 = every
                              sentence on the 32-row pool (0 = 16); does not change results */
  uint64_t seed;           /* Philox key for subsampling, window shrink, negatives */
  uint32_t workers;        /* DSGL form: compute threads the corpus is split over (0 = one per
                              resident wavefront, at most five per CU; dsgl.hpp:814-825 splits
                              over its threads) */
  uint32_t reserved;
} dg_sgns_params;

void dg_sgns_params_default(dg_sgns_params* p);

/* Build the vocabulary (ranked by count, ties by vertex id), initialise
 * syn0 (the reference LCG) and syn1 = 0, the unigram table and the sigmoid
 * table.  Replaces TrainModel's set-up (word2vec.cu:1246-1283) and
 * dsgl's preparatory_work (dsgl.hpp:511-519). */
int dg_sgns_init(dg_ctx* ctx, const dg_sgns_params* p);

typedef struct {
  uint64_t words;       /* corpus tokens consumed, subsampled ones included (word2vec.cu:1011-1020) */
  uint64_t kept;        /* tokens left after subsampling */
  uint64_t sentences;   /* walks */
  uint64_t launches;    /* trainer kernel launches */
  double ms;            /* device time of the training kernels */
  float alpha_end;      /* learning rate of the last launch */
  float reserved;
  uint64_t pool_rows;   /* batched: input rows held per sentence, summed (distinct
                           vertices of each sentence; rows loaded and written back once) */
} dg_sgns_stats;

#define DG_CORPUS_ROUND 0 /* the current round's corpus (one walk file of the reference) */
#define DG_CORPUS_TRAIN 1 /* the appended training corpus */

/* Train on a device corpus for p->iter epochs.  `alpha` overrides p->alpha
 * when > 0 (the reference sets a learning rate per walk file). */
int dg_sgns_train(dg_ctx* ctx, int corpus, const dg_sgns_params* p, float alpha,
                  dg_sgns_stats* out);

/* The learning rate of every kernel launch of the last dg_sgns_train (or
 * captured step) in launch order; *n receives the count, at most `cap` are
 * copied.  The linear schedule is dsgl.hpp:886-897's float expression,
 * alpha0 * (1 - words / (iter * corpus_words + 1)) clamped at alpha0 * 1e-4,
 * evaluated at each launch's first sentence (the reference steps it every
 * 10000 words).  Empty for the DSGL form (its rate steps inside the launch). */
int dg_sgns_last_alphas(dg_ctx* ctx, float* alphas, uint32_t cap, uint32_t* n);

/* Vocabulary: vertex id of each rank, and its size. */
int dg_sgns_vocab(dg_ctx* ctx, uint32_t* vertex_of_rank, uint32_t* vocab_size);

/* Copy or replace the FP32 (master) tables, rows in rank order, vocab x dim. */
int dg_sgns_get_tables(dg_ctx* ctx, float* syn0, float* syn1);
int dg_sgns_set_tables(dg_ctx* ctx, const float* syn0, const float* syn1);

/* Link-prediction scores: dot(syn0[u], syn0[v]) for vertex pairs (rank
 * lookup on the device; vertices outside the vocabulary score 0). */
int dg_sgns_score_pairs(dg_ctx* ctx, const uint32_t* src, const uint32_t* dst, uint64_t n,
                        float* scores);

/* Embedding file (dsgl.hpp:1443-1461): "V d\n", then per rank "id " and d
 * values "%f " (text) or d raw floats (binary), then "\n". */
int dg_emb_write(dg_ctx* ctx, const char* path, int binary);

/* Launch-shape autotuner for the batched forms: times the trainer on the
 * first sample_walks sentences of a device corpus (0 = 524288) for candidate
 * shapes (block_waves x pool_split, then minibatch) and writes the fastest
 * into *tuned (a copy of *p with those three fields set).  The tables and the
 * epoch counter (part of every draw's Philox key) are saved and restored:
 * tuning changes no training state.
 * words_per_s[0] is the default shape's rate, [1] the chosen one's (may be
 * NULL).  No reference counterpart: the reference's launch size is the
 * compile-time MAX_SENTENCE (src/word2vec.cu:35,1002-1040). */
int dg_sgns_autotune(dg_ctx* ctx, int corpus, const dg_sgns_params* p, uint64_t sample_walks,
                     dg_sgns_params* tuned, double* words_per_s);

/* ---- walk -> train as one HIP graph ------------------------------------
 * dg_walk_round(ctx, wp, round, walker_begin, walker_num, rs) followed by
 * dg_sgns_train(ctx, DG_CORPUS_ROUND, sp, alpha, ss), recorded by stream
 * capture and launched as one HIP graph: no host wait between the walk, the
 * packing and visit counts, the relative entropy and the trainer launches.
 * Same results as the two calls (the same bits in deterministic mode).
 * Batched trainer forms only (DG_ERR_UNSUPPORTED for the plain form).  Every
 * buffer is sized for any outcome of the round; a round longer than the
 * walker's token buffer is emptied on the device, the buffer grows and the
 * step runs again.  rs->total_ms is the whole step; rs->walk_ms and ss->ms the
 * walk and trainer kernels inside it.  Replaces the walker -> trainer hand-off
 * through walk files and a SyncQueue (include/type.hpp:29-61,
 * include/walk.hpp:686, src/word2vec.cu:1285-1300). */
int dg_pipeline_step(dg_ctx* ctx, const dg_walk_params* wp, uint32_t round, uint32_t walker_begin,
                     uint32_t walker_num, const dg_sgns_params* sp, float alpha,
                     dg_round_stats* rs, dg_sgns_stats* ss);
/* Graph bookkeeping of the context: instantiations, in-place updates, launches. */
int dg_pipeline_info(dg_ctx* ctx, uint64_t* instantiations, uint64_t* updates,
                     uint64_t* launches);

/* ---- differentiable tuple interface (forward / hand-written adjoint) ----
 * An autograd-Function-shaped pair over the tables of dg_sgns_init or
 * dg_sgns_tables_create.  A tuple is (center, context, neg_1..neg_K) in table
 * rows (ranks), tuples[n * (K + 2)].  x = syn0[context], y_0 = syn1[center],
 * y_k = syn1[neg_k], z_j = x . y_j and
 *   L = sum_t [ -log sig(z_0) - sum_k log sig(-z_k) ]   (exact sigmoid).
 * A negative equal to its tuple's centre is skipped, as skip_gram_kernel
 * skips it (src/word2vec.cu:366): its logit is reported, its loss term and
 * gradient are 0.  The reference's in-place step on the same pairs
 * (src/word2vec.cu:359-396) is W -= alpha * dL/dW one tuple at a time.
 * K <= 15; dim 64, 128 or 256; bf16 != 0 computes from BF16 copies of the
 * rows (FP32 products and sums; updates go to the FP32 rows).
 *
 * forward: logits[n * (1 + K)] (may be NULL) and the summed loss (FP64 sum
 * of FP32 terms in a fixed order; may be NULL). */
int dg_sgns_tuples_forward(dg_ctx* ctx, const uint32_t* tuples, uint64_t n, uint32_t K,
                           int bf16, float* logits, double* loss);
/* backward: dL/dz_j = sig(z_j) - label_j, or the caller's upstream gradient
 * dlogits[n * (1 + K)] when not NULL; row gradients dL/dW are summed per row
 * by a sorted-key segmented reduction in a fixed order (the result does not
 * depend on scheduling).  apply != 0 then updates W -= lr * dL/dW.  The
 * gradients stay readable through dg_sgns_grad_rows until the next backward.
 * n (K + 2) is at most 2^30 - 1 (DG_ERR_ARG beyond).  Should the sort's
 * bounded waits fail (a fault of the device, not of the input), the call
 * returns DG_ERR_HIP and leaves the tables unchanged. */
int dg_sgns_tuples_backward(dg_ctx* ctx, const uint32_t* tuples, uint64_t n, uint32_t K, int bf16,
                            const float* dlogits, float lr, int apply);
/* The forward and the backward with the loss's own gradient in one call:
 * the logits and loss of dg_sgns_tuples_forward and the row gradients of
 * dg_sgns_tuples_backward (dlogits = NULL), bit for bit, from one pass over
 * the rows instead of two; apply != 0 then updates W -= lr * dL/dW.  For
 * training loops whose loss is the tuple loss itself.  The reference has no
 * separate forward and backward: its skip_gram_kernel takes this step in
 * place, pair by pair (src/word2vec.cu:359-396).  logits[n * (1 + K)] and
 * loss may be NULL; the gradients stay readable through dg_sgns_grad_rows.
 * Limits and errors are those of the backward. */
int dg_sgns_tuples_step(dg_ctx* ctx, const uint32_t* tuples, uint64_t n, uint32_t K, int bf16, float lr,
                        int apply, float* logits, double* loss);
/* The forward without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid hipMalloc).  No counterpart in the reference.  The caller
 * passes device pointers: d_tuples[n * (K + 2)] (rows of the FP32 tables),
 * d_logits[n * (1 + K)] and d_loss (one double); the results are those of
 * dg_sgns_tuples_forward with bf16 = 0, bit for bit.  Working memory is a
 * fixed array in each device's memory (512 KiB: one loss partial per forward
 * wavefront, held until the process ends; devices of up to 512 CUs):
 * n is at most 1,048,576 (DG_ERR_ARG beyond).  The row ids are checked on the
 * device first (DG_ERR_ARG if one is out of range; this waits for the
 * context's stream once); the forward itself is queued on the context's
 * stream (dg_sync waits for it).  Calls on one device share that array:
 * calls that overlap on one device, from any contexts, race. */
int dg_sgns_tuples_forward_noalloc(dg_ctx* ctx, const uint32_t* d_tuples, uint64_t n, uint32_t K,
                                   float* d_logits, double* d_loss);
/* Gradient rows of the last backward: `table` 0 = syn0, 1 = syn1, rows by
 * rank; out[n_rows * dim] (rows that got no gradient read 0). */
int dg_sgns_grad_rows(dg_ctx* ctx, int table, const uint32_t* rows, uint64_t n_rows, float* out);

/* Tables without a vocabulary, for the tuple interface: `rows` x `dim` FP32
 * input (syn0) and output (syn1) rows; rank r is vertex r.  Replaces any
 * trainer state; training again needs dg_sgns_init.  No reference
 * counterpart: the reference builds its tables from a corpus
 * (src/word2vec.cu:844-878). */
int dg_sgns_tables_create(dg_ctx* ctx, uint32_t rows, uint32_t dim, const float* syn0,
                          const float* syn1);

/* ---- timing ------------------------------------------------------------ */

/* Per-kernel device time accumulated with HIP events on the context's stream
 * since the last reset.  `name` is a kernel family ("walk", "pack", "re",
 * "cn", "sgns_plain", "sgns_fwd", ...).  Returns launches and total ms. */
int dg_timing_get(dg_ctx* ctx, const char* name, uint64_t* launches, double* total_ms);
int dg_timing_reset(dg_ctx* ctx);

/* Streaming device-to-device copy of `bytes` (kernel dg_copy_kernel), `iters`
 * times; returns read+write GB/s.  Calibrates hardware byte counters of the
 * other kernels.  No reference counterpart. */
int dg_copy_bench(dg_ctx* ctx, uint64_t bytes, uint32_t iters, double* gbps);

/* The walker's FP64 functions evaluated on the device, for checking that
 * host and device round identically (the walk-termination and
 * convergence decisions depend on it): fn DG_MATH_SQRT, _LOG2, _LOG, _TANH
 * take x; DG_MATH_POW takes x and y.  out[i] = f(x[i] [, y[i]]).  No
 * reference counterpart. */
#define DG_MATH_SQRT 0
#define DG_MATH_LOG2 1
#define DG_MATH_LOG 2
#define DG_MATH_TANH 3
#define DG_MATH_POW 4
int dg_math_eval(dg_ctx* ctx, int fn, const double* x, const double* y, uint64_t n, double* out);

/* ---- host data utilities (synthetic inputs, file formats) -------------- */

/* Chung-Lu power-law graph: expected degrees w_i ~ (i+1)^(-1/(gamma-1))
 * scaled to 2*e_target and capped at sqrt(2*e_target); endpoints drawn
 * independently ~ w; self-loops and duplicates dropped; vertices below
 * min_degree topped up with extra draws.  Writes at most e_cap undirected
 * edges (src < dst) with data 0 and returns the count in *e_num. */
int dg_gen_chung_lu(uint32_t v_num, uint64_t e_target, double gamma, uint32_t min_degree,
                    uint64_t seed, dg_edge* edges, uint64_t e_cap, uint64_t* e_num);

/* Zipf(s) ids: out[i] in [0, v_num) with P(k) proportional to (k + 1)^-s,
 * by inverse CDF (double prefix sums, binary search); element i uses its own
 * Philox stream, so the output does not depend on the thread count. */
int dg_gen_zipf(uint32_t v_num, double s, uint64_t n, uint64_t seed, uint32_t* out);

/* N(0, sigma^2) floats by Box-Muller, one Philox block per pair of outputs. */
int dg_gen_normal(uint64_t n, double sigma, uint64_t seed, float* out);

/* R-MAT (Chakrabarti, Zhan, Faloutsos 2004) with 2^scale vertices and
 * edge_factor * 2^scale draws, quadrant probabilities (a, b, c, 1-a-b-c),
 * self-loops and duplicates dropped, weights uniform(0, 1] in `data`. */
int dg_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                uint64_t seed, dg_edge* edges, uint64_t e_cap, uint64_t* e_num);

/* Link-prediction split after src/tools/gtrain.cpp:226-386: move
 * round(frac * e_num) edges whose endpoints keep degree > 1 to the test set
 * (edges[] is permuted: train = edges[0 .. *n_train), positives follow), and
 * draw as many uniform non-edges (distinct, no self-pairs) into neg[]. */
int dg_lp_split(uint32_t v_num, dg_edge* edges, uint64_t e_num, double frac, uint64_t seed,
                uint64_t* n_train, dg_edge* neg, uint64_t neg_cap, uint64_t* n_neg);

/* Area under the ROC curve (Mann-Whitney form, ties counted 1/2). */
double dg_auc(const float* pos_scores, uint64_t n_pos, const float* neg_scores, uint64_t n_neg);

/* Reference binary edge file I/O (include/storage.hpp:15-38, :108-115). */
int dg_write_edges_bin(const char* path, const dg_edge* edges, uint64_t e_num);
int dg_read_edges_bin(const char* path, dg_edge* edges, uint64_t e_cap, uint64_t* e_num);

/* Text edge list -> reference binary edge file (src/tools/gconverter.cpp:
 * 45-76): weighted != 0 reads "%u %u %f" triples and writes 12-byte records
 * (Edge<real_t>, the input huge_walk reads); weighted == 0 reads "%u %u" pairs
 * and writes 8-byte records (Edge<EmptyData>: src, dst).  Reading stops at the
 * first entry that does not scan, as the reference's fscanf loop does.
 * *e_num receives the number of edges written. */
int dg_convert_text_edges(const char* in_path, const char* out_path, int weighted,
                          uint64_t* e_num);

/* The link-prediction files of src/tools/gtrain.cpp:345-384: the training
 * edges as "%u %u %u\n" (source, destination, the edge data as an unsigned
 * integer), the held-out pairs as "%u %u\n" (positives, then non-edges), and
 * their labels as "1\n" per positive and "0\n" per non-edge.  Any path may
 * be NULL to skip that file. */
int dg_lp_write_files(const char* train_path, const char* test_edges_path,
                      const char* test_labels_path, const dg_edge* train, uint64_t n_train,
                      const dg_edge* pos, uint64_t n_pos, const dg_edge* neg, uint64_t n_neg);

/* Learning-rate schedules of src/lr_scheduler.hpp:6-60, the value at a given
 * step (the reference's get_lr() returns the value at its step counter and
 * then advances it; the GPU trainer calls it once per walk file,
 * src/word2vec.cu:1279-1290):
 *   DG_LR_FIXED        initial
 *   DG_LR_EXPONENTIAL  initial * exp(-a * step)                (float, a = decay rate)
 *   DG_LR_COSINE       a + 0.5 (initial - a)(1 + cos(pi step / b))  (a = min lr, b = T_max)
 *   DG_LR_STEP         initial * a^(step / b)                  (a = factor, b = step size)
 * each in the reference's float/double mix.  Returns NaN for an unknown kind
 * or b <= 0 where b divides. */
#define DG_LR_FIXED 0
#define DG_LR_EXPONENTIAL 1
#define DG_LR_COSINE 2
#define DG_LR_STEP 3
float dg_lr_value(int kind, float initial, float a, float b, int step);

#ifdef __cplusplus
}
#endif

#endif /* DISTGER_HIP_H */
