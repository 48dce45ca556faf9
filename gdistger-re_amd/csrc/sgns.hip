// sgns.hip - skip-gram negative-sampling trainer on gfx950.
//
// Reference: the word2vec.cu trainer behind train_corpus_cuda
// (src/word2vec.cu:1246-1416, kernels :231-399) and the batched DSGL trainer
// (include/dsgl.hpp:783-1427).  The walk corpus is consumed where the walker
// left it in HBM; no text files, no host batching.
//
// Forms
//   * plain FP32 (DG_SGNS_PLAIN_FP32): skip_gram_kernel's semantics - one
//     wavefront per sentence, per position a shrunk window, per (context,
//     target) pair one dot product, the table sigmoid and in-place updates.
//     A wavefront owns all dims of a row (dim/64 per lane); the dot is the
//     reference's block reduction order (pairs at distance 64, then a fold
//     by 32, 16, ..., 1), so a single-wavefront run in sentence order gives
//     the oracle's bits.
//   * batched (DG_SGNS_BATCHED_FP32 / _BF16): one wavefront per sentence.
//     Each position draws K negatives shared by the contexts of its window
//     (the per-position minibatch of DSGL, dsgl.hpp:1100-1159).  The
//     sentence's distinct vertices form a pool whose input rows stay in
//     registers for the whole sentence; per position the target and negative
//     rows are staged once and the (pool x outputs) logits and Gram products
//     run on MFMA.  The sequential order of the position's updates is
//     reproduced to first order in the updates (a recursion over the
//     contexts in order on the Gram matrices), so the result follows the
//     per-pair order of the plain form instead of a per-position snapshot.
//     Output-row updates are atomics; input rows are written back once per
//     sentence.  Deterministic mode runs the sentences in order in one
//     wavefront.
//   * BF16: FP32 master rows in HBM, rounded to BF16 while staged; products
//     on BF16 MFMA with FP32 accumulation; updates applied in FP32.
#include <algorithm>
#include <atomic>
#include <thread>
#include <cmath>
#include <cstring>
#include <vector>

#include "dg_internal.h"
#include "dg_philox.h"

namespace {

constexpr int kExpTable = 1000;
constexpr int kMaxExp = 6;
constexpr int kMaxNeg = 16;     // -negative
// -window: the plain form takes any window up to the reference's sentence
// length cap (MAX_SENTENCE_LENGTH, src/word2vec.cu:32); the batched forms
// at most kMaxSentWindow
constexpr int kMaxWindow = 1000;
constexpr uint32_t kNoRank = 0xffffffffu;

struct SgnsDev {
  const int32_t* sen;       // subsampled sentences (vocabulary ranks)
  const uint64_t* sen_off;  // n_sen + 1
  const int32_t* table;     // unigram table
  const float* expt;        // sigmoid table (kExpTable + 1)
  float* syn0;
  float* syn1;
  uint32_t table_size;
  uint32_t vocab;
  uint32_t window;
  uint32_t negative;
  float alpha;
  // the negative-reuse form (__sgNegReuse, word2vec.cu:231-303): `negative`
  // negatives per sentence of the epoch, shared by all its positions; null
  // for the plain form
  const int32_t* reuse_negs;
};

// ---- wave helpers --------------------------------------------------------
__device__ __forceinline__ float wave_tree_sum(float s) {
  // s[l] = s[l] + s[l + off] for off = 32 .. 1; lane 0 holds the reference
  // block-reduction result, which is then broadcast
  for (int off = 32; off >= 1; off >>= 1) s = s + __shfl_down(s, off, 64);
  return __shfl(s, 0, 64);
}

// Per-lane part of a dot product over dims lane + 64 j, j < D, folded by
// halving (f[x] += f[x + FSIZE] of word2vec.cu:373-376).
template <int D>
__device__ __forceinline__ float lane_partial(const float (&x)[D], const float (&y)[D]) {
  float p[D];
#pragma unroll
  for (int j = 0; j < D; ++j) p[j] = x[j] * y[j];
#pragma unroll
  for (int n = D; n > 1; n /= 2)
#pragma unroll
    for (int j = 0; j < n / 2; ++j) p[j] = p[j] + p[j + n / 2];
  return p[0];
}

// The table lookup reads at a clamped index and selects afterwards, so it
// has no branch: inside [-kMaxExp, kMaxExp] the index is the reference's.
__device__ __forceinline__ float sigmoid_of(float f, const float* expt, int exact) {
  if (exact) return (float)(1.0 / (1.0 + exp(-(double)f)));
  const float fc = fminf(fmaxf(f, -(float)kMaxExp), (float)kMaxExp);
  const float v = expt[(int)((fc + kMaxExp) * (kExpTable / kMaxExp / 2))];
  return f > kMaxExp ? 1.0f : (f < -kMaxExp ? 0.0f : v);
}

// ---- subsampling (word2vec.cu:1011-1020 with a Philox draw per token) ------
// The draw of corpus token i in epoch e is word 0 of Philox(counter (0, i lo,
// i hi, SAMPLE), key dg_epoch_key(seed, e)).
// A token whose vertex id is not below rank_size has no rank (skipped).
// The walk's tokens eight at a time: their ids, then their ranks, then the
// keep limits, each as eight independent loads in flight (a thread per walk).
constexpr int kSubChunk = 8;

struct SubChunk {
  uint32_t r[kSubChunk];   // rank, kNoRank for a token without one or past the walk
  bool kept[kSubChunk];
};

__device__ __forceinline__ void subsample_chunk(const uint32_t* tokens, uint64_t i0, uint64_t end,
                                                const uint32_t* rank, uint32_t rank_size,
                                                const uint32_t* keep, uint64_t key, SubChunk& c) {
  uint32_t v[kSubChunk], kp[kSubChunk];
#pragma unroll
  for (int u = 0; u < kSubChunk; ++u) v[u] = i0 + u < end ? tokens[i0 + u] : 0xffffffffu;
#pragma unroll
  for (int u = 0; u < kSubChunk; ++u) c.r[u] = v[u] < rank_size ? rank[v[u]] : kNoRank;
#pragma unroll
  for (int u = 0; u < kSubChunk; ++u) kp[u] = c.r[u] != kNoRank ? keep[c.r[u]] : 0u;
#pragma unroll
  for (int u = 0; u < kSubChunk; ++u) {
    const uint64_t i = i0 + u;
    const dg_u32x4 o = dg_philox4x32_10(0, (uint32_t)i, (uint32_t)(i >> 32), DG_TAG_SAMPLE,
                                        (uint32_t)key, (uint32_t)(key >> 32));
    c.kept[u] = c.r[u] != kNoRank && (o.x0 & 0xFFFFu) <= kp[u];
  }
}

__global__ void subsample_count_kernel(const uint32_t* tokens, const uint64_t* offsets,
                                       uint64_t n_walks, const uint32_t* rank, uint32_t rank_size,
                                       const uint32_t* keep, uint64_t seed, uint32_t epoch,
                                       uint32_t* kept) {
  const uint64_t key = dg_epoch_key(seed, epoch);
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < n_walks;
       w += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t c = 0;
    const uint64_t end = offsets[w + 1];
    for (uint64_t i0 = offsets[w]; i0 < end; i0 += kSubChunk) {
      SubChunk ch;
      subsample_chunk(tokens, i0, end, rank, rank_size, keep, key, ch);
#pragma unroll
      for (int u = 0; u < kSubChunk; ++u) c += ch.kept[u] ? 1u : 0u;
    }
    kept[w] = c;
  }
}

__global__ void subsample_write_kernel(const uint32_t* tokens, const uint64_t* offsets,
                                       uint64_t n_walks, const uint32_t* rank, uint32_t rank_size,
                                       const uint32_t* keep, uint64_t seed, uint32_t epoch,
                                       const uint64_t* sen_off, int32_t* sen) {
  const uint64_t key = dg_epoch_key(seed, epoch);
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < n_walks;
       w += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t o_pos = sen_off[w];
    const uint64_t end = offsets[w + 1];
    for (uint64_t i0 = offsets[w]; i0 < end; i0 += kSubChunk) {
      SubChunk ch;
      subsample_chunk(tokens, i0, end, rank, rank_size, keep, key, ch);
#pragma unroll
      for (int u = 0; u < kSubChunk; ++u)
        if (ch.kept[u]) sen[o_pos++] = (int32_t)ch.r[u];
    }
  }
}

// ---- plain form: skip_gram_kernel (word2vec.cu:305-399), negative sampling --
// One wavefront per sentence; lane l owns dims l + 64 j.  `single` runs the
// whole batch in one wavefront in sentence order (the deterministic mode).
// REUSE: the negative-reuse form (a separate instance, so the plain form's
// code is unchanged by it).
template <int D, bool REUSE>
__global__ __launch_bounds__(256) void sg_plain_kernel(SgnsDev a, uint64_t s_begin, uint64_t s_end,
                                                       int single, int exact) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = single ? 1 : ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t s = s_begin + wave; s < s_end; s += nwaves) {
    const uint64_t sb = a.sen_off[s], se = a.sen_off[s + 1];
    unsigned long long next_random = (unsigned long long)(s - s_begin);  // blockIdx.x
    const int32_t* rn = REUSE ? a.reuse_negs + s * a.negative : nullptr;
    for (uint64_t pos = sb; pos < se; ++pos) {
      const int word = a.sen[pos];
      int b = 0;  // the negative-reuse form trains the full window (word2vec.cu:257)
      if constexpr (!REUSE) {
        next_random = next_random * 2514903917ULL + 11;  // word2vec.cu:321 (sic)
        b = (int)(next_random % a.window);
      }
      for (int aa = b; aa < (int)a.window * 2 + 1 - b; ++aa) {
        if (aa == (int)a.window) continue;
        const int64_t c = (int64_t)pos - (int64_t)a.window + aa;
        if (c < (int64_t)sb || c >= (int64_t)se) continue;
        float* xrow = a.syn0 + (uint64_t)a.sen[c] * (64 * D);
        float x[D], neu1e[D];
#pragma unroll
        for (int j = 0; j < D; ++j) {
          x[j] = xrow[lane + 64 * j];
          neu1e[j] = 0.0f;
        }
        for (uint32_t d = 0; d < a.negative + 1; ++d) {
          int target, label;
          if (d == 0) {
            target = word;
            label = 1;
          } else {
            if constexpr (REUSE) {
              target = rn[d - 1];  // the sentence's shared negatives (word2vec.cu:247-249)
            } else {
              next_random = next_random * 25214903917ULL + 11;
              target = a.table[(next_random >> 16) % a.table_size];
              if (target == 0) target = (int)(next_random % (a.vocab - 1) + 1);
            }
            if (target == word) continue;
            label = 0;
          }
          float* yrow = a.syn1 + (uint64_t)target * (64 * D);
          float y[D];
#pragma unroll
          for (int j = 0; j < D; ++j) y[j] = yrow[lane + 64 * j];
          const float f = wave_tree_sum(lane_partial<D>(x, y));
          float g;
          if (exact) {
            g = ((float)label - sigmoid_of(f, a.expt, 1)) * a.alpha;
          } else if (f > kMaxExp) {
            g = (float)(label - 1) * a.alpha;
          } else if (f < -kMaxExp) {
            g = (float)(label - 0) * a.alpha;
          } else {
            g = ((float)label - a.expt[(int)((f + kMaxExp) * (kExpTable / kMaxExp / 2))]) * a.alpha;
          }
#pragma unroll
          for (int j = 0; j < D; ++j) {
            neu1e[j] += g * y[j];
            atomicAdd(&yrow[lane + 64 * j], g * x[j]);
          }
        }
#pragma unroll
        for (int j = 0; j < D; ++j) atomicAdd(&xrow[lane + 64 * j], neu1e[j]);
      }
    }
  }
}

// ---- batched form: sentence-resident pools on MFMA ---------------------------
// One wavefront trains one sentence, position after position, in the order
// of skip_gram_kernel (word2vec.cu:305-399), with the K negatives of a
// position drawn once and shared by all its contexts (the per-position
// negatives of dsgl.hpp:1100-1159).
//
// The distinct vertices of the sentence (walks revisit vertices: a
// 22-vertex walk holds about 12) form a pool whose input rows live in
// registers for the whole sentence: the row as loaded (E) plus an FP32
// delta that the update product accumulates into directly, so every
// position reads the input rows as the earlier positions left them.  Per
// position, with X the pool rows, Y the target and negative rows (staged in
// LDS from HBM) and the contexts u_1 < u_2 < ... of the shrunk window:
//   S = X Y^T,  Gx = X X^T,  Gy = Y Y^T                        (MFMA)
//   per output column n (lanes 0..K), over the contexts in order:
//     s_i = S[u_i][n] + sum_{j<i} g_j[n] Gx[u_i][u_j]
//                     + sum_{j<i, u_j = u_i} sum_n' g_j[n'] Gy[n'][n]
//     g_i[n] = alpha w_n (label_n - sig(s_i))
//   Gp[u][n] = sum_{i: u_i = u} g_i[n]
//   H[u][v]  = sum_{j<i, u_i = u, u_j = v} g_i . g_j
//   M[n][n'] = sum_{j<i, u_j = u_i} g_i[n] g_j[n']
//   dX = Gp Y + H X (into the register deltas),
//   dY = Gp^T X + M Y (atomics to HBM)                          (MFMA)
// This is the sequential order to first order in the updates: the second
// term of s_i is the output row as updated by the earlier contexts of the
// position, the third the context row as updated by its own earlier
// occurrences in the window; H X is the part of a context row's update that
// comes from output rows already moved by earlier contexts (all of it while
// the output rows are still zero), M Y the part of an output row's update
// that comes from context rows already moved.  Changes of changes (and the
// coupling of two equal negatives of one position) are dropped.  Output rows are read
// from HBM at each position, so a row updated by one position is seen by
// the next ones.  The deltas are added to the FP32 input rows when the
// sentence (or a segment of it, if its windows hold more distinct vertices
// than the pool) is done.
//
// Rows live in HBM as FP32 only; the BF16 form rounds them to BF16 as they
// are staged and runs v_mfma_f32_16x16x16_bf16, the FP32 form runs
// v_mfma_f32_16x16x4f32, both accumulating in FP32.  Every staged row is
// also updated, so a BF16 copy in HBM would need a refresh (4 B read + 2 B
// written per element) for each 2 B it saves on the gather.
typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4_t;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4_t;
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

// Two floats rounded to BF16 (nearest, ties to even) and packed, a in the low
// half: one v_cvt_pk_bf16_f32.  For every finite float it gives the bits of
// (u + 0x7fff + ((u >> 16) & 1)) >> 16 on the float's bits u.
__device__ __forceinline__ uint32_t f2bf2(float a, float b) {
  const f32x2_t v = {a, b};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_t));
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

constexpr int kOutCols = 16;    // output columns: the target and up to 15 negatives
constexpr int kSpan = 64;       // positions of a segment's windows held at once
constexpr int kMaxSentWindow = 15;
constexpr int kMaxSlots = 2 * kMaxSentWindow;  // contexts of one position

// A 16 x 16 x 16 product over four-element k chunks.  For one MFMA call,
// lane l supplies chunk element `e` of k index 4 (l/16) + e of row or column
// l % 16; the caller passes the four k values of a chunk in order.  The FP32
// form issues four 16x16x4 steps (k = 4 (l/16) + e, e = 0..3), the BF16 form
// one 16x16x16 step; both sum the same products.
template <typename E>
struct Chunk;

template <>
struct Chunk<float> {
  typedef float4 v_t;  // four k values
  __device__ static f32x4_t mma(v_t a, v_t b, f32x4_t c) {
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.x, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b.y, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a.z, b.z, c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a.w, b.w, c, 0, 0, 0);
  }
  __device__ static v_t from(float a, float b, float c, float d) { return make_float4(a, b, c, d); }
  // column l16 of the four rows from blk (row stride LDN): k value e from row e
  template <int LDN>
  __device__ static v_t col4(const float* blk, int l16) {
    return make_float4(blk[l16], blk[l16 + LDN], blk[l16 + 2 * LDN], blk[l16 + 3 * LDN]);
  }
  __device__ static void put4(float* dst, float a, float b, float c, float d) {
    *reinterpret_cast<float4*>(dst) = make_float4(a, b, c, d);
  }
  // four consecutive stored elements
  __device__ static v_t load4(const float* src) { return *reinterpret_cast<const float4*>(src); }
};

template <>
struct Chunk<uint16_t> {
  union U {
    uint2 q;
    uint16_t h[4];
    bf16x4_t v;
  };
  typedef bf16x4_t v_t;
  __device__ static f32x4_t mma(v_t a, v_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
  }
  __device__ static v_t from(float a, float b, float c, float d) {
    U u;
    u.q = make_uint2(f2bf2(a, b), f2bf2(c, d));
    return u.v;
  }
  // column l16 of the four rows from blk (row stride LDN, 8-byte aligned) by
  // one transposing LDS read: in each group of 16 lanes, lane j supplies the
  // 8 bytes of row j / 4, columns 4 (j % 4) .. 4 (j % 4) + 3, and lane i
  // receives column i of the four rows
  template <int LDN>
  __device__ static v_t col4(const uint16_t* blk, int l16) {
    typedef short s16x4_t __attribute__((ext_vector_type(4)));
    const uint16_t* src = blk + (l16 >> 2) * LDN + 4 * (l16 & 3);
    const s16x4_t r = __builtin_amdgcn_ds_read_tr16_b64_v4i16((__attribute__((address_space(3))) s16x4_t*)src);
    return __builtin_bit_cast(v_t, r);
  }
  __device__ static void put4(uint16_t* dst, float a, float b, float c, float d) {
    *reinterpret_cast<uint2*>(dst) = make_uint2(f2bf2(a, b), f2bf2(c, d));
  }
  __device__ static v_t load4(const uint16_t* src) {
    U u;
    u.q = *reinterpret_cast<const uint2*>(src);
    return u.v;
  }
};

// word j of a Philox stream (block j/4, word j%4)
__device__ __forceinline__ uint32_t stream_word(uint64_t seed, uint32_t c1, uint32_t c2,
                                                uint32_t tag, uint32_t j) {
  const dg_u32x4 o = dg_philox4x32_10(j >> 2, c1, c2, tag, (uint32_t)seed, (uint32_t)(seed >> 32));
  const uint32_t r = j & 3u;
  return r == 0 ? o.x0 : (r == 1 ? o.x1 : (r == 2 ? o.x2 : o.x3));
}

struct DrawArgs {
  const int32_t* sen;
  const uint64_t* sen_off;  // of the epoch's sentences
  uint64_t s_begin, s_end;  // sentences of the launch (indices in the epoch's corpus)
  uint64_t seed;
  uint32_t epoch;
  const int32_t* table;
  uint32_t table_size;
  uint32_t vocab;
  uint32_t window;
  int kd;
  uint8_t* shrink;  // per kept position of the launch
  int32_t* negs;    // kd per kept position of the launch
  uint8_t* ucount;  // per sentence of the launch: distinct vertices, 255 if above 64
  unsigned long long* pool_rows;  // sum of the distinct vertices (sentences of L >= 2)
  uint32_t* wq;     // the launch's two sentence counters (16- and 32-row pools), zeroed here
};

// Draws of every position: the stream of position q of sentence s of the
// epoch's corpus (counter (blk, s, q, SGNS), key dg_epoch_key(seed, epoch)):
// word 0 gives b = word % window, word k negative k: table[word % table_size],
// 0 -> word % (vocab-1) + 1.
__global__ void sent_draw_kernel(DrawArgs a) {
  const uint64_t base = a.sen_off[a.s_begin];
  const uint64_t n = a.sen_off[a.s_end] - base;
  const int32_t* __restrict__ table = a.table;
  int32_t* __restrict__ negs = a.negs;
  const uint64_t key = dg_epoch_key(a.seed, a.epoch);
  const uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    // sentence of kept position base + i
    uint64_t lo = a.s_begin, hi = a.s_end;
    while (hi - lo > 1) {
      const uint64_t mid = (lo + hi) / 2;
      if (a.sen_off[mid] <= base + i) lo = mid; else hi = mid;
    }
    const uint32_t q = (uint32_t)(base + i - a.sen_off[lo]);
    const uint32_t c1 = (uint32_t)lo, c2 = q;
    // word j of the stream is word j % 4 of Philox block j / 4 (stream_word);
    // each block is evaluated once, and every table read is issued before
    // any result is stored
    uint32_t wd[kOutCols];
    dg_u32x4 blk = dg_philox4x32_10(0, c1, c2, DG_TAG_SGNS, k0, k1);
    a.shrink[i] = (uint8_t)(blk.x0 % a.window);
    int v[kOutCols];
#pragma unroll
    for (int k = 0; k < kOutCols - 1; ++k) {
      if (k >= a.kd) break;
      const uint32_t j = (uint32_t)(k + 1), r = j & 3u;
      if (r == 0) blk = dg_philox4x32_10(j >> 2, c1, c2, DG_TAG_SGNS, k0, k1);
      wd[k] = r == 0 ? blk.x0 : (r == 1 ? blk.x1 : (r == 2 ? blk.x2 : blk.x3));
      v[k] = table[wd[k] % a.table_size];
    }
#pragma unroll
    for (int k = 0; k < kOutCols - 1; ++k) {
      if (k >= a.kd) break;
      negs[i * a.kd + k] = v[k] == 0 ? (int)(wd[k] % (a.vocab - 1) + 1) : v[k];
    }
  }
}

// Distinct vertices per sentence (255 when the sentence is longer than 64).
__global__ void sent_count_kernel(DrawArgs a) {
  // the sentence kernels that follow on the stream take their last sentences
  // from these counters
  if (blockIdx.x == 0 && threadIdx.x < 2) a.wq[threadIdx.x] = 0u;
  unsigned long long acc = 0;  // pool rows of this thread's sentences
  for (uint64_t s = a.s_begin + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < a.s_end;
       s += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t sb = a.sen_off[s];
    const int L = (int)(a.sen_off[s + 1] - sb);
    int u = 255;
    if (L <= 32) {
      // the common case: the tokens once into registers, the pairs compared there
      int tk[32];
#pragma unroll
      for (int i = 0; i < 32; ++i) tk[i] = i < L ? a.sen[sb + i] : -1 - i;
      u = 0;
#pragma unroll
      for (int i = 0; i < 32; ++i) {
        bool seen = false;
#pragma unroll
        for (int j = 0; j < i; ++j) seen |= tk[j] == tk[i];
        u += (i < L && !seen) ? 1 : 0;
      }
    } else if (L <= 64) {
      u = 0;
      for (int i = 0; i < L; ++i) {
        const int v = a.sen[sb + i];
        bool seen = false;
        for (int j = 0; j < i; ++j) seen |= a.sen[sb + j] == v;
        u += !seen;
      }
    }
    a.ucount[s - a.s_begin] = (uint8_t)u;
    if (L >= 2) {
      // longer sentences: count them here as well
      int d = u;
      if (u == 255) {
        d = 0;
        for (int i = 0; i < L; ++i) {
          const int v = a.sen[sb + i];
          bool seen = false;
          for (int j = 0; j < i && !seen; ++j) seen = a.sen[sb + j] == v;
          d += !seen;
        }
      }
      acc += (unsigned long long)d;
    }
  }
  // one atomic per wavefront on the shared counter, not one per sentence
  for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0 && acc) atomicAdd(a.pool_rows, acc);
}

// A wave-uniform 64-bit value moved to scalar registers.
__device__ inline uint64_t uniform_u64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

__global__ void zero_u64_kernel(unsigned long long* p) {
  if (threadIdx.x == 0) *p = 0ull;
}

// The linear learning rate of dsgl.hpp:886-897 in its float arithmetic,
// alpha0 * (1 - words / (iter * train_words + 1)) clamped at alpha0 * 1e-4,
// with `words` the corpus words consumed before the launch (the current
// epoch's and every earlier epoch's of the call) and train_words the call's
// corpus words per epoch; host and device evaluate the same expression.
__host__ __device__ __forceinline__ float linear_alpha(float alpha0, uint64_t words,
                                                       uint64_t total_plus_one) {
  float al = alpha0 * (1.0f - (float)words / (float)total_plus_one);
  if (al < alpha0 * 0.0001f) al = alpha0 * 0.0001f;
  return al;
}

// Learning rate of every launch of an epoch from the device corpus offsets
// (the captured step does not know them on the host).
__global__ void alpha_kernel(const uint64_t* woff, uint64_t n_walks, uint64_t mb, uint32_t n_mb,
                             float alpha0, int linear, uint32_t iter, uint32_t ep, float* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_mb) return;
  if (!linear) {
    out[i] = alpha0;
    return;
  }
  const uint64_t n_tok = woff[n_walks] - woff[0];
  const uint64_t words = (uint64_t)ep * n_tok + (woff[(uint64_t)i * mb] - woff[0]);
  out[i] = linear_alpha(alpha0, words, (uint64_t)iter * n_tok + 1);
}

// Negatives of the negative-reuse form: `negative` per sentence of the epoch,
// drawn as the reference's host loop draws them (word2vec.cu:1057-1062):
// table[x % table_size], 0 -> x % (vocab - 1) + 1, with x word k of the
// Philox stream (counter (k / 4, sentence, 0, REUSE), key of the epoch).
__global__ void reuse_negs_kernel(uint64_t n_sen, uint32_t K, uint64_t key, const int32_t* table,
                                  uint32_t table_size, uint32_t vocab, int32_t* out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_sen * K;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t s = i / K;
    const uint32_t k = (uint32_t)(i - s * K);
    const uint32_t x = stream_word(key, (uint32_t)s, 0u, DG_TAG_SGNS_REUSE, k);
    const int v = table[x % table_size];
    out[i] = v == 0 ? (int)(x % (vocab - 1) + 1) : v;
  }
}

// Adds an epoch's kept-token count (the last subsampled offset) to *acc.
__global__ void kept_add_kernel(const uint64_t* sen_off, uint64_t n_walks, unsigned long long* acc) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *acc += (unsigned long long)sen_off[n_walks];
}

struct SentArgs {
  const int32_t* sen;
  const uint64_t* sen_off;
  uint64_t s_begin, s_end;
  const uint8_t* shrink;
  const int32_t* negs;
  const uint8_t* ucount;
  const float* expt;
  float* syn0;
  float* syn1;
  float alpha;
  const float* alpha_dev;  // when set, the learning rate is read here (device schedule)
  float wneg;
  int window;
  int kd;
  int exact;
  int deterministic;
  int all;  // take every sentence (else only those of this pool size)
  int block_waves;  // launch shape (host side): wavefronts per workgroup, 1, 2 or 4
  int pool_split;   // launch shape (host side): 16 = small sentences on the 16-row pool; 32 = all on 32 rows
  uint32_t* wq;     // this launch's sentence counter (zero at the launch)
};

template <typename E, int DIM, int P, int KC>
struct alignas(16) SentLds {
  static constexpr int LD = DIM + 16 / (int)sizeof(E);
  static constexpr int GXL = P + 4;   // Gx, H rows (16-byte aligned)
  static constexpr int GPL = KC + 4;  // Gp, M rows (16-byte aligned)
  E xc[P * LD];              // pool rows as used at this position
  E yb[KC * LD];             // output rows (target, negatives), rows past K zero
  // Each pair below shares storage: the first is read only up to the end of
  // a position's recursion, the second written only after it.
  union {
    float S[P * (KC + 1)];   // S[u][n]
    float Gp[P * GPL];       // Gp[u][n]
  };
  union {
    float Gx[P * GXL];       // Gx[u][v]
    float H[P * GXL];        // H[u][v] = sum over context pairs j < i (v_i = u, v_j = v) of g_i . g_j
  };
  union {
    float Gy[KC * (KC + 1)]; // Gy[n][n']
    float Mm[KC * GPL];      // M[n][n'] = sum over pairs j < i of one vertex of g_i[n] g_j[n']
  };
  float Gs[kMaxSlots * KC];  // g_i[n] by context slot
  int su[kMaxSlots];         // pool index of context slot i
  uint8_t p2p[kSpan];        // position - lo -> pool index
};

// P = 16: sentences of at most 16 distinct vertices (one segment each);
// P = 32: every other sentence, in segments whose windows hold <= 32.
// KC output columns: 8 (up to 7 negatives) or 16.
// Wavefronts per SIMD the register allocation aims at.  The 16-row pools of
// the small shapes fit three (a few values spill to scratch; the third
// wavefront hides more of each position's dependency chain than the spills
// cost: 18% faster on the 2.2M-vertex corpus); the larger ones two.  The
// 32-row BF16 pools with 8 output columns up to dim 128 fit two in 256
// registers without spilling; the other 32-row pools one.
template <typename E, int DIM, int P, int KC>
constexpr int sent_waves_per_eu() {
  if (P != 16) return (P == 32 && KC == 8 && sizeof(E) == 2 && DIM <= 128) ? 2 : 1;
  if (KC == 8 && (sizeof(E) == 2 ? DIM <= 128 : DIM <= 64)) return 3;
  return 2;
}

template <typename E, int DIM, int P, int KC, int NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(sent_waves_per_eu<E, DIM, P, KC>())))
void sg_sent_kernel(SentArgs t) {
  // Multiply-adds fused in this kernel (the file builds with contraction off,
  // which the plain form's bit-exact order needs): the batched forms are
  // checked against the FP64 evaluation within a tolerance, and fusing
  // shortens the recursion's per-context chain (the repeated-vertex
  // correction is 8 fused steps instead of 16).
#pragma clang fp contract(fast)
  typedef Chunk<E> C;
  typedef typename C::v_t V;
  typedef SentLds<E, DIM, P, KC> Lds;
  constexpr int LD = Lds::LD, GXL = Lds::GXL, GPL = Lds::GPL;
  constexpr int NT = DIM / 16;   // dim tiles
  constexpr int NPT = P / 16;    // pool tiles
  constexpr int YCH = KC * DIM / 4 / 64;  // float4 chunks per lane for KC output rows
  static_assert(KC % 4 == 0 && KC <= 16 && YCH >= 1, "shape");
  __shared__ Lds lds_all[NW];
  __shared__ float s_expt[kExpTable + 1];
  // the wavefront index as a scalar: the per-wavefront LDS base in scalar
  // registers (as a lane value it and the addresses formed from it take
  // vector registers, and several of them spilled)
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int l16 = lane & 15, g4 = lane >> 4;
  const bool krow = l16 < KC;      // lane supplies an output row (m index) of an MFMA
  const bool kchk = 4 * g4 < KC;   // lane supplies output columns (k index) of an MFMA
  const V vzero = C::from(0.f, 0.f, 0.f, 0.f);
  Lds& L = lds_all[wv];
  for (int i = threadIdx.x; i <= kExpTable; i += blockDim.x) s_expt[i] = t.expt[i];
  for (int i = lane; i < KC * LD; i += 64) L.yb[i] = (E)0;
  __syncthreads();
  const int W = t.window, K = t.kd;
  const float alpha = t.alpha_dev ? *t.alpha_dev : t.alpha;
  const uint64_t pos_base = t.sen_off[t.s_begin];
  uint64_t wave_id = (uint64_t)blockIdx.x * NW + wv;
  uint64_t nwaves = (uint64_t)gridDim.x * NW;
  if (t.deterministic) {
    nwaves = 1;
    if (wave_id != 0) wave_id = t.s_end;  // one wavefront, sentences in order
  }
  // Sentences: the first 15/16 of each wavefront's share by stride, the rest
  // one at a time from the launch's counter.  Sentences differ in cost, and
  // with the stride alone a launch lasted as long as its slowest wavefront:
  // 1.4 ms (3.6%) past the mean finish on the 2.2M-vertex corpus.  The
  // 32-row launch after the 16-row one takes the few sentences of more than
  // 16 distinct vertices: by stride only (its counter would be taken once per
  // skipped sentence, and contended atomics cost more than they balance).
  const uint64_t n_sen = t.s_end - t.s_begin;
  uint64_t n_static = n_sen;
  if (!t.deterministic && (t.all || P == 16)) {
    const uint64_t per = n_sen / nwaves;
    n_static = (per - per / 16) * nwaves;
  }
  uint64_t k_next = wave_id;
  while (true) {
    uint64_t s;
    if (k_next < n_static) {
      s = t.s_begin + k_next;
      k_next += nwaves;
    } else {
      if (n_static == n_sen) break;
      uint32_t j = 0;
      if (lane == 0) j = atomicAdd(t.wq, 1u);
      j = __builtin_amdgcn_readfirstlane(j);
      if (n_static + j >= n_sen) break;
      s = t.s_begin + n_static + j;
    }
    // sentence-wide values are wave-uniform: kept in scalar registers
    const int uc = __builtin_amdgcn_readfirstlane(t.ucount[s - t.s_begin]);
    if (!t.all && (P == 16) != (uc <= 16)) continue;
    const uint64_t sb = uniform_u64(t.sen_off[s]);
    const int Ls = __builtin_amdgcn_readfirstlane((int)(t.sen_off[s + 1] - sb));
    if (Ls < 2) continue;
    const int32_t* tok = t.sen + sb;
    const uint8_t* shr = t.shrink + (sb - pos_base);
    const int32_t* ngs = t.negs + (sb - pos_base) * (uint64_t)K;
    int q0 = 0;
    while (q0 < Ls) {
      // ---- pool of the segment [q0, q1): the vertices of positions
      // lo .. q1-1+W; lane u < np holds pool vertex u
      if (t.deterministic) __threadfence();
      const int lo = q0 - W > 0 ? q0 - W : 0;
      const int tk = lo + lane < Ls ? tok[lo + lane] : -1;  // tokens lo .. lo+63
      // positions lo, lo+1, ... join the pool in order until it is full; the
      // segment is every position whose whole window has joined
      int my_id = -1, np = 0, c_last = lo - 1;
      for (int c = lo; c < Ls && c - lo < kSpan; ++c) {
        const int v = __builtin_amdgcn_readlane(tk, c - lo);
        const uint64_t m = __ballot(lane < np && my_id == v);
        int idx;
        if (m) {
          idx = __ffsll((unsigned long long)m) - 1;
        } else {
          if (np == P) break;
          idx = np;
          if (lane == np) my_id = v;
          ++np;
        }
        if (lane == 0) L.p2p[c - lo] = (uint8_t)idx;
        c_last = c;
      }
      // 2 W + 1 <= P and 2 W < kSpan: the window of q0 always joins
      const int q1 = c_last == Ls - 1 ? Ls : c_last - W + 1;
      // ---- pool rows: lane (u = p*16 + l16, g4) holds dims 16 t + 4 g4 + e
      V ox[NPT][NT];           // as loaded, rounded to E
      f32x4_t dx[NPT][NT];     // FP32 deltas
#pragma unroll
      for (int p = 0; p < NPT; ++p) {
        const int u = p * 16 + l16;
        const int row = __shfl(my_id, u, 64);
#pragma unroll
        for (int tt = 0; tt < NT; ++tt) {
          float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
          if (u < np) v = *reinterpret_cast<const float4*>(t.syn0 + (uint64_t)row * DIM + 16 * tt + 4 * g4);
          ox[p][tt] = C::from(v.x, v.y, v.z, v.w);
          dx[p][tt] = f32x4_t{0.f, 0.f, 0.f, 0.f};
        }
      }
      // ---- positions of the segment
      // the draws of each position are loaded one position ahead
      int neg_nx = lane >= 1 && lane <= K ? ngs[(uint64_t)q0 * K + lane - 1] : -1;
      int b_nx = shr[q0];
      for (int q = q0; q < q1; ++q) {
        if (t.deterministic) __threadfence();
        // output rows: lane n <= K holds row n (target, then negatives)
        const int target = __builtin_amdgcn_readlane(tk, q - lo);
        const int orow = lane == 0 ? target : neg_nx;
        const int row0 = target;
        const int b = __builtin_amdgcn_readfirstlane(b_nx);
        if (q + 1 < q1) {
          // the row of draws as a wave-uniform address each position: a
          // per-lane pointer stepped across positions would be held (and
          // spilled) for the whole loop
          typedef const int32_t __attribute__((address_space(1))) * gptr_t;
          const gptr_t nq = reinterpret_cast<gptr_t>(
              uniform_u64(reinterpret_cast<uint64_t>(ngs + (uint64_t)(q + 1) * K)));
          neg_nx = lane >= 1 && lane <= K ? nq[lane - 1] : -1;
          b_nx = shr[q + 1];
        }
        // the output rows' loads go out first (rows past K stay zero); the
        // pool rows as they stand and Gx = X X^T, which need only the pool,
        // are formed while they are in flight
        float4 yv[YCH];
#pragma unroll
        for (int r = 0; r < YCH; ++r) {
          const int ch = r * 64 + lane;
          const int n = ch / (DIM / 4), c4 = ch % (DIM / 4);
          const int row = __shfl(orow, n, 64);
          yv[r] = make_float4(0.f, 0.f, 0.f, 0.f);
          if (n <= K) yv[r] = *reinterpret_cast<const float4*>(t.syn1 + (uint64_t)row * DIM + 4 * c4);
        }
#pragma unroll
        for (int p = 0; p < NPT; ++p)
#pragma unroll
          for (int tt = 0; tt < NT; ++tt) {
            float o[4];
            if constexpr (sizeof(E) == 4) {
              o[0] = ox[p][tt].x; o[1] = ox[p][tt].y; o[2] = ox[p][tt].z; o[3] = ox[p][tt].w;
            } else {
              typename Chunk<uint16_t>::U u;
              u.v = ox[p][tt];
              o[0] = bf2f(u.h[0]); o[1] = bf2f(u.h[1]); o[2] = bf2f(u.h[2]); o[3] = bf2f(u.h[3]);
            }
            C::put4(L.xc + (p * 16 + l16) * LD + 16 * tt + 4 * g4, o[0] + dx[p][tt][0],
                    o[1] + dx[p][tt][1], o[2] + dx[p][tt][2], o[3] + dx[p][tt][3]);
          }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
#pragma unroll
        for (int p = 0; p < NPT; ++p)
#pragma unroll
          for (int p2 = 0; p2 < NPT; ++p2) {
            f32x4_t gx = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int tt = 0; tt < NT; ++tt)
              gx = C::mma(C::load4(L.xc + (p * 16 + l16) * LD + 16 * tt + 4 * g4),
                          C::load4(L.xc + (p2 * 16 + l16) * LD + 16 * tt + 4 * g4), gx);
#pragma unroll
            for (int e = 0; e < 4; ++e) L.Gx[(p * 16 + 4 * g4 + e) * GXL + p2 * 16 + l16] = gx[e];
          }
#pragma unroll
        for (int r = 0; r < YCH; ++r) {
          const int ch = r * 64 + lane;
          const int n = ch / (DIM / 4), c4 = ch % (DIM / 4);
          if (n <= K) C::put4(L.yb + n * LD + 4 * c4, yv[r].x, yv[r].y, yv[r].z, yv[r].w);
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        // S^T = Y X^T, Gy = Y Y^T (chunk operands from LDS)
        {
          f32x4_t gy = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int tt = 0; tt < NT; ++tt) {
            const V yq = krow ? C::load4(L.yb + l16 * LD + 16 * tt + 4 * g4) : vzero;
            gy = C::mma(yq, yq, gy);
          }
          if (kchk && krow) {
#pragma unroll
            for (int e = 0; e < 4; ++e) L.Gy[(4 * g4 + e) * (KC + 1) + l16] = gy[e];
          }
        }
#pragma unroll
        for (int p = 0; p < NPT; ++p) {
          f32x4_t st = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int tt = 0; tt < NT; ++tt)
            st = C::mma(krow ? C::load4(L.yb + l16 * LD + 16 * tt + 4 * g4) : vzero,
                        C::load4(L.xc + (p * 16 + l16) * LD + 16 * tt + 4 * g4), st);
          // st[e] = S^T[n = 4 g4 + e][u = p*16 + l16]
          if (kchk) {
#pragma unroll
            for (int e = 0; e < 4; ++e) L.S[(p * 16 + l16) * (KC + 1) + 4 * g4 + e] = st[e];
          }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        // ---- the recursion over the contexts in order.  Lane l works on
        // output column n = l % KC and pool vertices v = l / KC + G j (G =
        // 64 / KC groups, J = P / G vertices per lane): corr[j] and gp[j] are
        // the running corrections and g sums of (n, v).  Per context u the
        // column's corr[n][u] comes from its owner lane (n, u % G) by one
        // shuffle, gp[k][u] of every column k by readlanes of that group.
        // the recursion is the position's serial chain: while it runs, this
        // wavefront's instructions issue ahead of the other wavefronts'
        // staging, products and atomics (1% per round on the 2.2M-vertex corpus)
        __builtin_amdgcn_s_setprio(1);
        int slot = 0;
        {
          constexpr int G = 64 / KC;
          constexpr int J = P / G;
          const int n = lane % KC, grp = lane / KC;
          const int orow_n = __shfl(orow, n, 64);
          const bool live = n <= K && (n == 0 || orow_n != row0);
          const float lab = n == 0 ? 1.0f : 0.0f;
          const float wt = n == 0 ? alpha : alpha * t.wneg;
          float gyr[KC];
#pragma unroll
          for (int k = 0; k < KC; ++k) gyr[k] = L.Gy[k * (KC + 1) + n];
          float corr[J], gp[J], mrow[KC];
#pragma unroll
          for (int j = 0; j < J; ++j) {
            corr[j] = 0.0f;
            gp[j] = 0.0f;
          }
#pragma unroll
          for (int k = 0; k < KC; ++k) mrow[k] = 0.0f;
          // the position's contexts in order: lane i holds the pool index of
          // context i (the window minus the position itself)
          const int w = W - b;
          const int c_lo = q - w > 0 ? q - w : 0;
          const int c_hi = q + w < Ls - 1 ? q + w : Ls - 1;
          const int nctx = c_hi - c_lo;  // c_hi - c_lo + 1 positions, q excluded
          int uctx = 0;
          {
            const int c = c_lo + lane + (c_lo + lane >= q ? 1 : 0);
            if (lane < nctx) uctx = L.p2p[c - lo];
          }
          // first context's operands
          int u = nctx > 0 ? __builtin_amdgcn_readlane(uctx, 0) : 0;
          float s_nu = L.S[u * (KC + 1) + n];
          float gxu[J];
#pragma unroll
          for (int j = 0; j < J; ++j) gxu[j] = L.Gx[u * GXL + grp + G * j];  // Gx symmetric
          // pool vertices already seen among this position's contexts: only
          // their g sums can be non-zero (a first occurrence adds nothing)
          uint64_t seen = 0;
          for (int i = 0; i < nctx; ++i) {
            // next context's operands, in flight during this one
            const int un = __builtin_amdgcn_readlane(uctx, i + 1 < nctx ? i + 1 : i);
            const float s_next = L.S[un * (KC + 1) + n];
            float gx_next[J];
#pragma unroll
            for (int j = 0; j < J; ++j) gx_next[j] = L.Gx[un * GXL + grp + G * j];
            const int ou = u % G, ju = u / G;
            float c_own = corr[0], g_own = gp[0];
#pragma unroll
            for (int j = 1; j < J; ++j) {
              c_own = j == ju ? corr[j] : c_own;
              g_own = j == ju ? gp[j] : g_own;
            }
            const float cu = __shfl(c_own, n + KC * ou, 64);
            float sv = s_nu + cu;
            const bool rep = (seen >> u) & 1ull;
            seen |= 1ull << u;
            float gpk[KC];  // g of the earlier occurrences of this vertex, per column
            if (rep) {
#pragma unroll
              for (int k = 0; k < KC; ++k) {
                // columns past K (and skipped ones) never receive a g: their
                // sums read 0, so every column is read without a branch
                gpk[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(g_own), k + KC * ou));
                sv += gpk[k] * gyr[k];
              }
            }
            const float gl = (lab - sigmoid_of(sv, s_expt, t.exact)) * wt;
            const float g = live ? gl : 0.0f;
            if (rep) {
#pragma unroll
              for (int k = 0; k < KC; ++k) mrow[k] += g * gpk[k];
            }
            // every lane of a column holds the same g, and u is uniform: the
            // stores need no lane mask
            L.Gs[slot * KC + n] = g;
            L.su[slot] = u;
            ++slot;
#pragma unroll
            for (int j = 0; j < J; ++j) {
              corr[j] += g * gxu[j];
              gp[j] += (j == ju && grp == ou) ? g : 0.0f;
            }
            u = un;
            s_nu = s_next;
#pragma unroll
            for (int j = 0; j < J; ++j) gxu[j] = gx_next[j];
          }
#pragma unroll
          for (int j = 0; j < J; ++j) L.Gp[(grp + G * j) * GPL + n] = gp[j];
          if (lane < KC) {
#pragma unroll
            for (int k = 0; k < KC; ++k) L.Mm[n * GPL + k] = mrow[k];
          }
        }
        __builtin_amdgcn_s_setprio(0);
        for (int i = lane; i < P * GXL; i += 64) L.H[i] = 0.0f;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        // H[v_i][v_j] += g_i . g_j for every pair of context slots j < i, the
        // pairs spread over the lanes: pair p = i (i - 1) / 2 + j
        {
          const int npairs = slot * (slot - 1) / 2;
          for (int p0 = 0; p0 < npairs; p0 += 64) {
            const int pr = p0 + lane;
            if (pr < npairs) {
              int i = (int)((1.0f + __builtin_sqrtf(1.0f + 8.0f * (float)pr)) * 0.5f);
              if (i * (i - 1) / 2 > pr) --i;
              if ((i + 1) * i / 2 <= pr) ++i;
              const int j = pr - i * (i - 1) / 2;
              float dot = 0.0f;
#pragma unroll
              for (int k = 0; k < KC; ++k) dot += L.Gs[i * KC + k] * L.Gs[j * KC + k];
              atomicAdd(&L.H[L.su[i] * GXL + L.su[j]], dot);
            }
          }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        // ---- dX = Gp Y + H X into the deltas (dX^T[d][u] layout)
#pragma unroll
        for (int p = 0; p < NPT; ++p) {
          V bgp = vzero;
          if (kchk) {
            const float4 gq = *reinterpret_cast<const float4*>(L.Gp + (p * 16 + l16) * GPL + 4 * g4);
            bgp = C::from(gq.x, gq.y, gq.z, gq.w);
          }
#pragma unroll
          for (int tt = 0; tt < NT; ++tt) {
            // groups past the KC output rows read rows 0..3 and take zeros
            const V yk = C::template col4<LD>(L.yb + (kchk ? 4 * g4 : 0) * LD + 16 * tt, l16);
            dx[p][tt] = C::mma(kchk ? yk : vzero, bgp, dx[p][tt]);
          }
#pragma unroll
          for (int p2 = 0; p2 < NPT; ++p2) {
            const float4 hq = *reinterpret_cast<const float4*>(L.H + (p * 16 + l16) * GXL + p2 * 16 + 4 * g4);
            const V bh = C::from(hq.x, hq.y, hq.z, hq.w);
#pragma unroll
            for (int tt = 0; tt < NT; ++tt) {
              const V ax = C::template col4<LD>(L.xc + (p2 * 16 + 4 * g4) * LD + 16 * tt, l16);
              dx[p][tt] = C::mma(ax, bh, dx[p][tt]);
            }
          }
        }
        // ---- dY = Gp^T X + M Y -> output rows
        int drow[4];
        bool dok[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int n = 4 * g4 + e;
          drow[e] = __shfl(orow, n, 64);
          dok[e] = n <= K && (n == 0 || drow[e] != row0);
        }
        V am = vzero;
        if (krow && kchk) {
          const float4 mq = *reinterpret_cast<const float4*>(L.Mm + l16 * GPL + 4 * g4);
          am = C::from(mq.x, mq.y, mq.z, mq.w);
        }
        // Gp^T operands, the same for every dim tile
        V agp[NPT];
#pragma unroll
        for (int p = 0; p < NPT; ++p) {
          agp[p] = vzero;
          if (krow) {
            const float* gc = L.Gp + (p * 16 + 4 * g4) * GPL + l16;
            agp[p] = C::from(gc[0], gc[GPL], gc[2 * GPL], gc[3 * GPL]);
          }
        }
#pragma unroll
        for (int tt = 0; tt < NT; ++tt) {
          f32x4_t dy = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int p = 0; p < NPT; ++p) {
            const V bx = C::template col4<LD>(L.xc + (p * 16 + 4 * g4) * LD + 16 * tt, l16);
            dy = C::mma(agp[p], bx, dy);
          }
          {
            const V yk = C::template col4<LD>(L.yb + (kchk ? 4 * g4 : 0) * LD + 16 * tt, l16);
            dy = C::mma(am, kchk ? yk : vzero, dy);
          }
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (dok[e]) atomicAdd(&t.syn1[(uint64_t)drow[e] * DIM + 16 * tt + l16], dy[e]);
        }
        __builtin_amdgcn_wave_barrier();
      }
      // ---- the segment's deltas into the input rows
#pragma unroll
      for (int p = 0; p < NPT; ++p) {
        const int u = p * 16 + l16;
        const int row = __shfl(my_id, u, 64);
        if (u < np) {
#pragma unroll
          for (int tt = 0; tt < NT; ++tt)
#pragma unroll
            for (int e = 0; e < 4; ++e)
              atomicAdd(&t.syn0[(uint64_t)row * DIM + 16 * tt + 4 * g4 + e], dx[p][tt][e]);
        }
      }
      q0 = q1;
    }
  }
}

__global__ void pair_score_kernel(const uint32_t* rank, const float* syn0, uint32_t dim,
                                  const uint32_t* src, const uint32_t* dst, uint64_t n, float* out) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t i = wave; i < n; i += nwaves) {
    const uint32_t ru = rank[src[i]], rv = rank[dst[i]];
    float s = 0.0f;
    if (ru != kNoRank && rv != kNoRank)
      for (uint32_t d = lane; d < dim; d += 64) s += syn0[(uint64_t)ru * dim + d] * syn0[(uint64_t)rv * dim + d];
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) out[i] = s;
  }
}

int grid_for(dg_ctx* ctx, uint64_t n, int block) {
  uint64_t g = (n + block - 1) / block;
  uint64_t cap = (uint64_t)ctx->num_cus * 32;
  if (g > cap) g = cap;
  return g == 0 ? 1 : (int)g;
}

}  // namespace

// ---------------------------------------------------------------------------
// host state
// ---------------------------------------------------------------------------
struct dg_sgns_state {
  dg_sgns_params p{};
  uint32_t vocab = 0;
  uint32_t dim = 0;
  uint64_t train_words = 0;
  std::vector<uint32_t> vertex_of_rank;
  std::vector<uint64_t> cn_of_rank;
  uint32_t* d_rank = nullptr;
  float* d_syn0 = nullptr;
  float* d_syn1 = nullptr;
  int32_t* d_table = nullptr;
  uint32_t table_size = 0;
  float* d_expt = nullptr;
  uint32_t* d_keep = nullptr;
  // per-call buffers
  int32_t* d_sen = nullptr;
  size_t sen_cap = 0;
  uint64_t* d_sen_off = nullptr;
  size_t sen_off_cap = 0;
  uint32_t* d_cnt = nullptr;
  size_t cnt_cap = 0;
  uint8_t* d_shrink = nullptr;  // batched: window shrink per kept position of a launch
  size_t shrink_cap = 0;
  int32_t* d_negs = nullptr;    // batched: negatives per kept position of a launch
  size_t negs_cap = 0;
  int32_t* d_reuse = nullptr;   // negative-reuse form: negatives per sentence of an epoch
  size_t reuse_cap = 0;
  uint8_t* d_ucount = nullptr;  // batched: distinct vertices per sentence of a launch
  uint32_t* d_wq = nullptr;     // batched: the sentence counters of a launch's two kernels
  size_t ucount_cap = 0;
  unsigned long long* d_pool_rows = nullptr;  // batched: distinct vertices per sentence, summed
  // captured pipeline step: learning rate per launch, kept tokens, and the plan
  float* d_alpha = nullptr;
  size_t alpha_cap = 0;
  unsigned long long* d_kept = nullptr;
  struct {
    dg_sgns_params p;
    float alpha0;
    uint32_t kd, epoch0, n_mb;
    uint64_t n_walks, mb, pos_bound;
  } g{};
  uint32_t epoch = 0;       // epochs trained so far: part of every draw's Philox key
  std::vector<float> last_alphas;  // learning rate of every launch of the last training call
  uint64_t version = 0;     // changes on every write of syn0 / syn1
  bool has_vocab = true;    // false for dg_sgns_tables_create tables
  uint32_t rank_size = 0;   // entries of d_rank (vertex ids accepted by score_pairs)
};

void dg_sgns_free(dg_ctx* ctx) {
  dg_tuples_invalidate(ctx);
  dg_sgns_state* s = ctx->sgns;
  if (!s) return;
  void* bufs[] = {s->d_rank,  s->d_syn0,    s->d_syn1, s->d_table,   s->d_expt,
                  s->d_keep,  s->d_sen,     s->d_sen_off, s->d_cnt, s->d_shrink,
                  s->d_negs,  s->d_ucount,  s->d_pool_rows, s->d_alpha, s->d_kept,
                  s->d_reuse, s->d_wq};
  for (void* b : bufs)
    if (b) hipFree(b);
  delete s;
  ctx->sgns = nullptr;
}

static uint64_t next_table_version() {
  static std::atomic<uint64_t> counter{0};
  return ++counter;
}

int dg_sgns_view(dg_ctx* ctx, dg_table_view* v) {
  dg_sgns_state* s = ctx->sgns;
  if (!s || !s->d_syn0) return dg_fail(ctx, DG_ERR_STATE, "no embedding tables (dg_sgns_init or dg_sgns_tables_create)");
  v->syn0 = s->d_syn0;
  v->syn1 = s->d_syn1;
  v->rows = s->vocab;
  v->dim = s->dim;
  v->version = s->version;
  return DG_OK;
}

uint64_t dg_sgns_written(dg_ctx* ctx) {
  return ctx->sgns ? (ctx->sgns->version = next_table_version()) : 0;
}

extern "C" void dg_sgns_params_default(dg_sgns_params* p) {
  if (!p) return;
  memset(p, 0, sizeof(*p));
  p->dim = 128;
  p->window = 10;
  p->negative = 5;
  p->min_count = 0;
  p->iter = 1;
  p->batch_size = 21;
  p->sample = 1e-3f;
  p->alpha = 0.01f;
  p->form = DG_SGNS_BATCHED_FP32;
  p->vocab_source = DG_VOCAB_DEGREE;
  p->alpha_schedule = DG_ALPHA_LINEAR;
  p->deterministic = 0;
  p->exact_sigmoid = 0;
  p->minibatch = 0;
  p->table_size = 100000000u;
  p->init_float_ops = 0;
  p->negative_draws = 0;
  p->block_waves = 0;
  p->pool_split = 0;
  p->seed = 1;
}

static int check_params(dg_ctx* ctx, const dg_sgns_params* p) {
  if (p->dim == 0 || p->dim % 64 != 0 || p->dim > 512)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "dim %u: must be a multiple of 64, at most 512", p->dim);
  if (p->dim != 64 && p->dim != 128 && p->dim != 256 && p->dim != 512)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "dim %u: supported dims are 64, 128, 256, 512", p->dim);
  if (p->window == 0 || p->window > (uint32_t)kMaxWindow)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "window %u: must be 1..%d", p->window, kMaxWindow);
  if (p->negative > (uint32_t)kMaxNeg)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "negative %u: at most %d", p->negative, kMaxNeg);
  if (p->table_size == 0) return dg_fail(ctx, DG_ERR_ARG, "table_size must be positive");
  if (p->form > DG_SGNS_NEGREUSE_FP32) return dg_fail(ctx, DG_ERR_ARG, "unknown form %u", p->form);
  if (p->negative_draws != 0 && p->negative_draws < p->negative)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "negative_draws %u: at least negative (%u)",
                   p->negative_draws, p->negative);
  if (p->form == DG_SGNS_DSGL_FP32) {
    if (p->negative > 14)
      return dg_fail(ctx, DG_ERR_UNSUPPORTED, "DSGL form: at most 14 negatives (16 output rows per step)");
    if (p->window > (uint32_t)kMaxSentWindow)
      return dg_fail(ctx, DG_ERR_UNSUPPORTED, "DSGL form: window at most %d", kMaxSentWindow);
    if (p->dim > 256) return dg_fail(ctx, DG_ERR_UNSUPPORTED, "DSGL form supports dim <= 256");
  }
  if (p->form == DG_SGNS_BATCHED_FP32 || p->form == DG_SGNS_BATCHED_BF16) {
    const uint32_t kd = p->negative_draws ? p->negative_draws : p->negative;
    if (kd + 1 > (uint32_t)kOutCols)
      return dg_fail(ctx, DG_ERR_UNSUPPORTED, "batched forms: at most %d negatives per position",
                     kOutCols - 1);
    if (p->window > (uint32_t)kMaxSentWindow)
      return dg_fail(ctx, DG_ERR_UNSUPPORTED, "batched forms: window at most %d", kMaxSentWindow);
  }
  if (p->form == DG_SGNS_BATCHED_FP32 && p->dim > 128)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "batched FP32 form supports dim <= 128");
  if (p->form == DG_SGNS_BATCHED_BF16 && p->dim > 256)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "batched BF16 form supports dim <= 256");
  if (p->block_waves != 0 && p->block_waves != 1 && p->block_waves != 2 && p->block_waves != 4)
    return dg_fail(ctx, DG_ERR_ARG, "block_waves %u: 0 (default), 1, 2 or 4", p->block_waves);
  if (p->pool_split != 0 && p->pool_split != 16 && p->pool_split != 32)
    return dg_fail(ctx, DG_ERR_ARG, "pool_split %u: 0 (default), 16 or 32", p->pool_split);
  return DG_OK;
}

// Reference vocabulary, initial rows and tables (word2vec.cu:537-556,
// :632-659, :791-809, :844-878, :1569-1577; dsgl.hpp:314-382, :433-457).
extern "C" int dg_sgns_init(dg_ctx* ctx, const dg_sgns_params* p) {
  if (!ctx || !p) return DG_ERR_ARG;
  if (!ctx->d_row_ptr) return dg_fail(ctx, DG_ERR_STATE, "no graph loaded");
  DG_CHECK(check_params(ctx, p));
  DG_HIP(ctx, hipSetDevice(ctx->device));
  dg_sgns_free(ctx);
  dg_sgns_state* s = new dg_sgns_state();
  ctx->sgns = s;
  s->p = *p;
  s->dim = p->dim;
  const uint32_t V = ctx->v_num;
  // word counts
  std::vector<uint64_t> cn(V);
  if (p->vocab_source == DG_VOCAB_VISITS) {
    if (!ctx->d_freq) return dg_fail(ctx, DG_ERR_STATE, "visit counts requested but no walk ran");
    std::vector<uint32_t> f(V);
    DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    DG_HIP(ctx, hipMemcpy(f.data(), ctx->d_freq, sizeof(uint32_t) * V, hipMemcpyDeviceToHost));
    for (uint32_t v = 0; v < V; ++v) cn[v] = f[v];
  } else {
    // huge_walk.cpp:43-46: in + out degree (= 2 x degree on the doubled graph)
    std::vector<uint64_t> rp(V + 1);
    DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    DG_HIP(ctx, hipMemcpy(rp.data(), ctx->d_row_ptr, sizeof(uint64_t) * (V + 1), hipMemcpyDeviceToHost));
    for (uint32_t v = 0; v < V; ++v) cn[v] = 2 * (rp[v + 1] - rp[v]);
  }
  // ranks: descending count, ties by vertex id; drop count < min_count except rank 0
  std::vector<uint32_t> order(V);
  for (uint32_t v = 0; v < V; ++v) order[v] = v;
  std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return cn[a] != cn[b] ? cn[a] > cn[b] : a < b;
  });
  std::vector<uint32_t> rank(V, kNoRank);
  s->vertex_of_rank.clear();
  s->cn_of_rank.clear();
  s->train_words = 0;
  // the DSGL form keeps every vertex (my_LearnVocabFromTrainFile, dsgl.hpp:314-350)
  const uint64_t min_count = p->form == DG_SGNS_DSGL_FP32 ? 0 : p->min_count;
  for (uint32_t a = 0; a < V; ++a) {
    const uint32_t v = order[a];
    if (cn[v] < min_count && a != 0) continue;
    rank[v] = (uint32_t)s->vertex_of_rank.size();
    s->vertex_of_rank.push_back(v);
    s->cn_of_rank.push_back(cn[v]);
    s->train_words += cn[v];
  }
  s->vocab = (uint32_t)s->vertex_of_rank.size();
  if (s->vocab < 2) return dg_fail(ctx, DG_ERR_ARG, "vocabulary needs at least 2 words");
  const uint32_t vocab = s->vocab, dim = s->dim;
  // syn0 from the reference LCG (word2vec.cu:844-878), syn1 = 0.  Element i
  // takes the LCG state after i + 1 steps from 1; the threads start from
  // states reached by jump-ahead (k steps of x -> A x + C compose to
  // x -> A^k x + C (A^(k-1) + ... + 1) mod 2^64), so the values do not
  // depend on the thread count.
  std::vector<float> syn0((size_t)vocab * dim);
  {
    const unsigned long long A = 25214903917ULL, Cc = 11ULL;
    auto jump = [&](unsigned long long k) {  // state after k steps from 1
      unsigned long long mul = 1, add = 0;     // accumulated map x -> mul x + add
      unsigned long long pm = A, pa = Cc;      // map of 2^b steps
      while (k) {
        if (k & 1) {
          mul = mul * pm;
          add = add * pm + pa;
        }
        pa = pa * pm + pa;
        pm = pm * pm;
        k >>= 1;
      }
      return mul * 1ULL + add;
    };
    const size_t n = syn0.size();
    unsigned nt = std::thread::hardware_concurrency();
    nt = nt == 0 ? 1 : (nt > 16 ? 16 : nt);
    if (n < ((size_t)1 << 22)) nt = 1;
    std::vector<std::thread> th;
    for (unsigned k = 0; k < nt; ++k) {
      const size_t lo = n * k / nt, hi = n * (k + 1) / nt;
      th.emplace_back([&, lo, hi] {
        unsigned long long nr = jump(lo);
        for (size_t i = lo; i < hi; ++i) {
          nr = nr * A + Cc;
          if (p->init_float_ops)
            syn0[i] = (((nr & 0xFFFF) / 65536.f) - 0.5f) / (float)dim;
          else
            syn0[i] = (float)(((double)((float)(nr & 0xFFFF) / (float)65536) - 0.5) / (double)(long long)dim);
        }
      });
    }
    for (auto& x : th) x.join();
  }
  // unigram table: word2vec.cu:537-552 (double cursor), or for the DSGL form
  // dsgl.hpp:121-147 (float cursor, the index clamped before its term is added)
  s->table_size = p->table_size;
  std::vector<int32_t> table(p->table_size);
  if (p->form == DG_SGNS_DSGL_FP32) {
    double pw = 0.;
    for (uint32_t a = 0; a < vocab; ++a) pw += std::pow((double)s->cn_of_rank[a], (double)0.75f);
    int i = 0;
    float d1 = (float)(std::pow((double)s->cn_of_rank[0], (double)0.75f) / pw);
    for (uint32_t a = 0; a < p->table_size; ++a) {
      table[a] = i;
      if (a / (float)p->table_size > d1) {
        i++;
        if (i >= (int)vocab) i = (int)vocab - 1;
        d1 += (float)(std::pow((double)s->cn_of_rank[i], (double)0.75f) / pw);
      }
    }
  } else {
    double pw = 0;
    for (uint32_t a = 0; a < vocab; ++a) pw += std::pow((double)s->cn_of_rank[a], 0.75);
    uint32_t i = 0;
    double d1 = std::pow((double)s->cn_of_rank[0], 0.75) / pw;
    for (uint32_t a = 0; a < p->table_size; ++a) {
      table[a] = (int32_t)i;
      if (a / (double)p->table_size > d1) {
        i++;
        if (i < vocab) d1 += std::pow((double)s->cn_of_rank[i], 0.75) / pw;
      }
      if (i >= vocab) i = vocab - 1;
    }
  }
  // sigmoid table: exp in double of a float argument, word2vec.cu:1569-1577
  std::vector<float> expt(kExpTable + 1);
  for (int i = 0; i < kExpTable; ++i) {
    expt[i] = (float)std::exp((double)((i / (float)kExpTable * 2 - 1) * kMaxExp));
    expt[i] = expt[i] / (expt[i] + 1);
  }
  expt[kExpTable] = expt[kExpTable - 1];
  // subsampling keep limits: float arithmetic of word2vec.cu:1015, or of
  // dsgl.hpp:922-926 (sqrt(r) + r, r = sample * train_words / cn; the word's
  // own count) for the DSGL form
  std::vector<uint32_t> keep(vocab);
  for (uint32_t w = 0; w < vocab; ++w) {
    if (!(p->sample > 0) || s->cn_of_rank[w] == 0) {
      keep[w] = 65536;
      continue;
    }
    const float st = p->sample * (float)s->train_words;
    const float c = (float)s->cn_of_rank[w];
    float ran;
    if (p->form == DG_SGNS_DSGL_FP32) {
      const float ratio = st / c;
      ran = std::sqrt(ratio) + ratio;
    } else {
      ran = (std::sqrt(c / st) + 1) * st / c;
    }
    const float lim = ran * 65536.0f;
    keep[w] = lim >= 65536.0f ? 65536u : (uint32_t)lim;
  }
  const size_t nrow = (size_t)vocab * dim;
  DG_HIP(ctx, hipMalloc((void**)&s->d_rank, sizeof(uint32_t) * V));
  DG_HIP(ctx, hipMalloc((void**)&s->d_syn0, sizeof(float) * nrow));
  DG_HIP(ctx, hipMalloc((void**)&s->d_syn1, sizeof(float) * nrow));
  DG_HIP(ctx, hipMalloc((void**)&s->d_table, sizeof(int32_t) * p->table_size));
  DG_HIP(ctx, hipMalloc((void**)&s->d_expt, sizeof(float) * (kExpTable + 1)));
  DG_HIP(ctx, hipMalloc((void**)&s->d_keep, sizeof(uint32_t) * vocab));
  DG_HIP(ctx, hipMemcpy(s->d_rank, rank.data(), sizeof(uint32_t) * V, hipMemcpyHostToDevice));
  DG_HIP(ctx, hipMemcpy(s->d_syn0, syn0.data(), sizeof(float) * nrow, hipMemcpyHostToDevice));
  DG_HIP(ctx, hipMemset(s->d_syn1, 0, sizeof(float) * nrow));
  DG_HIP(ctx, hipMemcpy(s->d_table, table.data(), sizeof(int32_t) * p->table_size, hipMemcpyHostToDevice));
  DG_HIP(ctx, hipMemcpy(s->d_expt, expt.data(), sizeof(float) * (kExpTable + 1), hipMemcpyHostToDevice));
  DG_HIP(ctx, hipMemcpy(s->d_keep, keep.data(), sizeof(uint32_t) * vocab, hipMemcpyHostToDevice));
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  s->rank_size = V;
  dg_sgns_written(ctx);
  return DG_OK;
}

extern "C" int dg_sgns_tables_create(dg_ctx* ctx, uint32_t rows, uint32_t dim, const float* syn0,
                                     const float* syn1) {
  if (!ctx || !syn0 || !syn1) return DG_ERR_ARG;
  if (rows < 1 || rows > (1u << 30)) return dg_fail(ctx, DG_ERR_ARG, "rows %u: must be in [1, 2^30]", rows);
  if (dim == 0 || dim % 64 != 0 || dim > 512)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "dim %u: must be a multiple of 64, at most 512", dim);
  DG_HIP(ctx, hipSetDevice(ctx->device));
  dg_sgns_free(ctx);
  dg_sgns_state* s = new dg_sgns_state();
  ctx->sgns = s;
  dg_sgns_params_default(&s->p);
  s->p.dim = dim;
  s->dim = dim;
  s->vocab = rows;
  s->has_vocab = false;
  s->rank_size = rows;
  s->vertex_of_rank.resize(rows);
  for (uint32_t r = 0; r < rows; ++r) s->vertex_of_rank[r] = r;
  const size_t nrow = (size_t)rows * dim;
  DG_HIP(ctx, hipMalloc((void**)&s->d_rank, sizeof(uint32_t) * rows));
  DG_HIP(ctx, hipMalloc((void**)&s->d_syn0, sizeof(float) * nrow));
  DG_HIP(ctx, hipMalloc((void**)&s->d_syn1, sizeof(float) * nrow));
  DG_HIP(ctx, hipMemcpy(s->d_rank, s->vertex_of_rank.data(), sizeof(uint32_t) * rows, hipMemcpyHostToDevice));
  DG_HIP(ctx, hipMemcpy(s->d_syn0, syn0, sizeof(float) * nrow, hipMemcpyHostToDevice));
  DG_HIP(ctx, hipMemcpy(s->d_syn1, syn1, sizeof(float) * nrow, hipMemcpyHostToDevice));
  dg_sgns_written(ctx);
  return DG_OK;
}

template <int D, bool REUSE>
static void launch_plain_form(dg_ctx* ctx, const SgnsDev& a, uint64_t sb, uint64_t se, int single,
                              int exact) {
  if (single) {
    sg_plain_kernel<D, REUSE><<<1, 64, 0, ctx->stream>>>(a, sb, se, 1, exact);
  } else {
    const unsigned blocks = (unsigned)((se - sb + 3) / 4);
    sg_plain_kernel<D, REUSE><<<blocks ? blocks : 1, 256, 0, ctx->stream>>>(a, sb, se, 0, exact);
  }
}

template <int D>
static void launch_plain(dg_ctx* ctx, const SgnsDev& a, uint64_t sb, uint64_t se, int single,
                         int exact) {
  if (a.reuse_negs)
    launch_plain_form<D, true>(ctx, a, sb, se, single, exact);
  else
    launch_plain_form<D, false>(ctx, a, sb, se, single, exact);
}

// Blocks resident per CU for a kernel, times the CUs.
template <typename Kern>
static unsigned resident_grid(dg_ctx* ctx, Kern k, int block) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, block, 0) != hipSuccess || per_cu < 1)
    per_cu = 1;
  return (unsigned)per_cu * (unsigned)ctx->num_cus;
}

template <typename E, int DIM, int P, int KC, int NW>
static void launch_sent(dg_ctx* ctx, const SentArgs& t) {
  auto k = sg_sent_kernel<E, DIM, P, KC, NW>;
  uint64_t g = t.deterministic ? 1 : resident_grid(ctx, k, 64 * NW);
  const uint64_t need = (t.s_end - t.s_begin + NW - 1) / NW;
  if (g > need) g = need;
  k<<<(unsigned)(g ? g : 1), 64 * NW, 0, ctx->stream>>>(t);
}

// Four wavefronts per workgroup where their LDS fits in one CU's 160 KB,
// else two.
template <typename E, int DIM, int P, int KC>
static void launch_sent_nw(dg_ctx* ctx, const SentArgs& t) {
  constexpr bool four = 4 * sizeof(SentLds<E, DIM, P, KC>) + sizeof(float) * (kExpTable + 1) <= 163840;
  switch (t.block_waves) {
    case 1: launch_sent<E, DIM, P, KC, 1>(ctx, t); break;
    case 4:
      if constexpr (four) {
        launch_sent<E, DIM, P, KC, 4>(ctx, t);
        break;
      }
      [[fallthrough]];
    default: launch_sent<E, DIM, P, KC, 2>(ctx, t); break;
  }
}

// Hogwild: sentences of <= 16 distinct vertices on the 16-row pool, the
// others on the 32-row pool (pool_split 16), or every sentence on the 32-row
// pool (pool_split 32).  Deterministic: one wavefront, every sentence in
// order, on the 32-row pool.
template <typename E, int DIM, int KC>
static void launch_sent_kc(dg_ctx* ctx, SentArgs t) {
  uint32_t* wq = t.wq;
  if (t.deterministic) {
    t.all = 1;
    t.block_waves = 1;
    launch_sent_nw<E, DIM, 32, KC>(ctx, t);
  } else if (t.pool_split == 32) {
    t.all = 1;
    launch_sent_nw<E, DIM, 32, KC>(ctx, t);
  } else {
    t.all = 0;
    launch_sent_nw<E, DIM, 16, KC>(ctx, t);
    t.wq = wq + 1;
    launch_sent_nw<E, DIM, 32, KC>(ctx, t);
  }
}

template <typename E, int DIM>
static void launch_sent_dim(dg_ctx* ctx, const SentArgs& t) {
  if (t.kd + 1 <= 8)
    launch_sent_kc<E, DIM, 8>(ctx, t);
  else
    launch_sent_kc<E, DIM, 16>(ctx, t);
}

template <typename E>
static void launch_sent_form(dg_ctx* ctx, const SentArgs& t, uint32_t dim) {
  switch (dim) {
    case 64: launch_sent_dim<E, 64>(ctx, t); break;
    case 128: launch_sent_dim<E, 128>(ctx, t); break;
    case 256:
      // FP32 pools of dim 256 exceed the register file; rejected by check_params
      if constexpr (sizeof(E) == 2) launch_sent_dim<E, 256>(ctx, t);
      break;
  }
}

// One launch of the batched trainer over sentences [s0, s1) of the subsampled
// epoch: the draws, the distinct-vertex counts, then the sentence kernels.
// The draw kernel loops over the launch's positions as the device offsets
// give them; pos_grid (an upper bound, 0 = none) only sizes its grid.
static int enqueue_batch(dg_ctx* ctx, dg_sgns_state* s, const dg_sgns_params& p, uint32_t kd,
                         uint32_t epoch, uint64_t s0, uint64_t s1,
                         uint64_t pos_grid, float alpha, const float* alpha_dev, bool timed) {
  DrawArgs da;
  da.sen = s->d_sen;
  da.sen_off = s->d_sen_off;
  da.s_begin = s0;
  da.s_end = s1;
  da.seed = p.seed;
  da.epoch = epoch;
  da.table = s->d_table;
  da.table_size = s->table_size;
  da.vocab = s->vocab;
  da.window = p.window;
  da.kd = (int)kd;
  da.shrink = s->d_shrink;
  da.negs = s->d_negs;
  da.ucount = s->d_ucount;
  da.pool_rows = s->d_pool_rows;
  da.wq = s->d_wq;
  SentArgs t;
  t.sen = s->d_sen;
  t.sen_off = s->d_sen_off;
  t.s_begin = s0;
  t.s_end = s1;
  t.shrink = s->d_shrink;
  t.negs = s->d_negs;
  t.ucount = s->d_ucount;
  t.expt = s->d_expt;
  t.syn0 = s->d_syn0;
  t.syn1 = s->d_syn1;
  t.alpha = alpha;
  t.alpha_dev = alpha_dev;
  t.wneg = (float)p.negative / (float)kd;
  t.window = (int)p.window;
  t.kd = (int)kd;
  t.exact = (int)p.exact_sigmoid;
  t.deterministic = (int)p.deterministic;
  t.all = 0;
  t.block_waves = p.block_waves ? (int)p.block_waves : 2;
  t.pool_split = p.pool_split ? (int)p.pool_split : 16;
  t.wq = s->d_wq;
  size_t tk = 0;
  if (timed) DG_CHECK(dg_time_begin(ctx, "sgns_prep", &tk));
  if (pos_grid) sent_draw_kernel<<<grid_for(ctx, pos_grid, 256), 256, 0, ctx->stream>>>(da);
  sent_count_kernel<<<grid_for(ctx, s1 - s0, 256), 256, 0, ctx->stream>>>(da);
  if (timed) DG_CHECK(dg_time_end(ctx, tk));
  if (timed)
    DG_CHECK(dg_time_begin(ctx, p.form == DG_SGNS_BATCHED_BF16 ? "sgns_bf16" : "sgns_fp32b", &tk));
  if (p.form == DG_SGNS_BATCHED_BF16)
    launch_sent_form<uint16_t>(ctx, t, s->dim);
  else
    launch_sent_form<float>(ctx, t, s->dim);
  DG_HIP(ctx, hipGetLastError());
  if (timed) DG_CHECK(dg_time_end(ctx, tk));
  return DG_OK;
}

static int train_impl(dg_ctx* ctx, const uint32_t* d_tok, const uint64_t* d_off, uint64_t n_walks,
                      uint64_t n_tok, const dg_sgns_params& p, float alpha_in,
                      dg_sgns_stats* out);

extern "C" int dg_sgns_train(dg_ctx* ctx, int corpus, const dg_sgns_params* pin, float alpha_in,
                             dg_sgns_stats* out) {
  if (!ctx || !pin) return DG_ERR_ARG;
  dg_sgns_state* s = ctx->sgns;
  if (!s) return dg_fail(ctx, DG_ERR_STATE, "dg_sgns_init has not run");
  if (!s->has_vocab)
    return dg_fail(ctx, DG_ERR_STATE, "tables from dg_sgns_tables_create have no vocabulary: run dg_sgns_init to train");
  DG_CHECK(check_params(ctx, pin));
  if (pin->dim != s->dim) return dg_fail(ctx, DG_ERR_ARG, "dim differs from dg_sgns_init");
  DG_HIP(ctx, hipSetDevice(ctx->device));
  const uint32_t* d_tok = corpus == DG_CORPUS_TRAIN ? ctx->d_train_tokens : ctx->d_tokens;
  const uint64_t* d_off = corpus == DG_CORPUS_TRAIN ? ctx->d_train_offsets : ctx->d_offsets;
  const uint64_t n_walks = corpus == DG_CORPUS_TRAIN ? ctx->train_walks : ctx->cur_walks;
  const uint64_t n_tok = corpus == DG_CORPUS_TRAIN ? ctx->train_tokens : ctx->cur_tokens;
  if (!d_tok || n_walks == 0) return dg_fail(ctx, DG_ERR_STATE, "empty corpus");
  return train_impl(ctx, d_tok, d_off, n_walks, n_tok, *pin, alpha_in, out);
}

// p->iter epochs over sentences [0, n_walks) of a device corpus (n_tok tokens).
static int train_impl(dg_ctx* ctx, const uint32_t* d_tok, const uint64_t* d_off, uint64_t n_walks,
                      uint64_t n_tok, const dg_sgns_params& p, float alpha_in,
                      dg_sgns_stats* out) {
  dg_sgns_state* s = ctx->sgns;
  std::vector<uint64_t> h_woff(n_walks + 1);
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  DG_HIP(ctx, hipMemcpy(h_woff.data(), d_off, sizeof(uint64_t) * (n_walks + 1), hipMemcpyDeviceToHost));
  const float alpha0 = alpha_in > 0 ? alpha_in : p.alpha;
  const int D = (int)(s->dim / 64);
  dg_sgns_stats st;
  memset(&st, 0, sizeof(st));
  double ms = 0.0;
  float alpha = alpha0;
  DG_CHECK(dg_reserve(ctx, &s->d_cnt, &s->cnt_cap, (size_t)n_walks + 1));
  DG_CHECK(dg_reserve(ctx, &s->d_sen_off, &s->sen_off_cap, (size_t)n_walks + 1));
  if (!s->d_pool_rows) DG_HIP(ctx, hipMalloc((void**)&s->d_pool_rows, sizeof(unsigned long long)));
  if (!s->d_wq) DG_HIP(ctx, hipMalloc((void**)&s->d_wq, 2 * sizeof(uint32_t)));
  zero_u64_kernel<<<1, 64, 0, ctx->stream>>>(s->d_pool_rows);
  s->last_alphas.clear();
  if (p.form == DG_SGNS_DSGL_FP32) {
    // DSGL: every epoch one persistent launch of the DSGL kernel (dsgl.hip)
    dg_dsgl_call c;
    c.tok = d_tok;
    c.off = d_off;
    c.n_walks = n_walks;
    c.n_tok = n_tok;
    c.max_walk = 0;
    for (uint64_t w = 0; w < n_walks; ++w)
      c.max_walk = std::max<uint64_t>(c.max_walk, h_woff[w + 1] - h_woff[w]);
    c.rank = s->d_rank;
    c.rank_size = s->rank_size;
    c.keep = s->d_keep;
    c.table = s->d_table;
    c.table_size = s->table_size;
    c.vocab = s->vocab;
    c.expt = s->d_expt;
    c.syn0 = s->d_syn0;
    c.syn1 = s->d_syn1;
    c.dim = s->dim;
    c.window = p.window;
    c.K = p.negative;
    c.iter = p.iter;
    c.workers = p.workers;
    c.alpha0 = alpha0;
    c.train_words = s->train_words;
    c.seed = p.seed;
    c.epoch0 = s->epoch;
    c.exact = (int)p.exact_sigmoid;
    c.deterministic = (int)p.deterministic;
    s->epoch += p.iter;
    uint64_t kept = 0;
    DG_CHECK(dg_dsgl_train(ctx, c, &kept, &ms, &alpha));
    dg_sgns_written(ctx);
    DG_CHECK(dg_time_resolve(ctx));
    st.words = (uint64_t)p.iter * n_tok;
    st.kept = kept;
    st.sentences = (uint64_t)p.iter * n_walks;
    st.launches = p.iter;
    st.ms = ms;
    st.alpha_end = alpha;
    if (out) *out = st;
    return DG_OK;
  }
  for (uint32_t ep = 0; ep < p.iter; ++ep) {
    const uint32_t epoch = s->epoch++;
    // subsample (two passes) into ranked sentences
    subsample_count_kernel<<<grid_for(ctx, n_walks, 256), 256, 0, ctx->stream>>>(
        d_tok, d_off, n_walks, s->d_rank, s->rank_size, s->d_keep, p.seed, epoch, s->d_cnt);
    DG_CHECK(dg_scan_lengths(ctx, s->d_cnt, n_walks, s->d_sen_off));
    uint64_t n_kept = 0;
    DG_HIP(ctx, hipMemcpyAsync(&n_kept, s->d_sen_off + n_walks, sizeof(uint64_t),
                               hipMemcpyDeviceToHost, ctx->stream));
    DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    DG_CHECK(dg_reserve(ctx, &s->d_sen, &s->sen_cap, (size_t)n_kept + 1));
    subsample_write_kernel<<<grid_for(ctx, n_walks, 256), 256, 0, ctx->stream>>>(
        d_tok, d_off, n_walks, s->d_rank, s->rank_size, s->d_keep, p.seed, epoch, s->d_sen_off,
        s->d_sen);
    DG_HIP(ctx, hipGetLastError());
    std::vector<uint64_t> h_soff(n_walks + 1);
    DG_HIP(ctx, hipMemcpy(h_soff.data(), s->d_sen_off, sizeof(uint64_t) * (n_walks + 1),
                          hipMemcpyDeviceToHost));
    st.kept += n_kept;
    SgnsDev a;
    a.sen = s->d_sen;
    a.sen_off = s->d_sen_off;
    a.table = s->d_table;
    a.expt = s->d_expt;
    a.syn0 = s->d_syn0;
    a.syn1 = s->d_syn1;
    a.table_size = s->table_size;
    a.vocab = s->vocab;
    a.window = p.window;
    a.negative = p.negative;
    a.reuse_negs = nullptr;
    if (p.form == DG_SGNS_NEGREUSE_FP32) {
      DG_CHECK(dg_reserve(ctx, &s->d_reuse, &s->reuse_cap, (size_t)n_walks * p.negative + 1));
      if (p.negative > 0) {
        reuse_negs_kernel<<<grid_for(ctx, n_walks * p.negative, 256), 256, 0, ctx->stream>>>(
            n_walks, p.negative, dg_epoch_key(p.seed, epoch), s->d_table, s->table_size, s->vocab,
            s->d_reuse);
        DG_HIP(ctx, hipGetLastError());
      }
      a.reuse_negs = s->d_reuse;
    }
    auto alpha_at = [&](uint64_t walk_index) -> float {
      const float al = p.alpha_schedule != DG_ALPHA_LINEAR
                           ? alpha0
                           : linear_alpha(alpha0, (uint64_t)ep * n_tok + (h_woff[walk_index] - h_woff[0]),
                                          (uint64_t)p.iter * n_tok + 1);
      s->last_alphas.push_back(al);
      return al;
    };
    if (p.form == DG_SGNS_PLAIN_FP32 || p.form == DG_SGNS_NEGREUSE_FP32) {
      // reference batches: <= 15000 sentences, closed at >= 299,980 tokens
      uint64_t sb = 0;
      while (sb < n_walks) {
        uint64_t se = sb, cnt = 0;
        const uint64_t base = h_soff[sb];
        while (se < n_walks && cnt < 15000) {
          ++cnt;
          ++se;
          if (h_soff[se] - base >= (15000 - 1) * 20) break;
        }
        a.alpha = alpha = alpha_at(sb);
        size_t tk;
        DG_CHECK(dg_time_begin(ctx, p.form == DG_SGNS_PLAIN_FP32 ? "sgns_plain" : "sgns_reuse", &tk));
        switch (D) {
          case 1: launch_plain<1>(ctx, a, sb, se, p.deterministic, p.exact_sigmoid); break;
          case 2: launch_plain<2>(ctx, a, sb, se, p.deterministic, p.exact_sigmoid); break;
          case 4: launch_plain<4>(ctx, a, sb, se, p.deterministic, p.exact_sigmoid); break;
          case 8: launch_plain<8>(ctx, a, sb, se, p.deterministic, p.exact_sigmoid); break;
        }
        DG_HIP(ctx, hipGetLastError());
        DG_CHECK(dg_time_end(ctx, tk));
        st.launches++;
        sb = se;
      }
    } else {
      const uint32_t kd = p.negative_draws ? p.negative_draws : p.negative;
      const uint64_t mb = p.minibatch ? p.minibatch : 131072;
      for (uint64_t s0 = 0; s0 < n_walks; s0 += mb) {
        const uint64_t s1 = n_walks - s0 < mb ? n_walks : s0 + mb;
        const uint64_t n_pos = h_soff[s1] - h_soff[s0];
        DG_CHECK(dg_reserve(ctx, &s->d_shrink, &s->shrink_cap, (size_t)n_pos + 1));
        DG_CHECK(dg_reserve(ctx, &s->d_negs, &s->negs_cap, (size_t)n_pos * kd + 1));
        DG_CHECK(dg_reserve(ctx, &s->d_ucount, &s->ucount_cap, (size_t)(s1 - s0) + 1));
        alpha = alpha_at(s0);
        DG_CHECK(enqueue_batch(ctx, s, p, kd, epoch, s0, s1, n_pos, alpha,
                               nullptr, true));
        st.launches++;
      }
    }
  }
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  dg_sgns_written(ctx);
  for (auto& pe : ctx->pending) {
    float m = 0.f;
    DG_HIP(ctx, hipEventElapsedTime(&m, pe.start, pe.stop));
    if (pe.name.rfind("sgns", 0) == 0) ms += m;
  }
  DG_CHECK(dg_time_resolve(ctx));
  {
    unsigned long long pr = 0;
    DG_HIP(ctx, hipMemcpy(&pr, s->d_pool_rows, sizeof(pr), hipMemcpyDeviceToHost));
    st.pool_rows = (p.form == DG_SGNS_BATCHED_FP32 || p.form == DG_SGNS_BATCHED_BF16) ? (uint64_t)pr : 0;
  }
  st.words = (uint64_t)p.iter * n_tok;
  st.sentences = (uint64_t)p.iter * n_walks;
  st.ms = ms;
  st.alpha_end = alpha;
  if (out) *out = st;
  return DG_OK;
}

// ---- the trainer epoch as the captured pipeline step runs it ---------------
// The corpus of the round being walked is not known to the host while the
// step is captured: subsampled buffers are sized by the walker's token bound,
// each launch's draw buffers by a bound on its positions (slot_tokens per
// sentence plus the walker's overflow region), and the learning rate of each
// launch comes from alpha_kernel on the device.

int dg_sgns_graph_check(dg_ctx* ctx, const dg_sgns_params* pin) {
  dg_sgns_state* s = ctx->sgns;
  if (!s) return dg_fail(ctx, DG_ERR_STATE, "dg_sgns_init has not run");
  if (!s->has_vocab)
    return dg_fail(ctx, DG_ERR_STATE, "tables from dg_sgns_tables_create have no vocabulary: run dg_sgns_init to train");
  DG_CHECK(check_params(ctx, pin));
  if (pin->dim != s->dim) return dg_fail(ctx, DG_ERR_ARG, "dim differs from dg_sgns_init");
  if (pin->form != DG_SGNS_BATCHED_FP32 && pin->form != DG_SGNS_BATCHED_BF16)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED,
                   "the captured pipeline step runs the batched forms (the plain forms batch by host offsets, "
                   "the DSGL form runs one launch per epoch)");
  return DG_OK;
}

int dg_sgns_graph_plan(dg_ctx* ctx, const dg_sgns_params* pin, float alpha_in, uint64_t n_walks,
                       uint64_t tok_bound, uint64_t slot_tokens, uint64_t overflow_tokens) {
  dg_sgns_state* s = ctx->sgns;
  DG_CHECK(dg_sgns_graph_check(ctx, pin));
  if (n_walks == 0) return dg_fail(ctx, DG_ERR_STATE, "empty corpus");
  auto& g = s->g;
  g.p = *pin;
  g.alpha0 = alpha_in > 0 ? alpha_in : pin->alpha;
  g.kd = pin->negative_draws ? pin->negative_draws : pin->negative;
  g.mb = pin->minibatch ? pin->minibatch : 131072;
  g.n_walks = n_walks;
  g.n_mb = (uint32_t)((n_walks + g.mb - 1) / g.mb);
  const uint64_t per_launch = (n_walks < g.mb ? n_walks : g.mb) * slot_tokens + overflow_tokens;
  g.pos_bound = per_launch < tok_bound ? per_launch : tok_bound;
  g.epoch0 = s->epoch;
  DG_CHECK(dg_reserve(ctx, &s->d_cnt, &s->cnt_cap, (size_t)n_walks + 1));
  DG_CHECK(dg_reserve(ctx, &s->d_sen_off, &s->sen_off_cap, (size_t)n_walks + 1));
  DG_CHECK(dg_reserve(ctx, &s->d_sen, &s->sen_cap, (size_t)tok_bound + 1));
  DG_CHECK(dg_reserve(ctx, &s->d_shrink, &s->shrink_cap, (size_t)g.pos_bound + 1));
  DG_CHECK(dg_reserve(ctx, &s->d_negs, &s->negs_cap, (size_t)g.pos_bound * g.kd + 1));
  DG_CHECK(dg_reserve(ctx, &s->d_ucount, &s->ucount_cap, (size_t)(n_walks < g.mb ? n_walks : g.mb) + 1));
  DG_CHECK(dg_reserve(ctx, &s->d_alpha, &s->alpha_cap, (size_t)g.n_mb * pin->iter));
  if (!s->d_pool_rows) DG_HIP(ctx, hipMalloc((void**)&s->d_pool_rows, sizeof(unsigned long long)));
  if (!s->d_wq) DG_HIP(ctx, hipMalloc((void**)&s->d_wq, 2 * sizeof(uint32_t)));
  if (!s->d_kept) DG_HIP(ctx, hipMalloc((void**)&s->d_kept, sizeof(unsigned long long)));
  return DG_OK;
}

int dg_sgns_graph_enqueue(dg_ctx* ctx, const uint32_t* d_tok, const uint64_t* d_off, bool timed) {
  dg_sgns_state* s = ctx->sgns;
  const auto& g = s->g;
  const dg_sgns_params& p = g.p;
  const uint64_t n_walks = g.n_walks;
  zero_u64_kernel<<<1, 64, 0, ctx->stream>>>(s->d_pool_rows);
  zero_u64_kernel<<<1, 64, 0, ctx->stream>>>(s->d_kept);
  for (uint32_t ep = 0; ep < p.iter; ++ep) {
    const uint32_t epoch = g.epoch0 + ep;
    subsample_count_kernel<<<grid_for(ctx, n_walks, 256), 256, 0, ctx->stream>>>(
        d_tok, d_off, n_walks, s->d_rank, s->rank_size, s->d_keep, p.seed, epoch, s->d_cnt);
    DG_CHECK(dg_scan_lengths(ctx, s->d_cnt, n_walks, s->d_sen_off));
    subsample_write_kernel<<<grid_for(ctx, n_walks, 256), 256, 0, ctx->stream>>>(
        d_tok, d_off, n_walks, s->d_rank, s->rank_size, s->d_keep, p.seed, epoch, s->d_sen_off,
        s->d_sen);
    kept_add_kernel<<<1, 64, 0, ctx->stream>>>(s->d_sen_off, n_walks, s->d_kept);
    float* alphas = s->d_alpha + (size_t)ep * g.n_mb;
    alpha_kernel<<<(g.n_mb + 255) / 256, 256, 0, ctx->stream>>>(
        d_off, n_walks, g.mb, g.n_mb, g.alpha0, p.alpha_schedule == DG_ALPHA_LINEAR ? 1 : 0, p.iter,
        ep, alphas);
    DG_HIP(ctx, hipGetLastError());
    uint32_t i = 0;
    for (uint64_t s0 = 0; s0 < n_walks; s0 += g.mb, ++i) {
      const uint64_t s1 = n_walks - s0 < g.mb ? n_walks : s0 + g.mb;
      DG_CHECK(enqueue_batch(ctx, s, p, g.kd, epoch, s0, s1,
                             g.pos_bound, g.alpha0, alphas + i, timed));
    }
  }
  return DG_OK;
}

int dg_sgns_graph_collect(dg_ctx* ctx, bool commit, double ms, dg_sgns_stats* out) {
  dg_sgns_state* s = ctx->sgns;
  const auto& g = s->g;
  if (!commit) return DG_OK;
  unsigned long long kept = 0, pr = 0;
  float alpha_end = 0.f;
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  DG_HIP(ctx, hipMemcpy(&kept, s->d_kept, sizeof(kept), hipMemcpyDeviceToHost));
  DG_HIP(ctx, hipMemcpy(&pr, s->d_pool_rows, sizeof(pr), hipMemcpyDeviceToHost));
  alpha_end = g.alpha0;  // no epoch (iter 0): nothing trained
  s->last_alphas.assign((size_t)g.p.iter * g.n_mb, 0.0f);
  if (g.p.iter > 0) {
    DG_HIP(ctx, hipMemcpy(s->last_alphas.data(), s->d_alpha, sizeof(float) * s->last_alphas.size(),
                          hipMemcpyDeviceToHost));
    alpha_end = s->last_alphas.back();
  }
  s->epoch = g.epoch0 + g.p.iter;
  dg_sgns_written(ctx);
  if (out) {
    memset(out, 0, sizeof(*out));
    out->words = (uint64_t)g.p.iter * ctx->cur_tokens;
    out->kept = kept;
    out->sentences = (uint64_t)g.p.iter * g.n_walks;
    out->launches = (uint64_t)g.p.iter * g.n_mb;
    out->ms = ms;
    out->alpha_end = alpha_end;
    out->pool_rows = pr;
  }
  return DG_OK;
}

// ---- launch-shape autotuner ------------------------------------------------
// Coordinate search on the caller's corpus: wavefronts per workgroup x pool
// split at the default launch size, then sentences per launch with the best
// shape.  Each candidate trains the sample twice (Hogwild) and keeps its faster
// device time; the tables and the epoch counter are copied before
// and restored after, so tuning leaves no trace in the training state.

extern "C" int dg_sgns_autotune(dg_ctx* ctx, int corpus, const dg_sgns_params* pin,
                                uint64_t sample_walks, dg_sgns_params* tuned,
                                double* words_per_s) {
  if (!ctx || !pin || !tuned) return DG_ERR_ARG;
  dg_sgns_state* s = ctx->sgns;
  if (!s) return dg_fail(ctx, DG_ERR_STATE, "dg_sgns_init has not run");
  if (!s->has_vocab)
    return dg_fail(ctx, DG_ERR_STATE, "tables from dg_sgns_tables_create have no vocabulary: run dg_sgns_init to train");
  DG_CHECK(check_params(ctx, pin));
  if (pin->dim != s->dim) return dg_fail(ctx, DG_ERR_ARG, "dim differs from dg_sgns_init");
  if (pin->form != DG_SGNS_BATCHED_FP32 && pin->form != DG_SGNS_BATCHED_BF16)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "the launch shapes tuned here are those of the batched forms");
  DG_HIP(ctx, hipSetDevice(ctx->device));
  const uint32_t* d_tok = corpus == DG_CORPUS_TRAIN ? ctx->d_train_tokens : ctx->d_tokens;
  const uint64_t* d_off = corpus == DG_CORPUS_TRAIN ? ctx->d_train_offsets : ctx->d_offsets;
  const uint64_t n_walks = corpus == DG_CORPUS_TRAIN ? ctx->train_walks : ctx->cur_walks;
  if (!d_tok || n_walks == 0) return dg_fail(ctx, DG_ERR_STATE, "empty corpus");
  uint64_t n = sample_walks ? sample_walks : 524288;
  if (n > n_walks) n = n_walks;
  uint64_t off[2] = {0, 0};
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  DG_HIP(ctx, hipMemcpy(&off[0], d_off, sizeof(uint64_t), hipMemcpyDeviceToHost));
  DG_HIP(ctx, hipMemcpy(&off[1], d_off + n, sizeof(uint64_t), hipMemcpyDeviceToHost));
  const uint64_t n_tok = off[1] - off[0];

  // the training state, copied aside
  const size_t bytes = sizeof(float) * (size_t)s->vocab * s->dim;
  float* save0 = nullptr;
  float* save1 = nullptr;
  DG_HIP(ctx, hipMalloc((void**)&save0, bytes));
  if (hipMalloc((void**)&save1, bytes) != hipSuccess) {
    hipFree(save0);
    return dg_fail(ctx, DG_ERR_NOMEM, "autotune: no room to save the tables");
  }
  const uint32_t epoch0 = s->epoch;
  int rc = DG_OK;
  auto restore = [&]() {
    hipMemcpyAsync(s->d_syn0, save0, bytes, hipMemcpyDeviceToDevice, ctx->stream);
    hipMemcpyAsync(s->d_syn1, save1, bytes, hipMemcpyDeviceToDevice, ctx->stream);
    hipStreamSynchronize(ctx->stream);
    hipFree(save0);
    hipFree(save1);
    s->epoch = epoch0;
    dg_sgns_written(ctx);
  };
  if (hipMemcpyAsync(save0, s->d_syn0, bytes, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess ||
      hipMemcpyAsync(save1, s->d_syn1, bytes, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess) {
    restore();
    return dg_fail(ctx, DG_ERR_HIP, "autotune: saving the tables failed");
  }

  dg_sgns_params base = *pin;
  base.iter = 1;
  base.deterministic = 0;
  // device time of one candidate: the faster of two runs
  auto rate = [&](const dg_sgns_params& c, double* wps) -> int {
    double best = 0.0;
    for (int rep = 0; rep < 2; ++rep) {
      dg_sgns_stats st;
      const int r = train_impl(ctx, d_tok, d_off, n, n_tok, c, 0.f, &st);
      if (r != DG_OK) return r;
      const double w = st.ms > 0 ? (double)st.words / (st.ms * 1e-3) : 0.0;
      if (w > best) best = w;
    }
    *wps = best;
    return DG_OK;
  };
  dg_sgns_params best = base;
  best.minibatch = base.minibatch ? base.minibatch : 131072;
  best.block_waves = base.block_waves ? base.block_waves : 2;
  best.pool_split = base.pool_split ? base.pool_split : 16;
  double best_rate = 0.0, default_rate = 0.0, warm = 0.0;
  rc = rate(best, &warm);  // code and caches warm
  if (rc == DG_OK) rc = rate(best, &default_rate);
  best_rate = default_rate;
  const uint32_t waves[3] = {1, 2, 4};
  const uint32_t splits[2] = {16, 32};
  for (int i = 0; i < 3 && rc == DG_OK; ++i)
    for (int j = 0; j < 2 && rc == DG_OK; ++j) {
      dg_sgns_params c = best;
      c.block_waves = waves[i];
      c.pool_split = splits[j];
      if (c.block_waves == best.block_waves && c.pool_split == best.pool_split) continue;
      double r = 0.0;
      rc = rate(c, &r);
      if (rc == DG_OK && r > best_rate) {
        best_rate = r;
        best.block_waves = c.block_waves;
        best.pool_split = c.pool_split;
      }
    }
  const uint32_t sizes[5] = {32768, 65536, 131072, 262144, 524288};
  const uint32_t mb0 = best.minibatch;
  for (int i = 0; i < 5 && rc == DG_OK; ++i) {
    if (sizes[i] == mb0 || sizes[i] > n) continue;
    dg_sgns_params c = best;
    c.minibatch = sizes[i];
    double r = 0.0;
    rc = rate(c, &r);
    if (rc == DG_OK && r > best_rate) {
      best_rate = r;
      best.minibatch = c.minibatch;
    }
  }
  restore();
  if (rc != DG_OK) return rc;
  *tuned = *pin;
  tuned->minibatch = best.minibatch;
  tuned->block_waves = best.block_waves;
  tuned->pool_split = best.pool_split;
  if (words_per_s) {
    words_per_s[0] = default_rate;
    words_per_s[1] = best_rate;
  }
  return DG_OK;
}

extern "C" int dg_sgns_last_alphas(dg_ctx* ctx, float* alphas, uint32_t cap, uint32_t* n) {
  if (!ctx) return DG_ERR_ARG;
  if (!ctx->sgns) return dg_fail(ctx, DG_ERR_STATE, "dg_sgns_init has not run");
  const std::vector<float>& v = ctx->sgns->last_alphas;
  if (n) *n = (uint32_t)v.size();
  if (alphas) memcpy(alphas, v.data(), sizeof(float) * std::min<size_t>(cap, v.size()));
  return DG_OK;
}

extern "C" int dg_sgns_vocab(dg_ctx* ctx, uint32_t* vertex_of_rank, uint32_t* vocab_size) {
  if (!ctx) return DG_ERR_ARG;
  if (!ctx->sgns) return dg_fail(ctx, DG_ERR_STATE, "dg_sgns_init has not run");
  if (vocab_size) *vocab_size = ctx->sgns->vocab;
  if (vertex_of_rank)
    memcpy(vertex_of_rank, ctx->sgns->vertex_of_rank.data(), sizeof(uint32_t) * ctx->sgns->vocab);
  return DG_OK;
}

extern "C" int dg_sgns_get_tables(dg_ctx* ctx, float* syn0, float* syn1) {
  if (!ctx) return DG_ERR_ARG;
  dg_sgns_state* s = ctx->sgns;
  if (!s) return dg_fail(ctx, DG_ERR_STATE, "dg_sgns_init has not run");
  DG_HIP(ctx, hipSetDevice(ctx->device));
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const size_t n = (size_t)s->vocab * s->dim;
  if (syn0) DG_HIP(ctx, hipMemcpy(syn0, s->d_syn0, sizeof(float) * n, hipMemcpyDeviceToHost));
  if (syn1) DG_HIP(ctx, hipMemcpy(syn1, s->d_syn1, sizeof(float) * n, hipMemcpyDeviceToHost));
  return DG_OK;
}

extern "C" int dg_sgns_set_tables(dg_ctx* ctx, const float* syn0, const float* syn1) {
  if (!ctx) return DG_ERR_ARG;
  dg_sgns_state* s = ctx->sgns;
  if (!s) return dg_fail(ctx, DG_ERR_STATE, "dg_sgns_init has not run");
  DG_HIP(ctx, hipSetDevice(ctx->device));
  const size_t n = (size_t)s->vocab * s->dim;
  if (syn0) DG_HIP(ctx, hipMemcpy(s->d_syn0, syn0, sizeof(float) * n, hipMemcpyHostToDevice));
  if (syn1) DG_HIP(ctx, hipMemcpy(s->d_syn1, syn1, sizeof(float) * n, hipMemcpyHostToDevice));
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  dg_sgns_written(ctx);
  return DG_OK;
}

extern "C" int dg_sgns_score_pairs(dg_ctx* ctx, const uint32_t* src, const uint32_t* dst,
                                   uint64_t n, float* scores) {
  if (!ctx || (n && (!src || !dst || !scores))) return DG_ERR_ARG;
  dg_sgns_state* s = ctx->sgns;
  if (!s) return dg_fail(ctx, DG_ERR_STATE, "dg_sgns_init has not run");
  if (n == 0) return DG_OK;
  for (uint64_t i = 0; i < n; ++i)
    if (src[i] >= s->rank_size || dst[i] >= s->rank_size)
      return dg_fail(ctx, DG_ERR_ARG, "pair %llu: vertex id out of range (%u vertices)",
                     (unsigned long long)i, s->rank_size);
  DG_HIP(ctx, hipSetDevice(ctx->device));
  uint32_t *ds = nullptr, *dd = nullptr;
  float* dout = nullptr;
  DG_HIP(ctx, hipMalloc((void**)&ds, sizeof(uint32_t) * n));
  DG_HIP(ctx, hipMalloc((void**)&dd, sizeof(uint32_t) * n));
  DG_HIP(ctx, hipMalloc((void**)&dout, sizeof(float) * n));
  DG_HIP(ctx, hipMemcpy(ds, src, sizeof(uint32_t) * n, hipMemcpyHostToDevice));
  DG_HIP(ctx, hipMemcpy(dd, dst, sizeof(uint32_t) * n, hipMemcpyHostToDevice));
  pair_score_kernel<<<grid_for(ctx, n * 64, 256), 256, 0, ctx->stream>>>(s->d_rank, s->d_syn0, s->dim,
                                                                          ds, dd, n, dout);
  DG_HIP(ctx, hipGetLastError());
  DG_HIP(ctx, hipMemcpy(scores, dout, sizeof(float) * n, hipMemcpyDeviceToHost));
  hipFree(ds);
  hipFree(dd);
  hipFree(dout);
  return DG_OK;
}

extern "C" int dg_emb_write(dg_ctx* ctx, const char* path, int binary) {
  if (!ctx || !path) return DG_ERR_ARG;
  dg_sgns_state* s = ctx->sgns;
  if (!s) return dg_fail(ctx, DG_ERR_STATE, "dg_sgns_init has not run");
  std::vector<float> syn0((size_t)s->vocab * s->dim);
  DG_CHECK(dg_sgns_get_tables(ctx, syn0.data(), nullptr));
  FILE* f = fopen(path, "wb");
  if (!f) return dg_fail(ctx, DG_ERR_IO, "cannot open %s", path);
  // dsgl.hpp:1443-1461
  fprintf(f, "%ld %ld\n", (long)s->vocab, (long)s->dim);
  for (uint32_t a = 0; a < s->vocab; ++a) {
    fprintf(f, "%u ", s->vertex_of_rank[a]);
    if (binary)
      fwrite(&syn0[(size_t)a * s->dim], sizeof(float), s->dim, f);
    else
      for (uint32_t b = 0; b < s->dim; ++b) fprintf(f, "%f ", syn0[(size_t)a * s->dim + b]);
    fprintf(f, "\n");
  }
  return fclose(f) == 0 ? DG_OK : dg_fail(ctx, DG_ERR_IO, "write to %s failed", path);
}
