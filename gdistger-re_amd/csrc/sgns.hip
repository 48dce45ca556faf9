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
//   * batched (DG_SGNS_BATCHED_*): a sentence is cut into tiles of T
//     consecutive positions (16, or 8 when K > 5).  Each position draws its
//     own K negatives, shared by the contexts of its window (the per-position
//     minibatch of DSGL, dsgl.hpp:1100-1159); the tile's rows are staged in
//     LDS once and the (contexts x outputs) logits and both rank-1 update
//     products run on MFMA, so a context row is read once per tile instead of
//     once per (context, target) pair.  All gradients of a launch come
//     from the tables as they were at its start ("snapshot" semantics) when
//     deterministic; they are then summed per row by a sorted-key segmented
//     scatter-add in a fixed order.  Non-deterministic mode adds them with
//     atomics as they are produced (Hogwild, like the reference kernels).
//   * BF16: rows stored in BF16 next to FP32 master rows; logits on BF16 MFMA.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "dg_internal.h"
#include "dg_philox.h"

namespace {

constexpr int kExpTable = 1000;
constexpr int kMaxExp = 6;
constexpr int kMaxNeg = 16;     // -negative
constexpr int kMaxWindow = 24;  // tile positions + 2 * window <= 64 staged context rows
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

__device__ __forceinline__ float sigmoid_of(float f, const float* expt, int exact) {
  if (exact) return (float)(1.0 / (1.0 + exp(-(double)f)));
  if (f > kMaxExp) return 1.0f;
  if (f < -kMaxExp) return 0.0f;
  return expt[(int)((f + kMaxExp) * (kExpTable / kMaxExp / 2))];
}

// ---- subsampling (word2vec.cu:1011-1020 with a Philox draw per token) ------
__global__ void subsample_count_kernel(const uint32_t* tokens, const uint64_t* offsets,
                                       uint64_t n_walks, const uint32_t* rank,
                                       const uint32_t* keep, uint64_t seed, uint32_t epoch,
                                       uint32_t* kept) {
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < n_walks;
       w += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t c = 0;
    for (uint64_t i = offsets[w]; i < offsets[w + 1]; ++i) {
      const uint32_t r = rank[tokens[i]];
      if (r == kNoRank) continue;
      const dg_u32x4 o = dg_philox4x32_10(0, (uint32_t)i, (uint32_t)(i >> 32) ^ (epoch << 24),
                                          DG_TAG_SAMPLE, (uint32_t)seed, (uint32_t)(seed >> 32));
      c += ((o.x0 & 0xFFFFu) <= keep[r]);
    }
    kept[w] = c;
  }
}

__global__ void subsample_write_kernel(const uint32_t* tokens, const uint64_t* offsets,
                                       uint64_t n_walks, const uint32_t* rank,
                                       const uint32_t* keep, uint64_t seed, uint32_t epoch,
                                       const uint64_t* sen_off, int32_t* sen) {
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < n_walks;
       w += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t o_pos = sen_off[w];
    for (uint64_t i = offsets[w]; i < offsets[w + 1]; ++i) {
      const uint32_t r = rank[tokens[i]];
      if (r == kNoRank) continue;
      const dg_u32x4 o = dg_philox4x32_10(0, (uint32_t)i, (uint32_t)(i >> 32) ^ (epoch << 24),
                                          DG_TAG_SAMPLE, (uint32_t)seed, (uint32_t)(seed >> 32));
      if ((o.x0 & 0xFFFFu) <= keep[r]) sen[o_pos++] = (int32_t)r;
    }
  }
}

// ---- plain form: skip_gram_kernel (word2vec.cu:305-399), negative sampling --
// One wavefront per sentence; lane l owns dims l + 64 j.  `single` runs the
// whole batch in one wavefront in sentence order (the deterministic mode).
template <int D>
__global__ __launch_bounds__(256) void sg_plain_kernel(SgnsDev a, uint64_t s_begin, uint64_t s_end,
                                                       int single, int exact) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = single ? 1 : ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t s = s_begin + wave; s < s_end; s += nwaves) {
    const uint64_t sb = a.sen_off[s], se = a.sen_off[s + 1];
    unsigned long long next_random = (unsigned long long)(s - s_begin);  // blockIdx.x
    for (uint64_t pos = sb; pos < se; ++pos) {
      const int word = a.sen[pos];
      next_random = next_random * 2514903917ULL + 11;  // word2vec.cu:321 (sic)
      const int b = (int)(next_random % a.window);
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
            next_random = next_random * 25214903917ULL + 11;
            target = a.table[(next_random >> 16) % a.table_size];
            if (target == 0) target = (int)(next_random % (a.vocab - 1) + 1);
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

// ---- batched form: position tiles on MFMA ----------------------------------
// A tile is T consecutive positions p0 .. p1-1 of a sentence.  Its context
// rows X (positions c_lo .. c_hi-1, c_lo = max(0, p0-W), c_hi = min(L, p1+W))
// and output rows Y (per position j, the columns j(1+K) .. j(1+K)+K hold its
// target and its K negatives) are staged in LDS, then
//   S  = X Y^T,  Gx = X X^T                                (MFMA)
//   per output column n (position p, label l, weight w), over the contexts
//   c_1 < c_2 < ... of the shrunk window of p:
//     s_k = S[c_k][n] + sum_{j<k} g_j Gx[c_k][c_j],   g_k = alpha w (l - sig(s_k))
//   dX = G Y,  dY = G^T X                                  (MFMA)
// s_k equals x_{c_k} . (y_n + sum_{j<k} g_j x_{c_j}): the output row as it
// stands after the updates of the earlier contexts of the same position,
// which is the order of skip_gram_kernel (word2vec.cu:360-396), where every
// (context, target) pair updates the target row before the next context is
// read.  The Gram matrix turns that dependency chain into a scalar
// recursion, so the row products stay on MFMA.  dX is taken with the rows of
// Y as staged (the context side of the reference update differs from this
// by terms of order alpha^2).
// Rows live in HBM as FP32 only.  The FP32 form stages them as they are and
// runs v_mfma_f32_16x16x4f32; the BF16 form rounds them to BF16 while staging
// and runs v_mfma_f32_16x16x32_bf16 (logits, Gram) and 16x16x16_bf16 (the
// update products, G rounded to BF16), both with FP32 accumulation, adding
// its FP32 results to the FP32 rows.  Every staged row is also updated by the tile that staged it, so a
// BF16 copy in HBM would need a refresh (4 B read + 2 B written per element)
// for each 2 B it saves on the gather, and would be stale in between.
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4_t;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4_t;

__device__ __forceinline__ uint16_t f2bf(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

template <typename E>
struct Mfma;

template <>
struct Mfma<float> {
  // logits / Gram products (k = dim) and update products (k = rows) both on
  // v_mfma_f32_16x16x4f32
  static constexpr int KS = 4;
  static constexpr int KU = 4;
  typedef float frag_t;
  typedef float ufrag_t;
  // M[r0 + lane%16][k0 + lane/16]
  __device__ static frag_t row_frag(const float* M, int ld, int r0, int k0, int lane) {
    return M[(r0 + (lane & 15)) * ld + k0 + (lane >> 4)];
  }
  __device__ static f32x4_t mma(frag_t a, frag_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // update products: A from the FP32 gradient matrix, B from staged rows
  __device__ static ufrag_t urow_f(const float* M, int ld, int r0, int k0, int lane) {
    return M[(r0 + (lane & 15)) * ld + k0 + (lane >> 4)];
  }
  __device__ static ufrag_t ucol_f(const float* M, int ld, int k0, int c0, int lane) {
    return M[(k0 + (lane >> 4)) * ld + c0 + (lane & 15)];
  }
  __device__ static ufrag_t ucol(const float* M, int ld, int k0, int c0, int lane) {
    return M[(k0 + (lane >> 4)) * ld + c0 + (lane & 15)];
  }
  __device__ static f32x4_t mmau(ufrag_t a, ufrag_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

template <>
struct Mfma<uint16_t> {
  // logits / Gram products on v_mfma_f32_16x16x32_bf16 (k = dim); update
  // products on v_mfma_f32_16x16x16_bf16, so row counts pad to 16 only
  static constexpr int KS = 32;
  static constexpr int KU = 16;
  typedef bf16x8_t frag_t;
  typedef bf16x4_t ufrag_t;
  union U {
    uint4 q;
    uint16_t h[8];
    bf16x8_t v;
  };
  union U4 {
    uint2 q;
    uint16_t h[4];
    bf16x4_t v;
  };
  // M[r0 + lane%16][k0 + 8 (lane/16) .. +7]
  __device__ static frag_t row_frag(const uint16_t* M, int ld, int r0, int k0, int lane) {
    U u;
    u.q = *reinterpret_cast<const uint4*>(M + (r0 + (lane & 15)) * ld + k0 + 8 * (lane >> 4));
    return u.v;
  }
  __device__ static f32x4_t mma(frag_t a, frag_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
  // FP32 M[r0 + lane%16][k0 + 4 (lane/16) .. +3] rounded to BF16 (16-byte aligned rows)
  __device__ static ufrag_t urow_f(const float* M, int ld, int r0, int k0, int lane) {
    const float4 a = *reinterpret_cast<const float4*>(M + (r0 + (lane & 15)) * ld + k0 + 4 * (lane >> 4));
    U4 u;
    u.h[0] = f2bf(a.x);
    u.h[1] = f2bf(a.y);
    u.h[2] = f2bf(a.z);
    u.h[3] = f2bf(a.w);
    return u.v;
  }
  // FP32 M[k0 + 4 (lane/16) + i][c0 + lane%16], i = 0..3, rounded to BF16
  __device__ static ufrag_t ucol_f(const float* M, int ld, int k0, int c0, int lane) {
    U4 u;
#pragma unroll
    for (int i = 0; i < 4; ++i) u.h[i] = f2bf(M[(k0 + 4 * (lane >> 4) + i) * ld + c0 + (lane & 15)]);
    return u.v;
  }
  // BF16 M[k0 + 4 (lane/16) + i][c0 + lane%16], i = 0..3
  __device__ static ufrag_t ucol(const uint16_t* M, int ld, int k0, int c0, int lane) {
    U4 u;
#pragma unroll
    for (int i = 0; i < 4; ++i) u.h[i] = M[(k0 + 4 * (lane >> 4) + i) * ld + c0 + (lane & 15)];
    return u.v;
  }
  __device__ static f32x4_t mmau(ufrag_t a, ufrag_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
  }
};

// Tile descriptor, built by tile_desc_kernel for every tile of a launch:
//   [0] p0, [1] c_lo, [2] nc (context rows), [3] np (positions),
//   [4 .. 4+T)            window shrink b of each position,
//   [4+T .. 4+T+CTX)      context rows (vocabulary ranks, -1 past nc),
//   [4+T+CTX .. +NOUT)    output rows per column (-1 past np (1+K)).
constexpr int kDescHead = 4;

struct DescArgs {
  const int32_t* sen;
  const uint64_t* sen_off;
  const int32_t* table;
  const uint2* tiles;  // {sentence index, tile index within the sentence}
  uint64_t n_tiles;
  uint64_t first_sentence;
  uint64_t seed;
  uint32_t epoch;
  uint32_t table_size;
  uint32_t vocab;
  int window;
  int kd;  // negatives drawn per position
  int T;
  int ctx;
  int nout;
  uint32_t stride;  // ints per descriptor
  int32_t* desc;
};

// word j of a Philox stream (block j/4, word j%4)
__device__ __forceinline__ uint32_t stream_word(uint64_t seed, uint32_t c1, uint32_t c2,
                                                uint32_t tag, uint32_t j) {
  const dg_u32x4 o = dg_philox4x32_10(j >> 2, c1, c2, tag, (uint32_t)seed, (uint32_t)(seed >> 32));
  const uint32_t r = j & 3u;
  return r == 0 ? o.x0 : (r == 1 ? o.x1 : (r == 2 ? o.x2 : o.x3));
}

// One wavefront per tile.  Draws: word q < np of the tile's stream gives
// b = word % window of position p0+q; word np + q K + k gives negative k of
// position q: table[word % table_size], 0 -> word % (vocab-1) + 1.
__global__ __launch_bounds__(256) void tile_desc_kernel(DescArgs a) {
  const int lane = threadIdx.x & 63;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const int K = a.kd, W = a.window;
  for (uint64_t ti = wave; ti < a.n_tiles; ti += nwaves) {
    const uint2 tl = a.tiles[ti];
    const uint64_t sb = a.sen_off[tl.x];
    const int L = (int)(a.sen_off[tl.x + 1] - sb);
    const int p0 = (int)tl.y * a.T;
    const int p1 = p0 + a.T < L ? p0 + a.T : L;
    const int np = p1 - p0;
    const int c_lo = p0 - W > 0 ? p0 - W : 0;
    const int c_hi = p1 + W < L ? p1 + W : L;
    const int nc = c_hi - c_lo;
    const uint32_t c1 = (uint32_t)(a.first_sentence + tl.x), c2 = tl.y | (a.epoch << 16);
    int32_t* d = a.desc + ti * a.stride;
    const int n_ent = kDescHead + a.T + a.ctx + a.nout;
    for (int e = lane; e < n_ent; e += 64) {
      int v;
      if (e < kDescHead) {
        v = e == 0 ? p0 : (e == 1 ? c_lo : (e == 2 ? nc : np));
      } else if (e < kDescHead + a.T) {
        const int q = e - kDescHead;
        v = q < np ? (int)(stream_word(a.seed, c1, c2, DG_TAG_SGNS, (uint32_t)q) % (uint32_t)W) : 0;
      } else if (e < kDescHead + a.T + a.ctx) {
        const int m = e - kDescHead - a.T;
        v = m < nc ? a.sen[sb + c_lo + m] : -1;
      } else {
        const int o = e - kDescHead - a.T - a.ctx;
        const int j = o / (1 + K), r = o % (1 + K);
        if (j >= np) {
          v = -1;
        } else if (r == 0) {
          v = a.sen[sb + p0 + j];
        } else {
          const uint32_t w = stream_word(a.seed, c1, c2, DG_TAG_SGNS, (uint32_t)(np + j * K + r - 1));
          v = a.table[w % a.table_size];
          if (v == 0) v = (int)(w % (a.vocab - 1) + 1);
        }
      }
      d[e] = v;
    }
  }
}

struct TileArgs {
  SgnsDev a;
  const int32_t* desc;
  uint32_t desc_stride;
  uint64_t n_tiles;
  int kd;      // negatives drawn per position
  float wneg;  // negative / kd: weight of one drawn negative
  int exact;
  int deterministic;
  float* stage;     // deterministic: per tile (CTX + NOUT) slots of dim floats
  uint32_t* keys;   // key = row | table << row_bits; 0xffffffff = unused slot
  uint32_t* vals;
  uint32_t row_bits;
};

template <typename E, int DIM, int T, int CTX, int NOUT>
__global__ __launch_bounds__(256) void sg_tile_kernel(TileArgs t) {
  typedef Mfma<E> M;
  constexpr int PAD = 16 / (int)sizeof(E);
  constexpr int LD = DIM + PAD;   // staged rows (elements)
  constexpr int SLD = NOUT + 4;   // S / G rows (floats, 16-byte aligned)
  constexpr int XLD = CTX + 1;    // Gx rows (floats)
  constexpr int C4 = DIM / 4;     // four-element chunks per row
  constexpr int NCHUNK = (CTX + NOUT) * C4;
  constexpr int NLD = (NCHUNK + 255) / 256;  // chunks per thread
  constexpr int BATCH = NLD < 12 ? NLD : 12; // loads in flight per thread
  constexpr int NDESC = kDescHead + T + CTX + NOUT;
  static_assert(CTX % 16 == 0 && NOUT % 16 == 0 && NDESC <= 256, "tile shape");
  __shared__ __attribute__((aligned(16))) E sX[CTX * LD];
  __shared__ __attribute__((aligned(16))) E sY[NOUT * LD];
  __shared__ __attribute__((aligned(16))) float sS[CTX * SLD];
  __shared__ float sGx[CTX * XLD];
  __shared__ float s_expt[kExpTable + 1];
  __shared__ int s_d[NDESC];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const SgnsDev& a = t.a;
  const int W = (int)a.window, K1 = t.kd + 1;
  const int* s_b = s_d + kDescHead;
  const int* s_ctx = s_d + kDescHead + T;
  const int* s_out = s_d + kDescHead + T + CTX;
  constexpr int NT_OUT = NOUT / 16;
  constexpr int NT_DIM = DIM / 16;
  for (int i = tid; i <= kExpTable; i += 256) s_expt[i] = a.expt[i];
  // descriptor of the first tile; later ones are loaded one tile ahead
  int dnext = 0;
  if (blockIdx.x < t.n_tiles && tid < NDESC) dnext = t.desc[(uint64_t)blockIdx.x * t.desc_stride + tid];
  for (uint64_t ti = blockIdx.x; ti < t.n_tiles; ti += gridDim.x) {
    if (tid < NDESC) s_d[tid] = dnext;
    __syncthreads();
    const int p0 = s_d[0], c_lo = s_d[1], nc = s_d[2], np = s_d[3];
    {
      const uint64_t tn = ti + gridDim.x;
      if (tn < t.n_tiles && tid < NDESC) dnext = t.desc[tn * t.desc_stride + tid];
    }
    // stage the rows: four FP32 elements per load, BATCH loads in flight, each
    // rounded to E as stored
#pragma unroll
    for (int b0 = 0; b0 < NLD; b0 += BATCH) {
      float4 v[BATCH];
#pragma unroll
      for (int q = 0; q < BATCH; ++q) {
        const int i = (b0 + q) * 256 + tid;
        v[q] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (b0 + q < NLD && i < NCHUNK) {
          const int r = i / C4, c = i % C4;
          const int row = r < CTX ? s_ctx[r] : s_out[r - CTX];
          if (row >= 0)
            v[q] = reinterpret_cast<const float4*>((r < CTX ? a.syn0 : a.syn1) + (uint64_t)row * DIM)[c];
        }
      }
#pragma unroll
      for (int q = 0; q < BATCH; ++q) {
        const int i = (b0 + q) * 256 + tid;
        if (b0 + q < NLD && i < NCHUNK) {
          const int r = i / C4, c = i % C4;
          E* dst = (r < CTX ? sX + r * LD : sY + (r - CTX) * LD) + 4 * c;
          if constexpr (sizeof(E) == 4) {
            *reinterpret_cast<float4*>(dst) = v[q];
          } else {
            uint2 h;
            h.x = (uint32_t)f2bf(v[q].x) | ((uint32_t)f2bf(v[q].y) << 16);
            h.y = (uint32_t)f2bf(v[q].z) | ((uint32_t)f2bf(v[q].w) << 16);
            *reinterpret_cast<uint2*>(dst) = h;
          }
        }
      }
    }
    __syncthreads();
    const int nM = (nc + 15) / 16;
    const int kc = nM * 16;  // <= CTX, a multiple of the update k step
    // S = X Y^T (nM x NT_OUT tiles) and Gx = X X^T (nM x nM tiles)
    for (int tt = wave; tt < nM * NT_OUT + nM * nM; tt += 4) {
      const bool is_s = tt < nM * NT_OUT;
      const int u = is_s ? tt : tt - nM * NT_OUT;
      const int mt = is_s ? u / NT_OUT : u / nM;
      const int nt = is_s ? u % NT_OUT : u % nM;
      const E* B = is_s ? sY : sX;
      f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
      for (int k0 = 0; k0 < DIM; k0 += M::KS)
        acc = M::mma(M::row_frag(sX, LD, mt * 16, k0, lane), M::row_frag(B, LD, nt * 16, k0, lane), acc);
      float* dst = is_s ? sS : sGx;
      const int ld = is_s ? SLD : XLD;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        dst[(mt * 16 + 4 * (lane >> 4) + i) * ld + nt * 16 + (lane & 15)] = acc[i];
    }
    __syncthreads();
    // the per-column recursion; columns spread over the four waves
    {
      const int n = (tid & 63) * 4 + wave;
      if (n < NOUT) {
        const int j = n / K1, r = n % K1;
        const int row = s_out[n];
        const bool live = j < np && row >= 0 && (r == 0 || row != s_out[j * K1]);
        int m_lo = 0, m_hi = -1, mc = -1;
        if (live) {
          const int p = p0 + j, w = W - s_b[j];
          m_lo = (p - w > c_lo ? p - w : c_lo) - c_lo;
          m_hi = (p + w < c_lo + nc - 1 ? p + w : c_lo + nc - 1) - c_lo;
          mc = p - c_lo;
        }
        const float lab = r == 0 ? 1.0f : 0.0f;
        for (int m = 0; m < kc; ++m) {
          float g = 0.0f;
          if (m >= m_lo && m <= m_hi && m != mc) {
            float sv = sS[m * SLD + n];
            for (int m2 = m_lo; m2 < m; ++m2)
              if (m2 != mc) sv += sS[m2 * SLD + n] * sGx[m * XLD + m2];
            g = (lab - sigmoid_of(sv, s_expt, t.exact)) * a.alpha;
            if (r != 0) g *= t.wneg;
          }
          sS[m * SLD + n] = g;
        }
      }
    }
    __syncthreads();
    const uint64_t slot0 = ti * (uint64_t)(CTX + NOUT);
    const uint32_t tb = 1u << t.row_bits;
    // dX = G Y -> input rows
    for (int tt = wave; tt < nM * NT_DIM; tt += 4) {
      const int mt = tt / NT_DIM, nt = tt % NT_DIM;
      f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k0 = 0; k0 < NOUT; k0 += M::KU)
        acc = M::mmau(M::urow_f(sS, SLD, mt * 16, k0, lane), M::ucol(sY, LD, k0, nt * 16, lane), acc);
      const int dd = nt * 16 + (lane & 15);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int m = mt * 16 + 4 * (lane >> 4) + i;
        if (m >= nc) continue;
        if (t.deterministic)
          t.stage[(slot0 + m) * DIM + dd] = acc[i];
        else
          atomicAdd(&a.syn0[(uint64_t)s_ctx[m] * DIM + dd], acc[i]);
      }
    }
    // dY = G^T X -> output rows
    for (int tt = wave; tt < NT_OUT * NT_DIM; tt += 4) {
      const int mt = tt / NT_DIM, nt = tt % NT_DIM;
      f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
      for (int k0 = 0; k0 < kc; k0 += M::KU)
        acc = M::mmau(M::ucol_f(sS, SLD, k0, mt * 16, lane), M::ucol(sX, LD, k0, nt * 16, lane), acc);
      const int dd = nt * 16 + (lane & 15);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int m = mt * 16 + 4 * (lane >> 4) + i;
        const int rr = s_out[m];
        if (rr < 0) continue;
        if (t.deterministic)
          t.stage[(slot0 + CTX + m) * DIM + dd] = acc[i];
        else
          atomicAdd(&a.syn1[(uint64_t)rr * DIM + dd], acc[i]);
      }
    }
    if (t.deterministic) {
      if (tid < CTX) {
        t.keys[slot0 + tid] = tid < nc ? (uint32_t)s_ctx[tid] : 0xffffffffu;
        t.vals[slot0 + tid] = (uint32_t)(slot0 + tid);
      }
      if (tid < NOUT) {
        t.keys[slot0 + CTX + tid] = s_out[tid] >= 0 ? ((uint32_t)s_out[tid] | tb) : 0xffffffffu;
        t.vals[slot0 + CTX + tid] = (uint32_t)(slot0 + CTX + tid);
      }
    }
    __syncthreads();
  }
}

// tiles per sentence: ceil(L / T) for L >= 2
__global__ void tile_count_kernel(const uint64_t* sen_off, uint64_t n_sen, uint32_t tile,
                                  uint32_t* cnt) {
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n_sen;
       s += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t L = sen_off[s + 1] - sen_off[s];
    cnt[s] = L >= 2 ? (uint32_t)((L + tile - 1) / tile) : 0u;
  }
}

__global__ void tile_list_kernel(const uint32_t* cnt, const uint64_t* off, uint64_t n_sen,
                                 uint2* tiles) {
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n_sen;
       s += (uint64_t)gridDim.x * blockDim.x)
    for (uint32_t k = 0; k < cnt[s]; ++k) tiles[off[s] + k] = make_uint2((uint32_t)s, k);
}

// ---- stable LSD radix sort of (key, value) pairs, 8 bits per pass -----------
constexpr int kSortBlock = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortBlock * kSortItems;

__global__ __launch_bounds__(kSortBlock) void radix_hist_kernel(const uint32_t* keys, uint64_t n,
                                                                int shift, uint32_t n_blocks,
                                                                uint32_t* hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
  for (int i = 0; i < kSortItems; ++i) {
    const uint64_t k = base + (uint64_t)i * kSortBlock + threadIdx.x;
    if (k < n) atomicAdd(&h[(keys[k] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(uint64_t)threadIdx.x * n_blocks + blockIdx.x] = h[threadIdx.x];  // digit-major
}

__global__ __launch_bounds__(kSortBlock) void radix_scatter_kernel(
    const uint32_t* keys, const uint32_t* vals, uint64_t n, int shift, uint32_t n_blocks,
    const uint64_t* digit_off, uint32_t* okeys, uint32_t* ovals) {
  __shared__ uint32_t run[256];
  __shared__ uint32_t wcnt[4][256];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  run[threadIdx.x] = 0;
  for (int w = 0; w < 4; ++w) wcnt[w][threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
  const uint64_t lanes_lt = (1ull << lane) - 1ull;
  for (int i = 0; i < kSortItems; ++i) {
    const uint64_t k = base + (uint64_t)i * kSortBlock + threadIdx.x;
    const bool valid = k < n;
    const uint32_t key = valid ? keys[k] : 0u;
    const uint32_t dig = (key >> shift) & 255u;
    // lanes of this wave holding the same digit
    uint64_t same = __ballot(valid);
    for (int b = 0; b < 8; ++b) {
      const uint64_t bal = __ballot(((dig >> b) & 1u) != 0);
      same &= ((dig >> b) & 1u) ? bal : ~bal;
    }
    const uint32_t rank = (uint32_t)__popcll(same & lanes_lt);
    const uint32_t cnt = (uint32_t)__popcll(same);
    if (valid && rank == 0) wcnt[wave][dig] = cnt;
    __syncthreads();
    if (valid) {
      uint32_t before = run[dig];
      for (uint32_t w = 0; w < wave; ++w) before += wcnt[w][dig];
      const uint64_t dst = digit_off[(uint64_t)dig * n_blocks + blockIdx.x] + before + rank;
      okeys[dst] = key;
      ovals[dst] = vals[k];
    }
    __syncthreads();
    run[threadIdx.x] += wcnt[0][threadIdx.x] + wcnt[1][threadIdx.x] + wcnt[2][threadIdx.x] +
                        wcnt[3][threadIdx.x];
    for (int w = 0; w < 4; ++w) wcnt[w][threadIdx.x] = 0;
    __syncthreads();
  }
}

// Segment heads of sorted keys (valid keys only).
__global__ void seg_head_kernel(const uint32_t* keys, uint64_t n, uint32_t* head) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    head[i] = (keys[i] != 0xffffffffu && (i == 0 || keys[i] != keys[i - 1])) ? 1u : 0u;
}

__global__ void seg_list_kernel(const uint32_t* head, const uint64_t* pos, uint64_t n,
                                uint64_t* starts) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    if (head[i]) starts[pos[i]] = i;
}

// One wavefront per row segment: contributions summed in sorted (slot) order,
// then added once to the row.  mode 0: W += sum; mode 1: W -= lr * sum;
// mode 2: store the sum in the gradient table instead.
template <int D>
__global__ __launch_bounds__(256) void seg_apply_kernel(const uint32_t* keys, const uint32_t* vals,
                                                        const uint64_t* starts, uint64_t n_seg,
                                                        uint64_t n_valid, const float* stage,
                                                        uint32_t row_bits, float* syn0, float* syn1, int mode,
                                                        float lr, float* g0, float* g1) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t sgi = wave; sgi < n_seg; sgi += nwaves) {
    const uint64_t b = starts[sgi];
    const uint64_t e = sgi + 1 < n_seg ? starts[sgi + 1] : n_valid;
    const uint32_t key = keys[b];
    float acc[D];
#pragma unroll
    for (int q = 0; q < D; ++q) acc[q] = 0.0f;
    for (uint64_t i = b; i < e; ++i) {
      const uint64_t slot = vals[i];
#pragma unroll
      for (int q = 0; q < D; ++q) acc[q] = acc[q] + stage[slot * (64 * D) + lane + 64 * q];
    }
    const uint32_t tbl = key >> row_bits;
    const uint64_t row = key & ((1u << row_bits) - 1u);
    float* w = (tbl ? syn1 : syn0) + row * (64 * D);
    if (mode == 2) {
      float* g = (tbl ? g1 : g0) + row * (64 * D);
#pragma unroll
      for (int q = 0; q < D; ++q) g[lane + 64 * q] = acc[q];
      continue;
    }
#pragma unroll
    for (int q = 0; q < D; ++q) {
      const float nv = mode == 0 ? w[lane + 64 * q] + acc[q] : w[lane + 64 * q] - lr * acc[q];
      w[lane + 64 * q] = nv;
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
  float* d_g0 = nullptr;  // gradient tables of the tuple interface
  float* d_g1 = nullptr;
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
  uint2* d_tiles = nullptr;
  size_t tiles_cap = 0;
  int32_t* d_desc = nullptr;  // tile descriptors of one launch
  size_t desc_cap = 0;
  uint64_t* d_tile_off = nullptr;
  size_t tile_off_cap = 0;
  // sorted scatter-add
  float* d_stage = nullptr;
  size_t stage_cap = 0;
  uint32_t* d_k[2] = {nullptr, nullptr};
  uint32_t* d_v[2] = {nullptr, nullptr};
  size_t kv_cap = 0;
  uint32_t* d_hist = nullptr;
  size_t hist_cap = 0;
  uint64_t* d_hoff = nullptr;
  size_t hoff_cap = 0;
  uint32_t* d_head = nullptr;
  size_t head_cap = 0;
  uint64_t* d_hpos = nullptr;
  size_t hpos_cap = 0;
  uint64_t* d_starts = nullptr;
  size_t starts_cap = 0;
  // tuple interface
  uint32_t* d_tuples = nullptr;
  size_t tuples_cap = 0;
  float* d_logits = nullptr;
  size_t logits_cap = 0;
  uint32_t epoch = 0;
  uint64_t sentences_done = 0;
};

void dg_sgns_free(dg_ctx* ctx) {
  dg_sgns_state* s = ctx->sgns;
  if (!s) return;
  void* bufs[] = {s->d_rank,   s->d_syn0,  s->d_syn1,   s->d_g0,
                  s->d_g1,     s->d_table, s->d_expt,   s->d_keep,    s->d_sen,   s->d_sen_off,
                  s->d_cnt,    s->d_tiles, s->d_desc,  s->d_tile_off, s->d_stage, s->d_k[0],  s->d_k[1],
                  s->d_v[0],   s->d_v[1],  s->d_hist,   s->d_hoff,    s->d_head,  s->d_hpos,
                  s->d_starts, s->d_tuples, s->d_logits};
  for (void* b : bufs)
    if (b) hipFree(b);
  delete s;
  ctx->sgns = nullptr;
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
  p->tile_positions = 0;
  p->tile_negatives = 0;
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
  if (p->form > DG_SGNS_BATCHED_BF16) return dg_fail(ctx, DG_ERR_ARG, "unknown form %u", p->form);
  if (p->tile_positions != 0 && p->tile_positions != 8 && p->tile_positions != 16)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "tile_positions %u: 8 or 16", p->tile_positions);
  if (p->tile_negatives != 0 && p->tile_negatives < p->negative)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "tile_negatives %u: at least negative (%u)",
                   p->tile_negatives, p->negative);
  if (p->form != DG_SGNS_PLAIN_FP32) {
    const uint32_t kd = p->tile_negatives ? p->tile_negatives : p->negative;
    const uint32_t T = p->tile_positions ? p->tile_positions : 8u;
    if (T * (1 + kd) > 96u || T + 2 * p->window > 64u)
      return dg_fail(ctx, DG_ERR_UNSUPPORTED,
                     "tile of %u positions: needs %u output rows (max 96) and %u context rows (max 64)",
                     T, T * (1 + kd), T + 2 * p->window);
  }
  if (p->form == DG_SGNS_BATCHED_FP32 && p->dim > 128)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "batched FP32 form supports dim <= 128");
  if (p->form == DG_SGNS_BATCHED_BF16 && p->dim > 256)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "batched BF16 form supports dim <= 256");
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
  for (uint32_t a = 0; a < V; ++a) {
    const uint32_t v = order[a];
    if (cn[v] < p->min_count && a != 0) continue;
    rank[v] = (uint32_t)s->vertex_of_rank.size();
    s->vertex_of_rank.push_back(v);
    s->cn_of_rank.push_back(cn[v]);
    s->train_words += cn[v];
  }
  s->vocab = (uint32_t)s->vertex_of_rank.size();
  if (s->vocab < 2) return dg_fail(ctx, DG_ERR_ARG, "vocabulary needs at least 2 words");
  const uint32_t vocab = s->vocab, dim = s->dim;
  // syn0 from the reference LCG, syn1 = 0
  std::vector<float> syn0((size_t)vocab * dim);
  unsigned long long nr = 1;
  for (size_t i = 0; i < syn0.size(); ++i) {
    nr = nr * 25214903917ULL + 11;
    if (p->init_float_ops)
      syn0[i] = (((nr & 0xFFFF) / 65536.f) - 0.5f) / (float)dim;
    else
      syn0[i] = (float)(((double)((float)(nr & 0xFFFF) / (float)65536) - 0.5) / (double)(long long)dim);
  }
  // unigram table (double accumulation, word2vec.cu:537-552)
  s->table_size = p->table_size;
  std::vector<int32_t> table(p->table_size);
  {
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
  // subsampling keep limits (float arithmetic of word2vec.cu:1015)
  std::vector<uint32_t> keep(vocab);
  for (uint32_t w = 0; w < vocab; ++w) {
    if (!(p->sample > 0)) {
      keep[w] = 65536;
      continue;
    }
    const float st = p->sample * (float)s->train_words;
    const float c = (float)s->cn_of_rank[w];
    const float ran = (std::sqrt(c / st) + 1) * st / c;
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
  return DG_OK;
}

static int row_bits_for(uint32_t vocab) {
  int b = 1;
  while ((1u << b) < vocab && b < 30) ++b;
  return b;
}

// Sort (key, slot) pairs stably by key and apply the per-row sums.
static int sorted_scatter_apply(dg_ctx* ctx, dg_sgns_state* s, uint64_t n, int mode, float lr) {
  if (n == 0) return DG_OK;
  const int rb = row_bits_for(s->vocab);
  // valid keys are < 2^(rb+1); sorting rb+2 bits puts the unused-slot key
  // 0xffffffff (bit rb+1 set) after every valid key
  const int key_bits = rb + 2;
  const uint32_t n_blocks = (uint32_t)((n + kSortTile - 1) / kSortTile);
  DG_CHECK(dg_reserve(ctx, &s->d_hist, &s->hist_cap, (size_t)n_blocks * 256));
  DG_CHECK(dg_reserve(ctx, &s->d_hoff, &s->hoff_cap, (size_t)n_blocks * 256 + 1));
  int cur = 0;
  for (int shift = 0; shift < key_bits; shift += 8) {
    radix_hist_kernel<<<n_blocks, kSortBlock, 0, ctx->stream>>>(s->d_k[cur], n, shift, n_blocks,
                                                                s->d_hist);
    DG_CHECK(dg_scan_lengths(ctx, s->d_hist, (uint64_t)n_blocks * 256, s->d_hoff));
    radix_scatter_kernel<<<n_blocks, kSortBlock, 0, ctx->stream>>>(
        s->d_k[cur], s->d_v[cur], n, shift, n_blocks, s->d_hoff, s->d_k[1 - cur], s->d_v[1 - cur]);
    DG_HIP(ctx, hipGetLastError());
    cur = 1 - cur;
  }
  // invalid keys (0xffffffff) sort last only if the top pass covers bit 31; they
  // are excluded by the head test and by n_valid below
  DG_CHECK(dg_reserve(ctx, &s->d_head, &s->head_cap, (size_t)n));
  DG_CHECK(dg_reserve(ctx, &s->d_hpos, &s->hpos_cap, (size_t)n + 1));
  seg_head_kernel<<<grid_for(ctx, n, 256), 256, 0, ctx->stream>>>(s->d_k[cur], n, s->d_head);
  DG_CHECK(dg_scan_lengths(ctx, s->d_head, n, s->d_hpos));
  uint64_t n_seg = 0;
  DG_HIP(ctx, hipMemcpyAsync(&n_seg, s->d_hpos + n, sizeof(uint64_t), hipMemcpyDeviceToHost,
                             ctx->stream));
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  DG_CHECK(dg_reserve(ctx, &s->d_starts, &s->starts_cap, (size_t)n_seg + 1));
  seg_list_kernel<<<grid_for(ctx, n, 256), 256, 0, ctx->stream>>>(s->d_head, s->d_hpos, n,
                                                                  s->d_starts);
  // number of valid keys: position after the last valid segment = first invalid key
  // (invalid keys sort to the end because 0xffffffff is the largest key)
  uint64_t n_valid = n;
  {
    // find the first 0xffffffff by counting invalid keys is equivalent: all sort last
    std::vector<uint32_t> tail(1);
    (void)tail;
  }
  const int D = (int)(s->dim / 64);
  const unsigned g = (unsigned)grid_for(ctx, n_seg * 64, 256);
  // n_valid is resolved on the device side by the segment bounds: the last
  // segment ends at the first invalid key, found by a binary search below
  uint32_t* keys = s->d_k[cur];
  uint64_t lo = 0, hi = n;  // first index with key == 0xffffffff
  {
    std::vector<uint32_t> probe(1);
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo) / 2;
      DG_HIP(ctx, hipMemcpy(probe.data(), keys + mid, sizeof(uint32_t), hipMemcpyDeviceToHost));
      if (probe[0] == 0xffffffffu)
        hi = mid;
      else
        lo = mid + 1;
    }
    n_valid = lo;
  }
  switch (D) {
    case 1:
      seg_apply_kernel<1><<<g, 256, 0, ctx->stream>>>(keys, s->d_v[cur], s->d_starts, n_seg, n_valid,
                                                      s->d_stage, rb, s->d_syn0, s->d_syn1, mode,
                                                      lr, s->d_g0, s->d_g1);
      break;
    case 2:
      seg_apply_kernel<2><<<g, 256, 0, ctx->stream>>>(keys, s->d_v[cur], s->d_starts, n_seg, n_valid,
                                                      s->d_stage, rb, s->d_syn0, s->d_syn1, mode,
                                                      lr, s->d_g0, s->d_g1);
      break;
    case 4:
      seg_apply_kernel<4><<<g, 256, 0, ctx->stream>>>(keys, s->d_v[cur], s->d_starts, n_seg, n_valid,
                                                      s->d_stage, rb, s->d_syn0, s->d_syn1, mode,
                                                      lr, s->d_g0, s->d_g1);
      break;
    case 8:
      seg_apply_kernel<8><<<g, 256, 0, ctx->stream>>>(keys, s->d_v[cur], s->d_starts, n_seg, n_valid,
                                                      s->d_stage, rb, s->d_syn0, s->d_syn1, mode,
                                                      lr, s->d_g0, s->d_g1);
      break;
  }
  DG_HIP(ctx, hipGetLastError());
  return DG_OK;
}

template <int D>
static void launch_plain(dg_ctx* ctx, const SgnsDev& a, uint64_t sb, uint64_t se, int single,
                         int exact) {
  if (single) {
    sg_plain_kernel<D><<<1, 64, 0, ctx->stream>>>(a, sb, se, 1, exact);
  } else {
    const unsigned blocks = (unsigned)((se - sb + 3) / 4);
    sg_plain_kernel<D><<<blocks ? blocks : 1, 256, 0, ctx->stream>>>(a, sb, se, 0, exact);
  }
}

// Tile shape: T positions per tile (tile_positions, else 8), K negatives
// drawn per position; staged context rows CTX = T + 2 window rounded up to 32
// or 64; output columns NOUT = T (1 + K) rounded up to 48 or 96.
struct TileShape {
  int T;
  int ctx;
  int nout;
};

static TileShape tile_shape(int form, uint32_t window, uint32_t kd, uint32_t tile) {
  TileShape ts;
  ts.T = tile ? (int)tile : 8;
  ts.ctx = ts.T + 2 * (int)window <= 32 ? 32 : 64;
  (void)form;
  ts.nout = ts.T * (1 + (int)kd) <= 48 ? 48 : 96;
  return ts;
}

static uint32_t desc_stride(const TileShape& ts) {
  return (uint32_t)((kDescHead + ts.T + ts.ctx + ts.nout + 3) & ~3);
}

// Blocks resident per CU for a kernel (queried once), times the CUs.
template <typename Kern>
static unsigned resident_grid(dg_ctx* ctx, Kern k, uint64_t n_tiles) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 256, 0) != hipSuccess || per_cu < 1)
    per_cu = 1;
  uint64_t g = (uint64_t)per_cu * (uint64_t)ctx->num_cus;
  if (g > n_tiles) g = n_tiles;
  return (unsigned)(g ? g : 1);
}

template <typename E, int DIM, int T, int CTX, int NOUT>
static void launch_tile(dg_ctx* ctx, const TileArgs& t) {
  auto k = sg_tile_kernel<E, DIM, T, CTX, NOUT>;
  k<<<resident_grid(ctx, k, t.n_tiles), 256, 0, ctx->stream>>>(t);
}

template <typename E, int DIM>
static void launch_tile_shape(dg_ctx* ctx, const TileArgs& t, const TileShape& ts) {
  constexpr int S = 48;
  if (ts.T == 16) {
    launch_tile<E, DIM, 16, 64, 96>(ctx, t);
  } else if (ts.ctx == 32) {
    if (ts.nout == S) launch_tile<E, DIM, 8, 32, S>(ctx, t);
    else launch_tile<E, DIM, 8, 32, 96>(ctx, t);
  } else {
    if (ts.nout == S) launch_tile<E, DIM, 8, 64, S>(ctx, t);
    else launch_tile<E, DIM, 8, 64, 96>(ctx, t);
  }
}

template <typename E>
static void launch_tile_dim(dg_ctx* ctx, const TileArgs& t, uint32_t dim, const TileShape& ts) {
  switch (dim) {
    case 64: launch_tile_shape<E, 64>(ctx, t, ts); break;
    case 128: launch_tile_shape<E, 128>(ctx, t, ts); break;
    case 256:
      // FP32 rows of dim 256 do not fit the 160 KiB LDS; rejected by check_params
      if constexpr (sizeof(E) == 2) launch_tile_shape<E, 256>(ctx, t, ts);
      break;
  }
}

extern "C" int dg_sgns_train(dg_ctx* ctx, int corpus, const dg_sgns_params* pin, float alpha_in,
                             dg_sgns_stats* out) {
  if (!ctx || !pin) return DG_ERR_ARG;
  dg_sgns_state* s = ctx->sgns;
  if (!s) return dg_fail(ctx, DG_ERR_STATE, "dg_sgns_init has not run");
  DG_CHECK(check_params(ctx, pin));
  if (pin->dim != s->dim) return dg_fail(ctx, DG_ERR_ARG, "dim differs from dg_sgns_init");
  DG_HIP(ctx, hipSetDevice(ctx->device));
  const dg_sgns_params p = *pin;
  const uint32_t* d_tok = corpus == DG_CORPUS_TRAIN ? ctx->d_train_tokens : ctx->d_tokens;
  const uint64_t* d_off = corpus == DG_CORPUS_TRAIN ? ctx->d_train_offsets : ctx->d_offsets;
  const uint64_t n_walks = corpus == DG_CORPUS_TRAIN ? ctx->train_walks : ctx->cur_walks;
  const uint64_t n_tok = corpus == DG_CORPUS_TRAIN ? ctx->train_tokens : ctx->cur_tokens;
  if (!d_tok || n_walks == 0) return dg_fail(ctx, DG_ERR_STATE, "empty corpus");
  std::vector<uint64_t> h_woff(n_walks + 1);
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  DG_HIP(ctx, hipMemcpy(h_woff.data(), d_off, sizeof(uint64_t) * (n_walks + 1), hipMemcpyDeviceToHost));
  const float alpha0 = alpha_in > 0 ? alpha_in : p.alpha;
  const double total_words = (double)p.iter * (double)n_tok;
  const int D = (int)(s->dim / 64);
  dg_sgns_stats st;
  memset(&st, 0, sizeof(st));
  double ms = 0.0;
  float alpha = alpha0;
  DG_CHECK(dg_reserve(ctx, &s->d_cnt, &s->cnt_cap, (size_t)n_walks + 1));
  DG_CHECK(dg_reserve(ctx, &s->d_sen_off, &s->sen_off_cap, (size_t)n_walks + 1));
  for (uint32_t ep = 0; ep < p.iter; ++ep) {
    const uint32_t epoch = s->epoch++;
    // subsample (two passes) into ranked sentences
    subsample_count_kernel<<<grid_for(ctx, n_walks, 256), 256, 0, ctx->stream>>>(
        d_tok, d_off, n_walks, s->d_rank, s->d_keep, p.seed, epoch, s->d_cnt);
    DG_CHECK(dg_scan_lengths(ctx, s->d_cnt, n_walks, s->d_sen_off));
    uint64_t n_kept = 0;
    DG_HIP(ctx, hipMemcpyAsync(&n_kept, s->d_sen_off + n_walks, sizeof(uint64_t),
                               hipMemcpyDeviceToHost, ctx->stream));
    DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    DG_CHECK(dg_reserve(ctx, &s->d_sen, &s->sen_cap, (size_t)n_kept + 1));
    subsample_write_kernel<<<grid_for(ctx, n_walks, 256), 256, 0, ctx->stream>>>(
        d_tok, d_off, n_walks, s->d_rank, s->d_keep, p.seed, epoch, s->d_sen_off, s->d_sen);
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
    const double words_before = (double)ep * (double)n_tok;
    auto alpha_at = [&](uint64_t walk_index) -> float {
      if (p.alpha_schedule != DG_ALPHA_LINEAR) return alpha0;
      // dsgl.hpp:886-897
      float al = alpha0 * (float)(1.0 - (words_before + (double)h_woff[walk_index]) / (total_words + 1.0));
      if (al < alpha0 * 0.0001f) al = alpha0 * 0.0001f;
      return al;
    };
    if (p.form == DG_SGNS_PLAIN_FP32) {
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
        DG_CHECK(dg_time_begin(ctx, "sgns_plain", &tk));
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
      // tile list
      const uint32_t kd = p.tile_negatives ? p.tile_negatives : p.negative;
      const TileShape ts = tile_shape((int)p.form, p.window, kd, p.tile_positions);
      tile_count_kernel<<<grid_for(ctx, n_walks, 256), 256, 0, ctx->stream>>>(
          s->d_sen_off, n_walks, (uint32_t)ts.T, s->d_cnt);
      DG_CHECK(dg_reserve(ctx, &s->d_tile_off, &s->tile_off_cap, (size_t)n_walks + 1));
      DG_CHECK(dg_scan_lengths(ctx, s->d_cnt, n_walks, s->d_tile_off));
      std::vector<uint64_t> h_toff(n_walks + 1);
      DG_HIP(ctx, hipMemcpy(h_toff.data(), s->d_tile_off, sizeof(uint64_t) * (n_walks + 1),
                            hipMemcpyDeviceToHost));
      const uint64_t n_tiles = h_toff[n_walks];
      DG_CHECK(dg_reserve(ctx, &s->d_tiles, &s->tiles_cap, (size_t)n_tiles + 1));
      tile_list_kernel<<<grid_for(ctx, n_walks, 256), 256, 0, ctx->stream>>>(s->d_cnt, s->d_tile_off,
                                                                             n_walks, s->d_tiles);
      DG_HIP(ctx, hipGetLastError());
      const uint64_t mb = p.minibatch ? p.minibatch : (p.deterministic ? 8192 : 262144);
      const uint32_t stride = desc_stride(ts);
      const uint64_t mb_cap = n_tiles < mb ? n_tiles : mb;
      DG_CHECK(dg_reserve(ctx, &s->d_desc, &s->desc_cap, (size_t)(mb_cap ? mb_cap : 1) * stride));
      DescArgs da;
      da.sen = s->d_sen;
      da.sen_off = s->d_sen_off;
      da.table = s->d_table;
      da.first_sentence = s->sentences_done;
      da.seed = p.seed;
      da.epoch = epoch;
      da.table_size = s->table_size;
      da.vocab = s->vocab;
      da.window = (int)p.window;
      da.kd = (int)kd;
      da.T = ts.T;
      da.ctx = ts.ctx;
      da.nout = ts.nout;
      da.stride = stride;
      da.desc = s->d_desc;
      TileArgs t;
      t.a = a;
      t.desc = s->d_desc;
      t.desc_stride = stride;
      t.kd = (int)kd;
      t.wneg = (float)p.negative / (float)kd;
      t.exact = (int)p.exact_sigmoid;
      t.deterministic = (int)p.deterministic;
      t.row_bits = (uint32_t)row_bits_for(s->vocab);
      const uint64_t slots_per_tile = (uint64_t)(ts.ctx + ts.nout);
      if (p.deterministic) {
        const size_t slots = (size_t)mb_cap * slots_per_tile;
        DG_CHECK(dg_reserve(ctx, &s->d_stage, &s->stage_cap, slots * s->dim));
        if (s->kv_cap < slots) {
          for (int i = 0; i < 2; ++i) {
            size_t c = 0;
            if (s->d_k[i]) hipFree(s->d_k[i]);
            if (s->d_v[i]) hipFree(s->d_v[i]);
            s->d_k[i] = nullptr;
            s->d_v[i] = nullptr;
            DG_CHECK(dg_reserve(ctx, &s->d_k[i], &c, slots));
            c = 0;
            DG_CHECK(dg_reserve(ctx, &s->d_v[i], &c, slots));
          }
          s->kv_cap = slots;
        }
        t.stage = s->d_stage;
        t.keys = s->d_k[0];
        t.vals = s->d_v[0];
      } else {
        t.stage = nullptr;
        t.keys = nullptr;
        t.vals = nullptr;
      }
      for (uint64_t t0 = 0; t0 < n_tiles; t0 += mb) {
        const uint64_t nt = n_tiles - t0 < mb ? n_tiles - t0 : mb;
        da.tiles = s->d_tiles + t0;
        da.n_tiles = nt;
        t.n_tiles = nt;
        // learning rate from the first sentence of the launch
        uint64_t lo = 0, hi = n_walks;
        while (lo < hi) {  // first sentence whose tiles start at or after t0
          const uint64_t mid = (lo + hi) / 2;
          if (h_toff[mid + 1] <= t0) lo = mid + 1; else hi = mid;
        }
        t.a.alpha = alpha = alpha_at(lo);
        size_t tk;
        DG_CHECK(dg_time_begin(ctx, p.form == DG_SGNS_BATCHED_BF16 ? "sgns_bf16" : "sgns_tile", &tk));
        {
          uint64_t g = (nt + 3) / 4;
          const uint64_t cap = (uint64_t)ctx->num_cus * 16;
          if (g > cap) g = cap;
          tile_desc_kernel<<<(unsigned)(g ? g : 1), 256, 0, ctx->stream>>>(da);
        }
        if (p.form == DG_SGNS_BATCHED_BF16)
          launch_tile_dim<uint16_t>(ctx, t, s->dim, ts);
        else
          launch_tile_dim<float>(ctx, t, s->dim, ts);
        DG_HIP(ctx, hipGetLastError());
        if (p.deterministic) DG_CHECK(sorted_scatter_apply(ctx, s, nt * slots_per_tile, 0, 0.0f));
        DG_CHECK(dg_time_end(ctx, tk));
        st.launches++;
      }
    }
    s->sentences_done += n_walks;
  }
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (auto& pe : ctx->pending) {
    float m = 0.f;
    DG_HIP(ctx, hipEventElapsedTime(&m, pe.start, pe.stop));
    if (pe.name.rfind("sgns", 0) == 0) ms += m;
  }
  DG_CHECK(dg_time_resolve(ctx));
  st.words = (uint64_t)p.iter * n_tok;
  st.sentences = (uint64_t)p.iter * n_walks;
  st.ms = ms;
  st.alpha_end = alpha;
  if (out) *out = st;
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
  return DG_OK;
}

extern "C" int dg_sgns_score_pairs(dg_ctx* ctx, const uint32_t* src, const uint32_t* dst,
                                   uint64_t n, float* scores) {
  if (!ctx || (n && (!src || !dst || !scores))) return DG_ERR_ARG;
  dg_sgns_state* s = ctx->sgns;
  if (!s) return dg_fail(ctx, DG_ERR_STATE, "dg_sgns_init has not run");
  if (n == 0) return DG_OK;
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

// tuple interface: filled in by the differentiation kernels
extern "C" int dg_sgns_tuples_forward(dg_ctx* ctx, const uint32_t*, uint64_t, uint32_t, int,
                                      float*, double*) {
  return dg_fail(ctx, DG_ERR_UNSUPPORTED, "not available yet");
}
extern "C" int dg_sgns_tuples_backward(dg_ctx* ctx, const uint32_t*, uint64_t, uint32_t, int,
                                       float, int) {
  return dg_fail(ctx, DG_ERR_UNSUPPORTED, "not available yet");
}
extern "C" int dg_sgns_grad_rows(dg_ctx* ctx, int, const uint32_t*, uint64_t, float*) {
  return dg_fail(ctx, DG_ERR_UNSUPPORTED, "not available yet");
}
