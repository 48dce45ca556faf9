// tuples.hip - the differentiable SGNS tuple interface on gfx950: an explicit
// forward (logits, loss) and a hand-written adjoint whose row gradients are
// summed by a sorted-key segmented reduction, so that the result does not
// depend on the schedule (no Hogwild races).
//
// A tuple is (center, context, neg_1..neg_K) in table rows; x = syn0[context],
// y_0 = syn1[center], y_k = syn1[neg_k]; z_j = x . y_j and
//   L = sum_t [ softplus(-z_0) + sum_k softplus(z_k) ].
// skip_gram_kernel (src/word2vec.cu:359-396) takes one SGD step on the same
// pairs in place: g_j = (label_j - sig(z_j)) * alpha, y_j += g_j x,
// x += sum_j g_j y_j, i.e. W -= alpha * dL/dW for one tuple at a time; a
// negative equal to the centre is skipped (src/word2vec.cu:366), here too.
//
// Pipeline of a backward call (n tuples, N = n (2 + K) gradient entries):
//   1. tuple_fwd_kernel<bwd>: one wavefront per tuple gathers x and the 1+K
//      y rows (D = dim/64 consecutive elements per lane), reduces the 1+K
//      dots together (a butterfly that halves the value set at each step),
//      and writes dL/dz, the context row's gradient sum_j dz_j y_j (a staged
//      row per tuple) and N (key, entry) pairs: key = table bit | row.
//   2. stable LSD radix sort of the pairs by key, 8 bits per pass: entries
//      of one row become a segment, in tuple order.
//   3. tuple_seg_kernel: the sorted entries are cut into blocks of 64; one
//      wavefront sums each block's runs of equal keys in order.  A run that
//      starts and ends inside its block is a finished row gradient; a run
//      that crosses a block edge leaves its partial sum, and
//      tuple_fix_kernel adds a crossing row's partials in block order.
//      Hot rows (Zipf heads) are thus split over many wavefronts with a
//      fixed summation order.
//   4. tuple_apply_kernel (optional): W -= lr * grad per row, after every
//      gradient is read from the old rows; the BF16 shadow rows follow.
// BF16 form: x and y are read from BF16 copies of the tables (half the
// gather bytes), products and sums in FP32, updates to the FP32 rows.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "dg_internal.h"

namespace {

constexpr int kSegBlock = 64;  // sorted entries per wavefront in the segmented sum
constexpr int kMaxTupleNeg = 15;

__device__ __forceinline__ uint16_t to_bf16(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// D consecutive elements of a row starting at element `first`, as floats.
template <typename E, int D>
struct RowIO;

template <int D>
struct RowIO<float, D> {
  __device__ static void load(const float* p, float (&v)[D]) {
    if constexpr (D == 1) {
      v[0] = p[0];
    } else if constexpr (D == 2) {
      const float2 q = *reinterpret_cast<const float2*>(p);
      v[0] = q.x;
      v[1] = q.y;
    } else {
      static_assert(D % 4 == 0, "D");
#pragma unroll
      for (int i = 0; i < D; i += 4) {
        const float4 q = *reinterpret_cast<const float4*>(p + i);
        v[i] = q.x;
        v[i + 1] = q.y;
        v[i + 2] = q.z;
        v[i + 3] = q.w;
      }
    }
  }
};

template <int D>
struct RowIO<uint16_t, D> {
  __device__ static void load(const uint16_t* p, float (&v)[D]) {
    if constexpr (D == 1) {
      v[0] = __uint_as_float((uint32_t)p[0] << 16);
    } else if constexpr (D == 2) {
      const uint32_t q = *reinterpret_cast<const uint32_t*>(p);
      v[0] = __uint_as_float(q << 16);
      v[1] = __uint_as_float(q & 0xffff0000u);
    } else if constexpr (D % 8 == 0) {
#pragma unroll
      for (int i = 0; i < D; i += 8) {
        const uint4 q = *reinterpret_cast<const uint4*>(p + i);
        const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          v[i + 2 * k] = __uint_as_float(w[k] << 16);
          v[i + 2 * k + 1] = __uint_as_float(w[k] & 0xffff0000u);
        }
      }
    } else {
      static_assert(D % 4 == 0, "D");
#pragma unroll
      for (int i = 0; i < D; i += 4) {
        const uint2 q = *reinterpret_cast<const uint2*>(p + i);
        v[i] = __uint_as_float(q.x << 16);
        v[i + 1] = __uint_as_float(q.x & 0xffff0000u);
        v[i + 2] = __uint_as_float(q.y << 16);
        v[i + 3] = __uint_as_float(q.y & 0xffff0000u);
      }
    }
  }
};

template <int D>
__device__ __forceinline__ void store_row(float* p, const float (&v)[D]) {
  if constexpr (D == 1) {
    p[0] = v[0];
  } else if constexpr (D == 2) {
    *reinterpret_cast<float2*>(p) = make_float2(v[0], v[1]);
  } else {
#pragma unroll
    for (int i = 0; i < D; i += 4)
      *reinterpret_cast<float4*>(p + i) = make_float4(v[i], v[i + 1], v[i + 2], v[i + 3]);
  }
}

// Sum NV per-lane values over each group of G lanes (G = 32 or 16) at once.
// Each butterfly step exchanges half of the remaining values with the lane
// `off` away, so after log2(NV) steps every lane holds one value index,
// idx = (lane % G) / (G / NV), summed over the lanes that differ in the
// exchanged bits; the remaining steps finish the sum.  Every lane of a group
// then holds its group's full sum of value idx.
template <int NV, int G>
__device__ __forceinline__ float group_multi_sum(float (&p)[NV], uint32_t lane) {
  static_assert(NV <= G, "values per group");
#pragma unroll
  for (int h = NV / 2, off = G / 2; h >= 1; h >>= 1, off >>= 1) {
    const bool up = (lane & (uint32_t)off) != 0;
#pragma unroll
    for (int i = 0; i < h; ++i) {
      const float keep = up ? p[i + h] : p[i];
      const float send = up ? p[i] : p[i + h];
      p[i] = keep + __shfl_xor(send, off, 64);
    }
  }
  float v = p[0];
#pragma unroll
  for (int off = G / 2 / NV; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ float softplus_f(float a) {
  return a > 0.0f ? a + log1pf(expf(-a)) : log1pf(expf(a));
}

struct TupleArgs {
  const uint32_t* tuples;  // n x (2 + K)
  uint64_t n;
  uint32_t K;
  const void* xt;          // syn0 (float or BF16 shadow)
  const void* yt;          // syn1
  float* logits;           // forward: n x (1 + K)
  float* tloss;            // forward: n
  const float* dlogits;    // backward: upstream dL/dz, or null for the loss's own
  float* dz;               // backward: n x (1 + K)
  float* stage;            // backward: n x dim, context-row gradient per tuple
  uint32_t* keys;          // backward: N
  uint32_t* vals;          // backward: N
  uint32_t row_bits;
};

// 64 / G tuples per wavefront, one per group of G lanes (D = dim / G
// consecutive elements per lane): G = 32 makes an FP32 row one 16-byte load
// per lane at dim 128, G = 16 does the same for a BF16 row (four tuples per
// wavefront, twice the rows in flight).  The next group of tuples' ids are
// loaded before the current one is reduced.
// BWD = false: logits and per-tuple loss.  BWD = true: dL/dz, the staged
// context-row gradient and the sort pairs (entries [0, n) for syn0 rows,
// [n, N) for syn1 rows, in tuple order).
template <typename E, int DIM, int NV, bool BWD, int G>
__global__ __launch_bounds__(256) void tuple_fwd_kernel(TupleArgs a) {
  constexpr int T = 64 / G;                 // tuples per wavefront
  constexpr int D = DIM / G;
  constexpr int kStride = G / NV;           // lanes per value index after the sum
  constexpr bool kKeepY = NV * D <= 64;     // else the stage pass reloads y rows
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t hl = lane & (uint32_t)(G - 1), hb = lane & ~(uint32_t)(G - 1);  // lane within the group, group base
  const uint32_t grp = lane / G;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const E* xt = static_cast<const E*>(a.xt);
  const E* yt = static_cast<const E*>(a.yt);
  const uint32_t K1 = a.K + 1, W = a.K + 2;
  const uint32_t idx = hl / kStride;  // value index this lane ends with
  const uint64_t groups = (a.n + T - 1) / T;
  uint64_t pr = wave;
  uint32_t id = 0;
  if (pr < groups) {
    const uint64_t t = (uint64_t)T * pr + grp;
    id = (t < a.n && hl < W) ? a.tuples[t * W + hl] : 0u;
  }
  for (; pr < groups; pr += nwaves) {
    const uint64_t t = (uint64_t)T * pr + grp;
    const bool tv = t < a.n;  // this group has a tuple (the last wavefront may be short)
    const uint32_t center = __shfl(id, (int)hb, 64);
    const uint32_t ctx = __shfl(id, (int)hb + 1, 64);
    const uint32_t my_target = __shfl(id, (int)hb + (idx == 0 ? 0 : (int)(1 + idx) & (G - 1)), 64);
    const bool my_live = tv && idx < K1 && (idx == 0 || my_target != center);
    float x[D];
    if (tv) RowIO<E, D>::load(xt + (uint64_t)ctx * DIM + hl * D, x);
    float y[kKeepY ? NV : 1][D];
    float p[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      p[j] = 0.0f;
      if ((uint32_t)j < K1) {
        const uint32_t r = __shfl(id, (int)hb + (j == 0 ? 0 : 1 + j), 64);
        float yj[D];
        if (tv) {
          RowIO<E, D>::load(yt + (uint64_t)r * DIM + hl * D, yj);
#pragma unroll
          for (int q = 0; q < D; ++q) p[j] += x[q] * yj[q];
        }
        if constexpr (kKeepY) {
#pragma unroll
          for (int q = 0; q < D; ++q) y[j][q] = tv ? yj[q] : 0.0f;
        }
      }
    }
    // ids of the next group of tuples, in flight during the sums below
    const uint64_t pn = pr + nwaves;
    uint32_t id_next = 0;
    if (pn < groups) {
      const uint64_t tn = (uint64_t)T * pn + grp;
      id_next = (tn < a.n && hl < W) ? a.tuples[tn * W + hl] : 0u;
    }
    const float z = group_multi_sum<NV, G>(p, lane);
    const bool writer = tv && (hl % kStride) == 0 && idx < K1;
    if constexpr (!BWD) {
      if (writer) a.logits[t * K1 + idx] = z;
      // loss: one lane per live value index, summed in index order
      const float l = (writer && my_live) ? softplus_f(idx == 0 ? -z : z) : 0.0f;
      float acc = 0.0f;
#pragma unroll
      for (int j = 0; j < NV; ++j) acc += __shfl(l, (int)hb + j * kStride, 64);
      if (tv && hl == 0) a.tloss[t] = acc;
    } else {
      float dz = 0.0f;
      if (my_live) {
        if (a.dlogits)
          dz = a.dlogits[t * K1 + idx];
        else
          dz = 1.0f / (1.0f + expf(-z)) - (idx == 0 ? 1.0f : 0.0f);
      }
      if (writer) {
        a.dz[t * K1 + idx] = dz;
        a.keys[a.n + t * K1 + idx] = (1u << a.row_bits) | my_target;
        a.vals[a.n + t * K1 + idx] = (uint32_t)(t * K1 + idx);
      }
      float c[D];
#pragma unroll
      for (int q = 0; q < D; ++q) c[q] = 0.0f;
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        if ((uint32_t)j < K1) {
          const float gj = __shfl(dz, (int)hb + j * kStride, 64);
          if constexpr (kKeepY) {
#pragma unroll
            for (int q = 0; q < D; ++q) c[q] += gj * y[j][q];
          } else {
            const uint32_t r = __shfl(id, (int)hb + (j == 0 ? 0 : 1 + j), 64);
            if (tv) {
              float yj[D];
              RowIO<E, D>::load(yt + (uint64_t)r * DIM + hl * D, yj);
#pragma unroll
              for (int q = 0; q < D; ++q) c[q] += gj * yj[q];
            }
          }
        }
      }
      if (tv) {
        store_row<D>(a.stage + t * DIM + hl * D, c);
        if (hl == 0) {
          a.keys[t] = ctx;
          a.vals[t] = (uint32_t)t;
        }
      }
    }
    id = id_next;
  }
}

// Deterministic loss in two levels: block k sums the k-th of kLossParts
// contiguous chunks (threads strided inside it, then a fixed tree), and one
// block sums the chunk totals with the same tree.
constexpr int kLossParts = 256;

__device__ double block_tree_sum(double v, double* part) {
  part[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  return part[0];
}

__global__ __launch_bounds__(256) void loss_part_kernel(const float* tloss, uint64_t n, double* parts) {
  __shared__ double part[256];
  const uint64_t lo = n * blockIdx.x / kLossParts, hi = n * (blockIdx.x + 1) / kLossParts;
  double s = 0.0;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += 256) s += (double)tloss[i];
  const double t = block_tree_sum(s, part);
  if (threadIdx.x == 0) parts[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void loss_final_kernel(const double* parts, double* out) {
  __shared__ double part[256];
  const double t = block_tree_sum(parts[threadIdx.x], part);
  if (threadIdx.x == 0) *out = t;
}

// ---- stable LSD radix sort of (key, value) pairs, 8 bits per pass -----------
constexpr int kSortBlock = 256;
constexpr int kSortItems = 4;
constexpr int kSortTile = kSortBlock * kSortItems;

// Per-tile digit counts, stored digit-major so that one exclusive scan gives
// every (digit, tile) its output offset.
__global__ __launch_bounds__(kSortBlock) void radix_hist_kernel(const uint32_t* keys, uint64_t n,
                                                                int shift, uint32_t n_tiles,
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
  hist[(uint64_t)threadIdx.x * n_tiles + blockIdx.x] = h[threadIdx.x];
}

// Per digit: exclusive scan of its per-tile counts (contiguous in the
// digit-major histogram) into offsets within the digit, and the digit total.
// One workgroup per digit.
__global__ __launch_bounds__(256) void radix_digit_scan_kernel(const uint32_t* hist, uint32_t n_tiles,
                                                               uint32_t* tile_off, uint32_t* digit_total) {
  __shared__ uint32_t sc[256];
  const uint32_t d = blockIdx.x;
  const uint32_t* h = hist + (uint64_t)d * n_tiles;
  uint32_t* o = tile_off + (uint64_t)d * n_tiles;
  uint32_t carry = 0;
  for (uint32_t c0 = 0; c0 < n_tiles; c0 += 256) {
    const uint32_t k = c0 + threadIdx.x;
    const uint32_t v = k < n_tiles ? h[k] : 0u;
    sc[threadIdx.x] = v;
    __syncthreads();
    for (int st = 1; st < 256; st <<= 1) {
      const uint32_t add = threadIdx.x >= (unsigned)st ? sc[threadIdx.x - st] : 0u;
      __syncthreads();
      sc[threadIdx.x] += add;
      __syncthreads();
    }
    if (k < n_tiles) o[k] = carry + sc[threadIdx.x] - v;
    carry += sc[255];
    __syncthreads();
  }
  if (threadIdx.x == 0) digit_total[d] = carry;
}

// Scatter a tile in input order: within a wave, lanes with the same digit are
// ranked by a ballot match on the digit bits; waves and rounds are ordered
// through LDS counters, so equal keys keep their input order.  The digit
// bases are the exclusive scan of the 256 digit totals, done per workgroup.
__global__ __launch_bounds__(kSortBlock) void radix_scatter_kernel(
    const uint32_t* keys, const uint32_t* vals, uint64_t n, int shift, uint32_t n_tiles,
    const uint32_t* tile_off, const uint32_t* digit_total, uint32_t* okeys, uint32_t* ovals) {
  __shared__ uint32_t run[256];
  __shared__ uint32_t base[256];
  __shared__ uint32_t wcnt[kSortBlock / 64][256];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t tot = digit_total[threadIdx.x];
  base[threadIdx.x] = tot;
  run[threadIdx.x] = 0;
  for (int w = 0; w < kSortBlock / 64; ++w) wcnt[w][threadIdx.x] = 0;
  __syncthreads();
  for (int st = 1; st < 256; st <<= 1) {
    const uint32_t add = threadIdx.x >= (unsigned)st ? base[threadIdx.x - st] : 0u;
    __syncthreads();
    base[threadIdx.x] += add;
    __syncthreads();
  }
  // inclusive -> exclusive, plus this tile's offset within the digit
  const uint32_t my_base = base[threadIdx.x] - tot + tile_off[(uint64_t)threadIdx.x * n_tiles + blockIdx.x];
  __syncthreads();
  base[threadIdx.x] = my_base;
  __syncthreads();
  const uint64_t tile0 = (uint64_t)blockIdx.x * kSortTile;
  const uint64_t lanes_lt = (1ull << lane) - 1ull;
  for (int i = 0; i < kSortItems; ++i) {
    const uint64_t k = tile0 + (uint64_t)i * kSortBlock + threadIdx.x;
    const bool valid = k < n;
    const uint32_t key = valid ? keys[k] : 0u;
    const uint32_t dig = (key >> shift) & 255u;
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
      const uint64_t dst = (uint64_t)base[dig] + before + rank;
      okeys[dst] = key;
      ovals[dst] = vals[k];
    }
    __syncthreads();
    uint32_t add = 0;
    for (int w = 0; w < kSortBlock / 64; ++w) add += wcnt[w][threadIdx.x];
    run[threadIdx.x] += add;
    for (int w = 0; w < kSortBlock / 64; ++w) wcnt[w][threadIdx.x] = 0;
    __syncthreads();
  }
}

// Row segments (runs of equal sorted keys) starting in each block of
// kSegBlock sorted entries; their exclusive scan numbers the segments.
__global__ __launch_bounds__(256) void seg_count_kernel(const uint32_t* keys, uint64_t N, uint32_t* bcnt) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint64_t n_blocks = (N + kSegBlock - 1) / kSegBlock;
  for (uint64_t b = wave; b < n_blocks; b += nwaves) {
    const uint64_t i = b * kSegBlock + lane;
    const bool head = i < N && (i == 0 || keys[i] != keys[i - 1]);
    const uint64_t heads = __ballot(head);
    if (lane == 0) bcnt[b] = (uint32_t)__popcll(heads);
  }
}

struct SegArgs {
  const uint32_t* keys;   // sorted, N
  const uint32_t* vals;   // entry ids in sorted order
  const uint64_t* boff;   // per block: segments starting before it (n_blocks + 1)
  uint64_t N;
  uint64_t n;             // tuples (entries [0, n) are syn0 entries)
  uint32_t K;
  uint32_t row_bits;
  const uint32_t* tuples;
  const float* dz;
  const float* stage;
  const void* xt;         // syn0 rows for the syn1 entries (float or BF16)
  uint32_t* seg_key;      // per segment: its key
  uint32_t* seg_first;    // per segment: sorted position of its first entry
  float* grad;            // per segment: the row gradient
  float* carry;           // per block: sum of its last run when that run continues
  float* headp;           // per block: sum of its first run when that run started earlier
  uint32_t* fin;          // per block: 1 if headp finishes a row
};

__device__ __forceinline__ uint32_t lane_u32(uint32_t v, uint32_t l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}
__device__ __forceinline__ float lane_f32(float v, uint32_t l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)l));
}

constexpr int kRowsInFlight = 8;

// Wave-uniform state of one block's walk over its sorted entries.
struct SegWalk {
  uint64_t b, base;
  uint32_t nv, pkey0, after;
  uint64_t ends;          // bit e: entry e ends a run of equal keys
  uint32_t piece_start;
  bool fin_here;
};

// Sum entries [lo, hi) of the block, all reading rows of element type T at
// `rows + src_e`; kRowsInFlight row loads are issued before they are added,
// in entry order.  A run end stores the run's sum: as a finished row
// gradient, as the block's carry (the run continues in the next block) or
// as the block's head part (the run started in an earlier block).
template <typename T, int DIM>
__device__ __forceinline__ void seg_run(const SegArgs& a, SegWalk& w, const T* rows, uint32_t lo,
                                        uint32_t hi, uint32_t lane, uint32_t key, uint32_t src,
                                        float scale, uint32_t ord, float (&acc)[DIM / 64]) {
  constexpr int D = DIM / 64;
  for (uint32_t e0 = lo; e0 < hi; e0 += kRowsInFlight) {
    float v[kRowsInFlight][D];
#pragma unroll
    for (int u = 0; u < kRowsInFlight; ++u) {
      if (e0 + u < hi) {
        const uint64_t se = (uint64_t)lane_u32(src, e0 + u) * DIM;
        RowIO<T, D>::load(rows + se + lane * D, v[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < kRowsInFlight; ++u) {
      const uint32_t e = e0 + u;
      if (e < hi) {
        const float ce = lane_f32(scale, e);
#pragma unroll
        for (int q = 0; q < D; ++q) acc[q] += ce * v[u][q];
        if ((w.ends >> e) & 1ull) {
          const uint32_t ke = lane_u32(key, e);
          const bool continued = w.piece_start == 0 && w.base > 0 && w.pkey0 == ke;
          const bool open = e + 1 == w.nv && w.after == ke;
          if (!continued && !open) {
            store_row<D>(a.grad + (uint64_t)lane_u32(ord, e) * DIM + lane * D, acc);
          } else if (open) {
            store_row<D>(a.carry + w.b * DIM + lane * D, acc);
          } else {
            store_row<D>(a.headp + w.b * DIM + lane * D, acc);
            w.fin_here = true;
          }
#pragma unroll
          for (int q = 0; q < D; ++q) acc[q] = 0.0f;
          w.piece_start = e + 1;
        }
      }
    }
  }
}

// Entry e of the sorted list contributes
//   syn0 entry (tuple t):        stage[t]               (sum_j dz_tj y_tj)
//   syn1 entry (t, j):           dz_tj * x[context_t]
// The table bit is the key's top bit, so a block's syn0 entries come first.
template <typename E, int DIM>
__global__ __launch_bounds__(256) void tuple_seg_kernel(SegArgs a) {
  constexpr int D = DIM / 64;
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint64_t n_blocks = (a.N + kSegBlock - 1) / kSegBlock;
  const E* xt = static_cast<const E*>(a.xt);
  const uint32_t K1 = a.K + 1, W = a.K + 2;
  for (uint64_t b = wave; b < n_blocks; b += nwaves) {
    SegWalk w;
    w.b = b;
    w.base = b * kSegBlock;
    w.nv = (uint32_t)(a.N - w.base < (uint64_t)kSegBlock ? a.N - w.base : kSegBlock);
    const uint64_t i = w.base + lane;
    const bool valid = lane < w.nv;
    const uint32_t key = valid ? a.keys[i] : 0xffffffffu;
    const uint32_t val = valid ? a.vals[i] : 0u;
    const uint32_t nkey = (i + 1 < a.N) ? a.keys[i + 1] : 0xffffffffu;
    w.pkey0 = w.base > 0 ? a.keys[w.base - 1] : 0xffffffffu;
    w.after = w.base + w.nv < a.N ? a.keys[w.base + w.nv] : 0xffffffffu;
    const uint32_t up_key = __shfl_up(key, 1, 64);  // every lane takes part
    const uint32_t prev_key = lane == 0 ? w.pkey0 : up_key;
    const bool head = valid && (i == 0 || prev_key != key);
    const uint64_t heads = __ballot(head);
    const uint64_t lanes_lt = (1ull << lane) - 1ull;
    const uint64_t bo = a.boff[b];
    if (head) {
      const uint64_t o = bo + (uint64_t)__popcll(heads & lanes_lt);
      a.seg_key[o] = key;
      a.seg_first[o] = (uint32_t)i;
    }
    // ordinal of this entry's segment: heads up to and including this lane,
    // after the bo segments that started in earlier blocks
    const uint32_t ord = valid ? (uint32_t)(bo + (uint64_t)__popcll(heads & (lanes_lt | (1ull << lane))) - 1) : 0u;
    const bool is0 = (key >> a.row_bits) == 0u;
    uint32_t src = 0;  // row index into the stage (syn0 entries) or into syn0 (syn1 entries)
    float scale = 0.0f;
    if (valid) {
      if (is0) {
        src = val;
        scale = 1.0f;
      } else {
        const uint32_t t = val / K1;
        src = a.tuples[(uint64_t)t * W + 1];
        scale = a.dz[val];
      }
    }
    w.ends = __ballot(valid && (key != nkey || lane + 1 == w.nv));
    const uint32_t n0 = (uint32_t)__popcll(__ballot(valid && is0));
    w.piece_start = 0;
    w.fin_here = false;
    float acc[D];
#pragma unroll
    for (int q = 0; q < D; ++q) acc[q] = 0.0f;
    seg_run<float, DIM>(a, w, a.stage, 0, n0, lane, key, src, scale, ord, acc);
    seg_run<E, DIM>(a, w, xt, n0, w.nv, lane, key, src, scale, ord, acc);
    if (lane == 0) a.fin[b] = w.fin_here ? 1u : 0u;
  }
}

// A row whose entries cross block edges: the carries of blocks b0 .. b-1
// (b0 = the block of the row's first entry) and the head part of block b.
// Sixteen wavefronts sum fixed contiguous chunks of the carries, in order;
// the chunk sums and the head part are then added in chunk order.
constexpr int kFixWaves = 16;

template <int DIM>
__global__ __launch_bounds__(64 * kFixWaves) void tuple_fix_kernel(SegArgs a) {
  constexpr int D = DIM / 64;
  __shared__ float part[kFixWaves][DIM];
  const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  const uint64_t n_blocks = (a.N + kSegBlock - 1) / kSegBlock;
  for (uint64_t b = blockIdx.x; b < n_blocks; b += gridDim.x) {
    if (!a.fin[b]) continue;  // uniform over the workgroup
    // block b's first entry continues a segment that started earlier: the
    // last one counted before block b
    const uint64_t ord = a.boff[b] - 1;
    const uint64_t b0 = a.seg_first[ord] / kSegBlock;
    const uint64_t cnt = b - b0;
    const uint64_t lo = b0 + cnt * wv / kFixWaves, hi = b0 + cnt * (wv + 1) / kFixWaves;
    float acc[D];
#pragma unroll
    for (int q = 0; q < D; ++q) acc[q] = 0.0f;
    for (uint64_t c0 = lo; c0 < hi; c0 += kRowsInFlight) {
      float v[kRowsInFlight][D];
#pragma unroll
      for (int u = 0; u < kRowsInFlight; ++u)
        if (c0 + u < hi) RowIO<float, D>::load(a.carry + (c0 + u) * DIM + lane * D, v[u]);
#pragma unroll
      for (int u = 0; u < kRowsInFlight; ++u)
        if (c0 + u < hi) {
#pragma unroll
          for (int q = 0; q < D; ++q) acc[q] += v[u][q];
        }
    }
#pragma unroll
    for (int q = 0; q < D; ++q) part[wv][lane * D + q] = acc[q];
    __syncthreads();
    if (wv == 0) {
      float tot[D];
      RowIO<float, D>::load(a.headp + b * DIM + lane * D, tot);
      float s[D];
#pragma unroll
      for (int q = 0; q < D; ++q) s[q] = 0.0f;
      for (int k = 0; k < kFixWaves; ++k) {
#pragma unroll
        for (int q = 0; q < D; ++q) s[q] += part[k][lane * D + q];
      }
#pragma unroll
      for (int q = 0; q < D; ++q) s[q] += tot[q];
      store_row<D>(a.grad + ord * DIM + lane * D, s);
    }
    __syncthreads();
  }
}

// W -= lr * grad for every segment; the BF16 shadow rows follow the new
// FP32 values.  One wavefront per segment.
template <int DIM>
__global__ __launch_bounds__(256) void tuple_apply_kernel(const uint32_t* seg_key, const float* grad,
                                                          const uint64_t* n_seg_p, uint32_t row_bits,
                                                          float lr, float* syn0, float* syn1,
                                                          uint16_t* bf0, uint16_t* bf1) {
  constexpr int D = DIM / 64;
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint64_t n_seg = *n_seg_p;
  for (uint64_t sg = wave; sg < n_seg; sg += nwaves) {
    const uint32_t key = seg_key[sg];
    const bool t1 = (key >> row_bits) != 0u;
    const uint64_t row = key & ((1u << row_bits) - 1u);
    float* w = (t1 ? syn1 : syn0) + row * DIM + lane * D;
    float g[D], cur[D];
    RowIO<float, D>::load(grad + sg * DIM + lane * D, g);
    RowIO<float, D>::load(w, cur);
#pragma unroll
    for (int q = 0; q < D; ++q) cur[q] = cur[q] - lr * g[q];
    store_row<D>(w, cur);
    uint16_t* sh = t1 ? bf1 : bf0;
    if (sh) {
#pragma unroll
      for (int q = 0; q < D; ++q) sh[row * DIM + lane * D + q] = to_bf16(cur[q]);
    }
  }
}

// Gradient rows by (table, row): binary search over the sorted segment keys.
__global__ void grad_rows_kernel(const uint32_t* seg_key, const float* grad, const uint64_t* n_seg_p,
                                 uint32_t row_bits, uint32_t table, const uint32_t* rows,
                                 uint64_t n_rows, uint32_t dim, float* out) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint64_t n_seg = *n_seg_p;
  for (uint64_t r = wave; r < n_rows; r += nwaves) {
    const uint32_t want = (table << row_bits) | rows[r];
    uint64_t lo = 0, hi = n_seg;
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo) / 2;
      if (seg_key[mid] < want)
        lo = mid + 1;
      else
        hi = mid;
    }
    const bool hit = lo < n_seg && seg_key[lo] == want;
    for (uint32_t d = lane; d < dim; d += 64)
      out[r * dim + d] = hit ? grad[lo * dim + d] : 0.0f;
  }
}

__global__ void to_bf16_kernel(const float* src, uint64_t n, uint16_t* dst) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = to_bf16(src[i]);
}

int waves_grid(dg_ctx* ctx, uint64_t waves) {
  uint64_t g = (waves + 3) / 4;  // 4 wavefronts per 256-thread block
  const uint64_t cap = (uint64_t)ctx->num_cus * 32;
  if (g > cap) g = cap;
  return g == 0 ? 1 : (int)g;
}

int row_bits_for(uint32_t rows) {
  int b = 1;
  while ((1ull << b) < rows) ++b;
  return b;
}

}  // namespace

struct dg_tuple_state {
  uint32_t* d_tuples = nullptr;
  size_t tuples_cap = 0;
  float* d_logits = nullptr;
  size_t logits_cap = 0;
  float* d_tloss = nullptr;
  size_t tloss_cap = 0;
  double* d_loss = nullptr;
  float* d_dl = nullptr;  // upstream dL/dz
  size_t dl_cap = 0;
  float* d_dz = nullptr;
  size_t dz_cap = 0;
  float* d_stage = nullptr;
  size_t stage_cap = 0;
  uint32_t* d_k[2] = {nullptr, nullptr};
  uint32_t* d_v[2] = {nullptr, nullptr};
  size_t k_cap[2] = {0, 0};
  size_t v_cap[2] = {0, 0};
  uint32_t* d_hist = nullptr;
  size_t hist_cap = 0;
  uint32_t* d_hoff = nullptr;   // per (digit, tile): offset within the digit
  size_t hoff_cap = 0;
  uint32_t* d_dtot = nullptr;   // per digit: total count
  size_t dtot_cap = 0;
  uint32_t* d_head = nullptr;
  size_t head_cap = 0;
  uint64_t* d_off = nullptr;
  size_t off_cap = 0;
  uint32_t* d_seg_key = nullptr;
  size_t seg_key_cap = 0;
  uint32_t* d_seg_first = nullptr;
  size_t seg_first_cap = 0;
  double* d_loss_parts = nullptr;
  float* d_grad = nullptr;
  size_t grad_cap = 0;
  float* d_carry = nullptr;
  size_t carry_cap = 0;
  float* d_headp = nullptr;
  size_t headp_cap = 0;
  uint32_t* d_fin = nullptr;
  size_t fin_cap = 0;
  // BF16 copies of the tables and the table version they were made from
  uint16_t* d_bf0 = nullptr;
  uint16_t* d_bf1 = nullptr;
  size_t bf_cap = 0;
  uint64_t bf_version = 0;
  // gradients of the last backward
  bool have_grads = false;
  uint32_t grad_rb = 0;
  uint32_t grad_dim = 0;
  uint32_t grad_rows = 0;
  uint64_t grad_N = 0;
  uint64_t grad_blocks = 0;  // segment count is d_off[grad_blocks]
  int sorted = 0;  // which of d_k / d_v holds the sorted pairs
};

void dg_tuples_free(dg_ctx* ctx) {
  dg_tuple_state* t = ctx->tuples;
  if (!t) return;
  void* bufs[] = {t->d_tuples, t->d_logits, t->d_tloss, t->d_loss,  t->d_dl,   t->d_dz,
                  t->d_stage,  t->d_k[0],   t->d_k[1],  t->d_v[0],  t->d_v[1], t->d_hist,
                  t->d_hoff,   t->d_dtot,   t->d_head,   t->d_off,   t->d_seg_key, t->d_seg_first, t->d_loss_parts, t->d_grad, t->d_carry,
                  t->d_headp,  t->d_fin,    t->d_bf0,   t->d_bf1};
  for (void* b : bufs)
    if (b) hipFree(b);
  delete t;
  ctx->tuples = nullptr;
}

void dg_tuples_invalidate(dg_ctx* ctx) {
  if (ctx->tuples) ctx->tuples->have_grads = false;
}

namespace {

dg_tuple_state* tuple_state(dg_ctx* ctx) {
  if (!ctx->tuples) ctx->tuples = new dg_tuple_state();
  return ctx->tuples;
}

// Validate a tuple batch on the host (every id is dereferenced on the device)
// and copy it to the device.
int upload_tuples(dg_ctx* ctx, dg_tuple_state* st, const dg_table_view& tv, const uint32_t* tuples,
                  uint64_t n, uint32_t K) {
  if (K > (uint32_t)kMaxTupleNeg)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "K = %u negatives per tuple: at most %d", K, kMaxTupleNeg);
  if (tv.dim != 64 && tv.dim != 128 && tv.dim != 256)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "tuple interface: dim %u (64, 128 or 256)", tv.dim);
  const uint64_t W = K + 2;
  if (n == 0 || n * W >= (1ull << 31))
    return dg_fail(ctx, DG_ERR_ARG, "tuple count %llu: must be in [1, 2^31 / (K + 2))",
                   (unsigned long long)n);
  for (uint64_t i = 0; i < n * W; ++i)
    if (tuples[i] >= tv.rows)
      return dg_fail(ctx, DG_ERR_ARG, "tuple %llu: row %u out of range (%u rows)",
                     (unsigned long long)(i / W), tuples[i], tv.rows);
  DG_CHECK(dg_reserve(ctx, &st->d_tuples, &st->tuples_cap, (size_t)(n * W)));
  DG_HIP(ctx, hipMemcpyAsync(st->d_tuples, tuples, sizeof(uint32_t) * n * W, hipMemcpyHostToDevice,
                             ctx->stream));
  return DG_OK;
}

// BF16 copies of both tables, remade when the FP32 tables changed.
int ensure_bf16(dg_ctx* ctx, dg_tuple_state* st, const dg_table_view& tv) {
  const size_t n = (size_t)tv.rows * tv.dim;
  if (st->d_bf0 && st->bf_cap >= n && st->bf_version == tv.version) return DG_OK;
  if (st->bf_cap < n) {
    if (st->d_bf0) hipFree(st->d_bf0);
    if (st->d_bf1) hipFree(st->d_bf1);
    st->d_bf0 = st->d_bf1 = nullptr;
    st->bf_cap = 0;
    DG_HIP(ctx, hipMalloc((void**)&st->d_bf0, sizeof(uint16_t) * n));
    DG_HIP(ctx, hipMalloc((void**)&st->d_bf1, sizeof(uint16_t) * n));
    st->bf_cap = n;
  }
  const int g = waves_grid(ctx, n / 64 + 1);
  to_bf16_kernel<<<g, 256, 0, ctx->stream>>>(tv.syn0, n, st->d_bf0);
  to_bf16_kernel<<<g, 256, 0, ctx->stream>>>(tv.syn1, n, st->d_bf1);
  DG_HIP(ctx, hipGetLastError());
  st->bf_version = tv.version;
  return DG_OK;
}

// Lanes per tuple: 8 for BF16 rows with K + 2 <= 8 ids per tuple (a BF16
// row is half an FP32 row, so twice the tuples fit the same registers), 16
// for K + 2 <= 16, otherwise 32.
template <typename E, int DIM, bool BWD>
void launch_fwd(dg_ctx* ctx, const TupleArgs& a) {
  if (sizeof(E) == 2 && DIM >= 128 && a.K + 2 <= 8) {
    const int g = waves_grid(ctx, (a.n + 7) / 8);  // a tuple per 8 lanes
    tuple_fwd_kernel<E, DIM, 8, BWD, 8><<<g, 256, 0, ctx->stream>>>(a);
    return;
  }
  if (a.K + 2 <= 16) {
    const int g = waves_grid(ctx, (a.n + 3) / 4);  // a tuple per 16 lanes
    if (a.K + 1 <= 8)
      tuple_fwd_kernel<E, DIM, 8, BWD, 16><<<g, 256, 0, ctx->stream>>>(a);
    else
      tuple_fwd_kernel<E, DIM, 16, BWD, 16><<<g, 256, 0, ctx->stream>>>(a);
    return;
  }
  const int g = waves_grid(ctx, (a.n + 1) / 2);  // a tuple per half-wavefront
  if (a.K + 1 <= 8)
    tuple_fwd_kernel<E, DIM, 8, BWD, 32><<<g, 256, 0, ctx->stream>>>(a);
  else
    tuple_fwd_kernel<E, DIM, 16, BWD, 32><<<g, 256, 0, ctx->stream>>>(a);
}

template <bool BWD>
void launch_fwd_form(dg_ctx* ctx, const TupleArgs& a, int bf16, uint32_t dim) {
  if (bf16) {
    if (dim == 64) launch_fwd<uint16_t, 64, BWD>(ctx, a);
    else if (dim == 128) launch_fwd<uint16_t, 128, BWD>(ctx, a);
    else launch_fwd<uint16_t, 256, BWD>(ctx, a);
  } else {
    if (dim == 64) launch_fwd<float, 64, BWD>(ctx, a);
    else if (dim == 128) launch_fwd<float, 128, BWD>(ctx, a);
    else launch_fwd<float, 256, BWD>(ctx, a);
  }
}

template <typename E>
void launch_seg(dg_ctx* ctx, const SegArgs& a, uint32_t dim, int g, int gf) {
  if (dim == 64) {
    tuple_seg_kernel<E, 64><<<g, 256, 0, ctx->stream>>>(a);
    tuple_fix_kernel<64><<<gf, 64 * kFixWaves, 0, ctx->stream>>>(a);
  } else if (dim == 128) {
    tuple_seg_kernel<E, 128><<<g, 256, 0, ctx->stream>>>(a);
    tuple_fix_kernel<128><<<gf, 64 * kFixWaves, 0, ctx->stream>>>(a);
  } else {
    tuple_seg_kernel<E, 256><<<g, 256, 0, ctx->stream>>>(a);
    tuple_fix_kernel<256><<<gf, 64 * kFixWaves, 0, ctx->stream>>>(a);
  }
}

}  // namespace

extern "C" int dg_sgns_tuples_forward(dg_ctx* ctx, const uint32_t* tuples, uint64_t n, uint32_t K,
                                      int bf16, float* logits, double* loss) {
  if (!ctx || !tuples) return DG_ERR_ARG;
  dg_table_view tv;
  DG_CHECK(dg_sgns_view(ctx, &tv));
  DG_HIP(ctx, hipSetDevice(ctx->device));
  dg_tuple_state* st = tuple_state(ctx);
  DG_CHECK(upload_tuples(ctx, st, tv, tuples, n, K));
  if (bf16) DG_CHECK(ensure_bf16(ctx, st, tv));
  const uint32_t K1 = K + 1;
  DG_CHECK(dg_reserve(ctx, &st->d_logits, &st->logits_cap, (size_t)(n * K1)));
  DG_CHECK(dg_reserve(ctx, &st->d_tloss, &st->tloss_cap, (size_t)n));
  if (!st->d_loss) DG_HIP(ctx, hipMalloc((void**)&st->d_loss, sizeof(double)));
  if (!st->d_loss_parts) DG_HIP(ctx, hipMalloc((void**)&st->d_loss_parts, sizeof(double) * kLossParts));
  TupleArgs a{};
  a.tuples = st->d_tuples;
  a.n = n;
  a.K = K;
  a.xt = bf16 ? (const void*)st->d_bf0 : (const void*)tv.syn0;
  a.yt = bf16 ? (const void*)st->d_bf1 : (const void*)tv.syn1;
  a.logits = st->d_logits;
  a.tloss = st->d_tloss;
  size_t tk;
  DG_CHECK(dg_time_begin(ctx, "tuple_fwd", &tk));
  launch_fwd_form<false>(ctx, a, bf16, tv.dim);
  loss_part_kernel<<<kLossParts, 256, 0, ctx->stream>>>(st->d_tloss, n, st->d_loss_parts);
  loss_final_kernel<<<1, 256, 0, ctx->stream>>>(st->d_loss_parts, st->d_loss);
  DG_HIP(ctx, hipGetLastError());
  DG_CHECK(dg_time_end(ctx, tk));
  if (logits)
    DG_HIP(ctx, hipMemcpyAsync(logits, st->d_logits, sizeof(float) * n * K1, hipMemcpyDeviceToHost,
                               ctx->stream));
  double h_loss = 0.0;
  DG_HIP(ctx, hipMemcpyAsync(&h_loss, st->d_loss, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  DG_CHECK(dg_time_resolve(ctx));
  if (loss) *loss = h_loss;
  return DG_OK;
}

namespace {

// ---- the allocation-free forward ---------------------------------------------
// Scratch of dg_sgns_tuples_forward_noalloc: the loss's partial sums
// (kLossParts doubles), a range-check flag, then one loss term per tuple.
// Each device holds its own copy; host code reaches it only through the
// address hipGetSymbolAddress gives on that device (dg_tuples_scratch_setup).
constexpr uint64_t kNoAllocMaxTuples = 1ull << 20;
constexpr int kNoAllocParts = 2 * kLossParts;    // floats holding the partial sums
constexpr int kNoAllocFlag = kNoAllocParts;      // float index of the flag
constexpr int kNoAllocTerms = kNoAllocParts + 64;  // first loss term (256-byte aligned)
static __device__ __align__(256) float scratch[kNoAllocTerms + kNoAllocMaxTuples];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_dev[kMaxScratchDevices] = {};

__global__ void noalloc_clear_kernel(uint32_t* flag) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *flag = 0u;
}

__global__ void tuple_check_kernel(const uint32_t* tuples, uint64_t count, uint32_t rows, uint32_t* flag) {
  bool bad = false;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * blockDim.x)
    bad |= tuples[i] >= rows;
  if (__ballot(bad) != 0 && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

}  // namespace

int dg_tuples_scratch_setup(dg_ctx* ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxScratchDevices)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "device %d: the allocation-free scratch covers devices "
                   "0..%d", ctx->device, kMaxScratchDevices - 1);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[ctx->device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return dg_fail(ctx, DG_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_dev[ctx->device] = static_cast<float*>(p);
  }
  ctx->noalloc_scratch = g_scratch_dev[ctx->device];
  return DG_OK;
}

extern "C" int dg_sgns_tuples_forward_noalloc(dg_ctx* ctx, const uint32_t* d_tuples, uint64_t n,
                                              uint32_t K, float* d_logits, double* d_loss) {
  if (!ctx || !d_tuples || !d_logits || !d_loss) return DG_ERR_ARG;
  if (!ctx->noalloc_scratch) return dg_fail(ctx, DG_ERR_STATE, "no allocation-free scratch on this context");
  dg_table_view tv;
  DG_CHECK(dg_sgns_view(ctx, &tv));
  if (K > (uint32_t)kMaxTupleNeg)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "K = %u negatives per tuple: at most %d", K, kMaxTupleNeg);
  if (tv.dim != 64 && tv.dim != 128 && tv.dim != 256)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "tuple interface: dim %u (64, 128 or 256)", tv.dim);
  if (n == 0 || n > kNoAllocMaxTuples)
    return dg_fail(ctx, DG_ERR_ARG, "%llu tuples: the allocation-free forward takes 1 to %llu",
                   (unsigned long long)n, (unsigned long long)kNoAllocMaxTuples);
  DG_HIP(ctx, hipSetDevice(ctx->device));
  float* base = ctx->noalloc_scratch;
  double* parts = reinterpret_cast<double*>(base);
  uint32_t* flag = reinterpret_cast<uint32_t*>(base + kNoAllocFlag);
  float* terms = base + kNoAllocTerms;
  // every row id is dereferenced by the forward: check them on the device
  const uint64_t count = n * (K + 2);
  noalloc_clear_kernel<<<1, 64, 0, ctx->stream>>>(flag);
  tuple_check_kernel<<<waves_grid(ctx, count / 64 + 1), 256, 0, ctx->stream>>>(d_tuples, count, tv.rows,
                                                                               flag);
  DG_HIP(ctx, hipGetLastError());
  uint32_t bad = 0;
  DG_HIP(ctx, hipMemcpyAsync(&bad, flag, sizeof(bad), hipMemcpyDeviceToHost, ctx->stream));
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (bad) return dg_fail(ctx, DG_ERR_ARG, "a tuple row is out of range (%u rows)", tv.rows);
  TupleArgs a{};
  a.tuples = d_tuples;
  a.n = n;
  a.K = K;
  a.xt = tv.syn0;
  a.yt = tv.syn1;
  a.logits = d_logits;
  a.tloss = terms;
  launch_fwd_form<false>(ctx, a, 0, tv.dim);
  loss_part_kernel<<<kLossParts, 256, 0, ctx->stream>>>(terms, n, parts);
  loss_final_kernel<<<1, 256, 0, ctx->stream>>>(parts, d_loss);
  DG_HIP(ctx, hipGetLastError());
  return DG_OK;
}

extern "C" int dg_sgns_tuples_backward(dg_ctx* ctx, const uint32_t* tuples, uint64_t n, uint32_t K,
                                       int bf16, const float* dlogits, float lr, int apply) {
  if (!ctx || !tuples) return DG_ERR_ARG;
  dg_table_view tv;
  DG_CHECK(dg_sgns_view(ctx, &tv));
  DG_HIP(ctx, hipSetDevice(ctx->device));
  dg_tuple_state* st = tuple_state(ctx);
  st->have_grads = false;
  DG_CHECK(upload_tuples(ctx, st, tv, tuples, n, K));
  if (bf16) DG_CHECK(ensure_bf16(ctx, st, tv));
  const uint32_t K1 = K + 1, dim = tv.dim;
  const uint64_t N = n * (K + 2);
  const int rb = row_bits_for(tv.rows);
  DG_CHECK(dg_reserve(ctx, &st->d_dz, &st->dz_cap, (size_t)(n * K1)));
  DG_CHECK(dg_reserve(ctx, &st->d_stage, &st->stage_cap, (size_t)(n * dim)));
  for (int b = 0; b < 2; ++b) {
    DG_CHECK(dg_reserve(ctx, &st->d_k[b], &st->k_cap[b], (size_t)N));
    DG_CHECK(dg_reserve(ctx, &st->d_v[b], &st->v_cap[b], (size_t)N));
  }
  if (dlogits) {
    DG_CHECK(dg_reserve(ctx, &st->d_dl, &st->dl_cap, (size_t)(n * K1)));
    DG_HIP(ctx, hipMemcpyAsync(st->d_dl, dlogits, sizeof(float) * n * K1, hipMemcpyHostToDevice,
                               ctx->stream));
  }
  TupleArgs a{};
  a.tuples = st->d_tuples;
  a.n = n;
  a.K = K;
  a.xt = bf16 ? (const void*)st->d_bf0 : (const void*)tv.syn0;
  a.yt = bf16 ? (const void*)st->d_bf1 : (const void*)tv.syn1;
  a.dlogits = dlogits ? st->d_dl : nullptr;
  a.dz = st->d_dz;
  a.stage = st->d_stage;
  a.keys = st->d_k[0];
  a.vals = st->d_v[0];
  a.row_bits = (uint32_t)rb;
  size_t tk;
  DG_CHECK(dg_time_begin(ctx, "tuple_prep", &tk));
  launch_fwd_form<true>(ctx, a, bf16, dim);
  DG_HIP(ctx, hipGetLastError());
  DG_CHECK(dg_time_end(ctx, tk));
  // sort (key, entry) pairs on rb + 1 key bits
  const uint32_t n_tiles = (uint32_t)((N + kSortTile - 1) / kSortTile);
  DG_CHECK(dg_reserve(ctx, &st->d_hist, &st->hist_cap, (size_t)n_tiles * 256));
  DG_CHECK(dg_reserve(ctx, &st->d_hoff, &st->hoff_cap, (size_t)n_tiles * 256));
  DG_CHECK(dg_reserve(ctx, &st->d_dtot, &st->dtot_cap, (size_t)256));
  DG_CHECK(dg_reserve(ctx, &st->d_head, &st->head_cap, (size_t)(N / kSegBlock + 1)));
  DG_CHECK(dg_reserve(ctx, &st->d_off, &st->off_cap, (size_t)(N / kSegBlock + 2)));
  DG_CHECK(dg_time_begin(ctx, "tuple_sort", &tk));
  int cur = 0;
  for (int shift = 0; shift < rb + 1; shift += 8) {
    radix_hist_kernel<<<n_tiles, kSortBlock, 0, ctx->stream>>>(st->d_k[cur], N, shift, n_tiles,
                                                               st->d_hist);
    radix_digit_scan_kernel<<<256, 256, 0, ctx->stream>>>(st->d_hist, n_tiles, st->d_hoff, st->d_dtot);
    radix_scatter_kernel<<<n_tiles, kSortBlock, 0, ctx->stream>>>(
        st->d_k[cur], st->d_v[cur], N, shift, n_tiles, st->d_hoff, st->d_dtot, st->d_k[1 - cur],
        st->d_v[1 - cur]);
    DG_HIP(ctx, hipGetLastError());
    cur = 1 - cur;
  }
  const uint64_t n_blocks = (N + kSegBlock - 1) / kSegBlock;
  seg_count_kernel<<<waves_grid(ctx, n_blocks), 256, 0, ctx->stream>>>(st->d_k[cur], N, st->d_head);
  DG_CHECK(dg_scan_lengths(ctx, st->d_head, n_blocks, st->d_off));
  DG_HIP(ctx, hipGetLastError());
  DG_CHECK(dg_time_end(ctx, tk));
  // segmented sums
  DG_CHECK(dg_reserve(ctx, &st->d_seg_key, &st->seg_key_cap, (size_t)N));
  DG_CHECK(dg_reserve(ctx, &st->d_seg_first, &st->seg_first_cap, (size_t)N));
  DG_CHECK(dg_reserve(ctx, &st->d_grad, &st->grad_cap, (size_t)(N * dim)));
  DG_CHECK(dg_reserve(ctx, &st->d_carry, &st->carry_cap, (size_t)(n_blocks * dim)));
  DG_CHECK(dg_reserve(ctx, &st->d_headp, &st->headp_cap, (size_t)(n_blocks * dim)));
  DG_CHECK(dg_reserve(ctx, &st->d_fin, &st->fin_cap, (size_t)n_blocks));
  SegArgs sa{};
  sa.keys = st->d_k[cur];
  sa.vals = st->d_v[cur];
  sa.boff = st->d_off;
  sa.N = N;
  sa.n = n;
  sa.K = K;
  sa.row_bits = (uint32_t)rb;
  sa.tuples = st->d_tuples;
  sa.dz = st->d_dz;
  sa.stage = st->d_stage;
  sa.xt = bf16 ? (const void*)st->d_bf0 : (const void*)tv.syn0;
  sa.seg_key = st->d_seg_key;
  sa.seg_first = st->d_seg_first;
  sa.grad = st->d_grad;
  sa.carry = st->d_carry;
  sa.headp = st->d_headp;
  sa.fin = st->d_fin;
  const int g = waves_grid(ctx, n_blocks);
  const int gf = (int)std::min<uint64_t>(n_blocks, (uint64_t)ctx->num_cus * 4);
  DG_CHECK(dg_time_begin(ctx, "tuple_seg", &tk));
  if (bf16)
    launch_seg<uint16_t>(ctx, sa, dim, g, gf);
  else
    launch_seg<float>(ctx, sa, dim, g, gf);
  DG_HIP(ctx, hipGetLastError());
  DG_CHECK(dg_time_end(ctx, tk));
  st->sorted = cur;
  st->grad_rb = (uint32_t)rb;
  st->grad_dim = dim;
  st->grad_rows = tv.rows;
  st->grad_N = N;
  st->grad_blocks = n_blocks;
  if (apply) {
    DG_CHECK(dg_time_begin(ctx, "tuple_apply", &tk));
    const bool shadow = st->d_bf0 && st->bf_version == tv.version;
    uint16_t* b0 = shadow ? st->d_bf0 : nullptr;
    uint16_t* b1 = shadow ? st->d_bf1 : nullptr;
    const int ga = waves_grid(ctx, N);
    if (dim == 64)
      tuple_apply_kernel<64><<<ga, 256, 0, ctx->stream>>>(st->d_seg_key, st->d_grad, st->d_off + n_blocks, rb,
                                                          lr, tv.syn0, tv.syn1, b0, b1);
    else if (dim == 128)
      tuple_apply_kernel<128><<<ga, 256, 0, ctx->stream>>>(st->d_seg_key, st->d_grad, st->d_off + n_blocks, rb,
                                                           lr, tv.syn0, tv.syn1, b0, b1);
    else
      tuple_apply_kernel<256><<<ga, 256, 0, ctx->stream>>>(st->d_seg_key, st->d_grad, st->d_off + n_blocks, rb,
                                                           lr, tv.syn0, tv.syn1, b0, b1);
    DG_HIP(ctx, hipGetLastError());
    DG_CHECK(dg_time_end(ctx, tk));
    const uint64_t v = dg_sgns_written(ctx);
    if (shadow) st->bf_version = v;  // the shadow rows were updated with the FP32 rows
  }
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  DG_CHECK(dg_time_resolve(ctx));
  st->have_grads = true;
  return DG_OK;
}

extern "C" int dg_sgns_grad_rows(dg_ctx* ctx, int table, const uint32_t* rows, uint64_t n_rows,
                                 float* out) {
  if (!ctx || (n_rows && (!rows || !out)) || (table != 0 && table != 1)) return DG_ERR_ARG;
  dg_tuple_state* st = ctx->tuples;
  if (!st || !st->have_grads) return dg_fail(ctx, DG_ERR_STATE, "no gradients: run dg_sgns_tuples_backward");
  if (n_rows == 0) return DG_OK;
  for (uint64_t i = 0; i < n_rows; ++i)
    if (rows[i] >= st->grad_rows)
      return dg_fail(ctx, DG_ERR_ARG, "row %u out of range (%u rows)", rows[i], st->grad_rows);
  DG_HIP(ctx, hipSetDevice(ctx->device));
  uint32_t* d_rows = nullptr;
  float* d_out = nullptr;
  DG_HIP(ctx, hipMalloc((void**)&d_rows, sizeof(uint32_t) * n_rows));
  DG_HIP(ctx, hipMalloc((void**)&d_out, sizeof(float) * n_rows * st->grad_dim));
  int rc = DG_OK;
  hipError_t e = hipMemcpy(d_rows, rows, sizeof(uint32_t) * n_rows, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    grad_rows_kernel<<<waves_grid(ctx, n_rows), 256, 0, ctx->stream>>>(
        st->d_seg_key, st->d_grad, st->d_off + st->grad_blocks, st->grad_rb, (uint32_t)table, d_rows, n_rows,
        st->grad_dim, d_out);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpy(out, d_out, sizeof(float) * n_rows * st->grad_dim, hipMemcpyDeviceToHost);
  if (e != hipSuccess) rc = dg_fail(ctx, DG_ERR_HIP, "grad_rows: %s", hipGetErrorString(e));
  hipFree(d_rows);
  hipFree(d_out);
  return rc;
}
