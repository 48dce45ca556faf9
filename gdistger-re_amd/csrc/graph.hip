// graph.hip - device-resident CSR for the walker, common-neighbour counts and
// per-entry acceptance thresholds.
//
// Replaces GraphEngine::load_graph (include/graph.hpp:642-824) for a single
// partition: the whole graph lives in one GPU's HBM, so the MPI edge shuffle
// and vertex-range partitioning (:542-640, :723-773) disappear.  Neighbour
// order is part of the walk result (the walker indexes adj.begin + i,
// include/walk.hpp:899), so the CSR keeps the reference order: the doubled
// edge list [e0, rev e0, e1, rev e1, ...] (:671-686), stable by source
// (build_edge_container, :509-540).
//
// HBM layout
//   d_row_ptr : uint64[V+1]  row offsets (edge_id_t)
//   d_nbr     : uint32[m]    neighbour ids, reference order
//   d_cn      : uint32[m]    common-neighbour count |N(u) & N(v) \ {u,v}|
//   d_adj     : uint2[m]     walker view {neighbour, acceptance threshold}:
//                            8 bytes per entry like the reference
//                            AdjUnit<real_t> (include/graph.hpp:32-37)
//   d_deg     : uint32[V]    degree (= row length)
//
// Common neighbours (src/tools/gcn.cpp:250-373, the offline tool whose
// output the reference walker reads back as edge data, walk.hpp:920) are
// computed on the device over the distinct neighbours: one wavefront per
// source vertex hashes N(u) into LDS, then probes the neighbour lists 64
// entries at a time and counts hits with a ballot; a source with more than
// kSmallDeg neighbours gets a hash table in global memory and its entries are
// shared out to many wavefronts.  On a simple undirected graph each pair is
// counted once, from its larger-degree endpoint scanning the smaller list.
//
// Acceptance threshold: the walker bounces back when tanh(pa) <= p with
// p = u / (2^31 - 1), u the 31-bit draw (walk.hpp:928-938).  pa depends only
// on (deg u, deg v, cn), so per entry the smallest u that satisfies the test
// is found once by bisection (u / M is monotone in u under correctly rounded
// division), and the walker tests `u >= threshold` - the same decision bit
// for bit, without tanh or the neighbour's degree in the step loop.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>

#include "dg_internal.h"
#include "dg_math.h"

namespace {

constexpr uint32_t kEmpty = 0xffffffffu;
constexpr int kSmallDeg = 1024;          // wave path: hash of 2048 slots per wave
constexpr int kSmallSlots = 2048;

// Multiplicative (Fibonacci) hashing keeps the high bits of the product:
// the low bits depend only on the low bits of w, and vertex ids of generated
// graphs (R-MAT in particular) are far from uniform in their low bits.
__device__ __forceinline__ uint32_t slot_hash(uint32_t w, uint32_t mask) {
  return (w * 2654435761u) >> __builtin_clz(mask);  // mask = 2^k - 1, k >= 6
}

__device__ __forceinline__ void hash_insert(uint32_t* tbl, uint32_t mask, uint32_t w) {
  uint32_t s = slot_hash(w, mask);
  while (true) {
    uint32_t prev = atomicCAS(&tbl[s], kEmpty, w);
    if (prev == kEmpty || prev == w) return;
    s = (s + 1) & mask;
  }
}

__device__ __forceinline__ bool hash_find(const uint32_t* tbl, uint32_t mask, uint32_t w) {
  uint32_t s = slot_hash(w, mask);
  while (true) {
    uint32_t x = tbl[s];
    if (x == w) return true;
    if (x == kEmpty) return false;
    s = (s + 1) & mask;
  }
}

// A pair's count is |N(u) & N(v) \ {u, v}| over the distinct neighbours.
// With SYM (simple undirected graph, rev[k] = position of the reverse entry
// of k) each pair is counted once, from its larger-degree endpoint (ties by
// id): the smaller list is scanned against the hash of the larger one and
// the count written to both entries.  Cost: the sum over edges of
// min(deg u, deg v) instead of the sum of squared degrees.
template <bool SYM>
__device__ __forceinline__ bool other_side(const uint64_t* rp, uint32_t u, uint32_t du, uint32_t v) {
  if (!SYM || v == u) return false;
  const uint32_t dv = (uint32_t)(rp[v + 1] - rp[v]);
  return dv > du || (dv == du && v > u);
}

// One wavefront per source vertex with degree <= kSmallDeg.
template <bool SYM>
__global__ __launch_bounds__(256) void cn_small_kernel(const uint64_t* rp, const uint32_t* nbr,
                                                       uint32_t v_num, const uint64_t* rev,
                                                       uint32_t* cn) {
  __shared__ uint32_t s_tbl[4][kSmallSlots];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t* tbl = s_tbl[wave];
  const uint32_t waves = gridDim.x * 4;
  for (uint32_t u = blockIdx.x * 4 + wave; u < v_num; u += waves) {
    const uint64_t b = rp[u], e = rp[u + 1];
    const uint32_t du = (uint32_t)(e - b);
    if (du > (uint32_t)kSmallDeg || du == 0) continue;  // wave-uniform
    uint32_t slots = 64;
    while (slots < 2 * du) slots <<= 1;
    const uint32_t mask = slots - 1;
    for (uint32_t i = lane; i < slots; i += 64) tbl[i] = kEmpty;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    for (uint32_t i = lane; i < du; i += 64) {
      const uint32_t w = nbr[b + i];
      if (w != u) hash_insert(tbl, mask, w);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (uint64_t k = b; k < e; ++k) {
      const uint32_t v = nbr[k];
      if (other_side<SYM>(rp, u, du, v)) continue;  // wave-uniform
      const uint64_t vb = rp[v], ve = rp[v + 1];
      uint32_t c = 0;
      for (uint64_t j0 = vb; j0 < ve; j0 += 64) {  // wave-uniform trip count
        bool hit = false;
        if (j0 + lane < ve) {
          const uint32_t w = nbr[j0 + lane];
          hit = (w != v) && (w != u) && hash_find(tbl, mask, w);
        }
        c += (uint32_t)__popcll(__ballot(hit));
      }
      if (lane == 0) {
        cn[k] = c;
        if (SYM && v != u) cn[rev[k]] = c;
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// Sources with degree > kSmallDeg (hubs): N(u) goes into a hash table of its
// own in global memory (hub_hash_kernel, one workgroup per hub), and the
// hub's entries are cut into items of kHubItem consecutive entries, one
// wavefront per item, so a hub's work spreads over many wavefronts and every
// neighbour list is scanned once.
constexpr uint32_t kHubItem = 64;

__global__ __launch_bounds__(256) void hub_hash_kernel(const uint64_t* rp, const uint32_t* nbr,
                                                       const uint32_t* big, uint32_t n_big,
                                                       const uint64_t* toff, const uint32_t* tsize,
                                                       uint32_t* tab) {
  for (uint32_t bi = blockIdx.x; bi < n_big; bi += gridDim.x) {
    const uint32_t u = big[bi];
    uint32_t* t = tab + toff[bi];
    const uint32_t mask = tsize[bi] - 1;
    for (uint64_t i = rp[u] + threadIdx.x; i < rp[u + 1]; i += 256) {
      const uint32_t w = nbr[i];
      if (w != u) hash_insert(t, mask, w);
    }
  }
}

template <bool SYM>
__global__ __launch_bounds__(256) void hub_cn_kernel(const uint64_t* rp, const uint32_t* nbr,
                                                     const uint32_t* big, const uint64_t* toff,
                                                     const uint32_t* tsize, const uint32_t* tab,
                                                     const uint32_t* it_hub, const uint64_t* it_k0,
                                                     uint64_t n_items, const uint64_t* rev, uint32_t* cn) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t it = wave; it < n_items; it += nwaves) {
    const uint32_t bi = it_hub[it];
    const uint32_t u = big[bi];
    const uint64_t b = rp[u], e = rp[u + 1];
    const uint32_t du = (uint32_t)(e - b);
    const uint32_t* t = tab + toff[bi];
    const uint32_t mask = tsize[bi] - 1;
    const uint64_t k0 = it_k0[it];
    const uint64_t k1 = k0 + kHubItem < e ? k0 + kHubItem : e;
    for (uint64_t k = k0; k < k1; ++k) {
      const uint32_t v = nbr[k];
      if (other_side<SYM>(rp, u, du, v)) continue;  // wave-uniform
      const uint64_t vb = rp[v], ve = rp[v + 1];
      uint32_t c = 0;
      for (uint64_t j = vb; j < ve; j += 64) {
        bool hit = false;
        if (j + lane < ve) {
          const uint32_t w = nbr[j + lane];
          hit = (w != v) && (w != u) && hash_find(t, mask, w);
        }
        c += (uint32_t)__popcll(__ballot(hit));
      }
      if (lane == 0) {
        cn[k] = c;
        if (SYM && v != u) cn[rev[k]] = c;
      }
    }
  }
}

__global__ void fill_kernel(uint32_t* p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void degree_kernel(const uint64_t* rp, uint32_t v_num, uint32_t* deg) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < v_num; v += gridDim.x * blockDim.x)
    deg[v] = (uint32_t)(rp[v + 1] - rp[v]);
}

// Smallest u in [0, 2^31 - 1] with tanh(pa) <= u / (2^31 - 1), pa as in
// walk.hpp:933 (times the edge weight w on a weighted graph: tanh(pa * w));
// 2^31 (never reached by a 31-bit draw) if no u qualifies.
__device__ __forceinline__ uint32_t bounce_threshold(int sd, int dd, int cnv, const float* w,
                                                     uint64_t e) {
  const double M = 2147483647.0;
  const double pa = (1.0 / (double)(sd - cnv)) * (double)(sd > dd ? sd : dd) /
                    (double)(sd < dd ? sd : dd);
  const double pn = dg_tanh(w ? pa * (double)w[e] : pa);
  if (pn != pn) return 0x80000000u;
  if (!(pn <= 1.0)) return 0x80000000u;
  uint32_t lo = 0, hi = 2147483647u;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (pn <= (double)mid / M)
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

__global__ void threshold_kernel(const uint64_t* rp, const uint32_t* nbr, const uint32_t* cn,
                                 const float* w, const uint32_t* deg, uint32_t v_num, uint64_t m,
                                 uint2* adj) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m;
       e += (uint64_t)gridDim.x * blockDim.x) {
    // source vertex of entry e: last u with rp[u] <= e
    uint32_t lo = 0, hi = v_num - 1;
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo + 1) / 2;
      if (rp[mid] <= e)
        lo = mid;
      else
        hi = mid - 1;
    }
    const uint32_t v = nbr[e];
    adj[e] = make_uint2(v, bounce_threshold((int)deg[lo], (int)deg[v], (int)cn[e], w, e));
  }
}

int grid_for(dg_ctx* ctx, uint64_t n, int block) {
  uint64_t g = (n + block - 1) / block;
  uint64_t cap = (uint64_t)ctx->num_cus * 32;
  if (g > cap) g = cap;
  return g == 0 ? 1 : (int)g;
}

}  // namespace

void dg_graph_free(dg_ctx* ctx) {
  void* bufs[] = {ctx->d_row_ptr, ctx->d_nbr, ctx->d_cn, ctx->d_adj, ctx->d_deg, ctx->d_hot_slot,
                  ctx->d_hot_id};
  for (void* p : bufs)
    if (p) hipFree(p);
  ctx->d_hot_slot = nullptr;
  ctx->d_hot_id = nullptr;
  ctx->n_hot = 0;
  ctx->d_row_ptr = nullptr;
  ctx->d_nbr = nullptr;
  ctx->d_cn = nullptr;
  ctx->d_adj = nullptr;
  ctx->d_deg = nullptr;
  ctx->v_num = 0;
  ctx->m = 0;
}

// Device CSR from the host arrays; CN from the edge data or computed here.
// Rows in parallel on the host: lo .. hi of [0, n) per thread.
template <typename F>
static void host_rows(uint32_t n, F f) {
  unsigned t = std::thread::hardware_concurrency();
  t = t == 0 ? 1 : (t > 16 ? 16 : t);
  if (n < 4096) t = 1;
  std::vector<std::thread> th;
  for (unsigned i = 0; i < t; ++i) {
    const uint32_t lo = (uint32_t)((uint64_t)n * i / t), hi = (uint32_t)((uint64_t)n * (i + 1) / t);
    th.emplace_back([=] { f(lo, hi); });
  }
  for (auto& x : th) x.join();
}

// True if no row lists a neighbour twice or the vertex itself.
static bool rows_simple(uint32_t v_num, const std::vector<uint64_t>& rp, const std::vector<uint32_t>& nbr) {
  std::atomic<bool> simple{true};
  host_rows(v_num, [&](uint32_t lo, uint32_t hi) {
    std::vector<uint32_t> row;
    for (uint32_t u = lo; u < hi && simple.load(std::memory_order_relaxed); ++u) {
      row.assign(nbr.begin() + rp[u], nbr.begin() + rp[u + 1]);
      std::sort(row.begin(), row.end());
      for (size_t i = 0; i < row.size(); ++i)
        if (row[i] == u || (i && row[i] == row[i - 1])) {
          simple = false;
          break;
        }
    }
  });
  return simple;
}

// Common-neighbour counts on the device.  A simple undirected graph counts
// each pair once (SYM, with the reverse positions `rev`); a simple directed
// graph counts every entry; any other graph is counted on its distinct
// neighbour lists (sorted, without the vertex itself), and each entry takes
// the count of its pair - the set semantics of gcn.cpp:329-346.
static int compute_cn(dg_ctx* ctx, uint32_t v_num, const std::vector<uint64_t>& rp,
                      const std::vector<uint32_t>& nbr, const std::vector<uint64_t>& rev,
                      bool undirected) {
  const bool simple = rows_simple(v_num, rp, nbr);
  const std::vector<uint64_t>* crp = &rp;
  const std::vector<uint32_t>* cnbr = &nbr;
  std::vector<uint64_t> srp;
  std::vector<uint32_t> snbr;
  if (!simple) {  // distinct neighbour lists
    srp.assign(v_num + 1, 0);
    std::vector<std::vector<uint32_t>> rows(v_num);
    host_rows(v_num, [&](uint32_t lo, uint32_t hi) {
      for (uint32_t u = lo; u < hi; ++u) {
        std::vector<uint32_t>& r = rows[u];
        r.assign(nbr.begin() + rp[u], nbr.begin() + rp[u + 1]);
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end()), r.end());
        r.erase(std::remove(r.begin(), r.end(), u), r.end());
      }
    });
    for (uint32_t u = 0; u < v_num; ++u) srp[u + 1] = srp[u] + rows[u].size();
    snbr.resize(srp[v_num]);
    for (uint32_t u = 0; u < v_num; ++u) std::copy(rows[u].begin(), rows[u].end(), snbr.begin() + srp[u]);
    crp = &srp;
    cnbr = &snbr;
  }
  const uint64_t cm = (*crp)[v_num];
  const bool sym = simple && undirected;
  uint64_t *d_rp = ctx->d_row_ptr, *d_rev = nullptr;
  uint32_t *d_nbr = ctx->d_nbr, *d_cnv = ctx->d_cn;
  if (!simple) {
    DG_HIP(ctx, hipMalloc((void**)&d_rp, sizeof(uint64_t) * (v_num + 1)));
    DG_HIP(ctx, hipMalloc((void**)&d_nbr, sizeof(uint32_t) * (cm ? cm : 1)));
    DG_HIP(ctx, hipMalloc((void**)&d_cnv, sizeof(uint32_t) * (cm ? cm : 1)));
    DG_HIP(ctx, hipMemcpy(d_rp, crp->data(), sizeof(uint64_t) * (v_num + 1), hipMemcpyHostToDevice));
    if (cm) DG_HIP(ctx, hipMemcpy(d_nbr, cnbr->data(), sizeof(uint32_t) * cm, hipMemcpyHostToDevice));
  }
  if (sym && cm) {
    DG_HIP(ctx, hipMalloc((void**)&d_rev, sizeof(uint64_t) * cm));
    DG_HIP(ctx, hipMemcpyAsync(d_rev, rev.data(), sizeof(uint64_t) * cm, hipMemcpyHostToDevice, ctx->stream));
  }
  // hubs: hash-table offsets and work items
  std::vector<uint32_t> big, tsize, it_hub;
  std::vector<uint64_t> toff, it_k0;
  uint64_t tslots = 0;
  for (uint32_t v = 0; v < v_num; ++v) {
    const uint64_t d = (*crp)[v + 1] - (*crp)[v];
    if (d <= (uint64_t)kSmallDeg) continue;
    uint32_t sz = 64;
    while (sz < 2 * d) sz <<= 1;
    const uint32_t bi = (uint32_t)big.size();
    big.push_back(v);
    tsize.push_back(sz);
    toff.push_back(tslots);
    tslots += sz;
    for (uint64_t k0 = (*crp)[v]; k0 < (*crp)[v + 1]; k0 += kHubItem) {
      it_hub.push_back(bi);
      it_k0.push_back(k0);
    }
  }
  uint32_t *d_big = nullptr, *d_tsize = nullptr, *d_tab = nullptr, *d_it_hub = nullptr;
  uint64_t *d_toff = nullptr, *d_it_k0 = nullptr;
  if (!big.empty()) {
    DG_HIP(ctx, hipMalloc((void**)&d_big, sizeof(uint32_t) * big.size()));
    DG_HIP(ctx, hipMalloc((void**)&d_tsize, sizeof(uint32_t) * big.size()));
    DG_HIP(ctx, hipMalloc((void**)&d_toff, sizeof(uint64_t) * big.size()));
    DG_HIP(ctx, hipMalloc((void**)&d_tab, sizeof(uint32_t) * tslots));
    DG_HIP(ctx, hipMalloc((void**)&d_it_hub, sizeof(uint32_t) * it_hub.size()));
    DG_HIP(ctx, hipMalloc((void**)&d_it_k0, sizeof(uint64_t) * it_k0.size()));
    DG_HIP(ctx, hipMemcpy(d_big, big.data(), sizeof(uint32_t) * big.size(), hipMemcpyHostToDevice));
    DG_HIP(ctx, hipMemcpy(d_tsize, tsize.data(), sizeof(uint32_t) * big.size(), hipMemcpyHostToDevice));
    DG_HIP(ctx, hipMemcpy(d_toff, toff.data(), sizeof(uint64_t) * big.size(), hipMemcpyHostToDevice));
    DG_HIP(ctx, hipMemcpy(d_it_hub, it_hub.data(), sizeof(uint32_t) * it_hub.size(), hipMemcpyHostToDevice));
    DG_HIP(ctx, hipMemcpy(d_it_k0, it_k0.data(), sizeof(uint64_t) * it_k0.size(), hipMemcpyHostToDevice));
  }
  size_t tc;
  DG_CHECK(dg_time_begin(ctx, "cn", &tc));
  const unsigned gs = (unsigned)grid_for(ctx, (uint64_t)v_num, 4);
  if (!big.empty()) {
    fill_kernel<<<grid_for(ctx, tslots, 256), 256, 0, ctx->stream>>>(d_tab, tslots, kEmpty);
    hub_hash_kernel<<<(unsigned)std::min<size_t>(big.size(), (size_t)ctx->num_cus * 8), 256, 0, ctx->stream>>>(
        d_rp, d_nbr, d_big, (uint32_t)big.size(), d_toff, d_tsize, d_tab);
  }
  const unsigned gh = (unsigned)grid_for(ctx, (uint64_t)it_hub.size() * 64, 256);
  if (sym) {
    cn_small_kernel<true><<<gs, 256, 0, ctx->stream>>>(d_rp, d_nbr, v_num, d_rev, d_cnv);
    if (!big.empty())
      hub_cn_kernel<true><<<gh, 256, 0, ctx->stream>>>(d_rp, d_nbr, d_big, d_toff, d_tsize, d_tab, d_it_hub,
                                                       d_it_k0, it_hub.size(), d_rev, d_cnv);
  } else {
    cn_small_kernel<false><<<gs, 256, 0, ctx->stream>>>(d_rp, d_nbr, v_num, nullptr, d_cnv);
    if (!big.empty())
      hub_cn_kernel<false><<<gh, 256, 0, ctx->stream>>>(d_rp, d_nbr, d_big, d_toff, d_tsize, d_tab, d_it_hub,
                                                        d_it_k0, it_hub.size(), nullptr, d_cnv);
  }
  DG_HIP(ctx, hipGetLastError());
  DG_CHECK(dg_time_end(ctx, tc));
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  void* hub_bufs[] = {d_big, d_tsize, d_toff, d_tab, d_it_hub, d_it_k0};
  for (void* q : hub_bufs)
    if (q) hipFree(q);
  if (d_rev) hipFree(d_rev);
  if (!simple) {
    // each original entry takes the count of its pair in the distinct lists
    std::vector<uint32_t> scn(cm), cn(rp[v_num]);
    if (cm) DG_HIP(ctx, hipMemcpy(scn.data(), d_cnv, sizeof(uint32_t) * cm, hipMemcpyDeviceToHost));
    host_rows(v_num, [&](uint32_t lo, uint32_t hi) {
      for (uint32_t u = lo; u < hi; ++u) {
        const uint32_t* row = snbr.data() + srp[u];
        const uint64_t len = srp[u + 1] - srp[u];
        for (uint64_t k = rp[u]; k < rp[u + 1]; ++k) {
          const uint32_t v = nbr[k];
          if (v == u) {  // |N(u) \ {u}|
            cn[k] = (uint32_t)len;
            continue;
          }
          const uint32_t* it = std::lower_bound(row, row + len, v);
          cn[k] = scn[srp[u] + (uint64_t)(it - row)];
        }
      }
    });
    hipFree(d_rp);
    hipFree(d_nbr);
    hipFree(d_cnv);
    if (rp[v_num])
      DG_HIP(ctx, hipMemcpy(ctx->d_cn, cn.data(), sizeof(uint32_t) * rp[v_num], hipMemcpyHostToDevice));
  }
  return DG_OK;
}

static int finish_graph(dg_ctx* ctx, uint32_t v_num, const std::vector<uint64_t>& rp,
                        const std::vector<uint32_t>& nbr, const std::vector<uint32_t>& cn_data,
                        const std::vector<float>& weight, const std::vector<uint64_t>& rev,
                        bool undirected, int cn_mode) {
  dg_graph_free(ctx);
  const uint64_t m = rp[v_num];
  DG_HIP(ctx, hipMalloc((void**)&ctx->d_row_ptr, sizeof(uint64_t) * (v_num + 1)));
  DG_HIP(ctx, hipMalloc((void**)&ctx->d_nbr, sizeof(uint32_t) * (m ? m : 1)));
  DG_HIP(ctx, hipMalloc((void**)&ctx->d_cn, sizeof(uint32_t) * (m ? m : 1)));
  DG_HIP(ctx, hipMalloc((void**)&ctx->d_adj, sizeof(uint2) * (m ? m : 1)));
  DG_HIP(ctx, hipMalloc((void**)&ctx->d_deg, sizeof(uint32_t) * v_num));
  DG_HIP(ctx, hipMemcpyAsync(ctx->d_row_ptr, rp.data(), sizeof(uint64_t) * (v_num + 1),
                             hipMemcpyHostToDevice, ctx->stream));
  {
    // hot set for visit counting: the kHotMax highest-degree vertices (ties
    // by id); walks visit vertices in proportion to degree
    const uint32_t n_hot = v_num < kHotMax ? v_num : kHotMax;
    std::vector<uint32_t> order(v_num);
    for (uint32_t v = 0; v < v_num; ++v) order[v] = v;
    auto deg_of = [&](uint32_t v) { return rp[v + 1] - rp[v]; };
    std::partial_sort(order.begin(), order.begin() + n_hot, order.end(), [&](uint32_t a, uint32_t b) {
      return deg_of(a) != deg_of(b) ? deg_of(a) > deg_of(b) : a < b;
    });
    std::vector<uint16_t> slot(v_num, kNotHot);
    for (uint32_t k = 0; k < n_hot; ++k) slot[order[k]] = (uint16_t)k;
    DG_HIP(ctx, hipMalloc((void**)&ctx->d_hot_slot, sizeof(uint16_t) * (v_num ? v_num : 1)));
    DG_HIP(ctx, hipMalloc((void**)&ctx->d_hot_id, sizeof(uint32_t) * (n_hot ? n_hot : 1)));
    DG_HIP(ctx, hipMemcpy(ctx->d_hot_slot, slot.data(), sizeof(uint16_t) * v_num, hipMemcpyHostToDevice));
    DG_HIP(ctx, hipMemcpy(ctx->d_hot_id, order.data(), sizeof(uint32_t) * n_hot, hipMemcpyHostToDevice));
    ctx->n_hot = n_hot;
  }
  if (m)
    DG_HIP(ctx, hipMemcpyAsync(ctx->d_nbr, nbr.data(), sizeof(uint32_t) * m,
                               hipMemcpyHostToDevice, ctx->stream));
  ctx->v_num = v_num;
  ctx->m = m;
  degree_kernel<<<grid_for(ctx, v_num, 256), 256, 0, ctx->stream>>>(ctx->d_row_ptr, v_num,
                                                                    ctx->d_deg);
  DG_HIP(ctx, hipGetLastError());
  if (cn_mode == DG_CN_COMPUTE || cn_mode == DG_CN_COMPUTE_WEIGHTED) {
    DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    DG_CHECK(compute_cn(ctx, v_num, rp, nbr, rev, undirected));
  } else if (m) {
    DG_HIP(ctx, hipMemcpyAsync(ctx->d_cn, cn_data.data(), sizeof(uint32_t) * m,
                               hipMemcpyHostToDevice, ctx->stream));
  }
  // weights are only needed for the thresholds, so they are not kept
  float* d_w = nullptr;
  if (cn_mode == DG_CN_COMPUTE_WEIGHTED && m) {
    DG_HIP(ctx, hipMalloc((void**)&d_w, sizeof(float) * m));
    DG_HIP(ctx, hipMemcpyAsync(d_w, weight.data(), sizeof(float) * m, hipMemcpyHostToDevice,
                               ctx->stream));
  }
  size_t tt;
  DG_CHECK(dg_time_begin(ctx, "threshold", &tt));
  if (m)
    threshold_kernel<<<grid_for(ctx, m, 256), 256, 0, ctx->stream>>>(
        ctx->d_row_ptr, ctx->d_nbr, ctx->d_cn, d_w, ctx->d_deg, v_num, m, ctx->d_adj);
  DG_HIP(ctx, hipGetLastError());
  DG_CHECK(dg_time_end(ctx, tt));
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (d_w) hipFree(d_w);
  return DG_OK;
}

extern "C" int dg_graph_from_edges(dg_ctx* ctx, uint32_t v_num, const dg_edge* edges,
                                   uint64_t e_num, int undirected, int cn_mode) {
  if (!ctx || (!edges && e_num) || v_num == 0) return DG_ERR_ARG;
  if (cn_mode != DG_CN_FROM_DATA && cn_mode != DG_CN_COMPUTE && cn_mode != DG_CN_COMPUTE_WEIGHTED)
    return dg_fail(ctx, DG_ERR_ARG, "unknown cn_mode %d", cn_mode);
  DG_HIP(ctx, hipSetDevice(ctx->device));
  for (uint64_t i = 0; i < e_num; ++i)
    if (edges[i].src >= v_num || edges[i].dst >= v_num)
      return dg_fail(ctx, DG_ERR_ARG, "edge %llu (%u, %u) out of range for %u vertices",
                     (unsigned long long)i, edges[i].src, edges[i].dst, v_num);
  // include/graph.hpp:671-699 then :509-540: doubled list, counted, placed in order
  const uint64_t m = e_num * (undirected ? 2 : 1);
  std::vector<uint64_t> rp(v_num + 1, 0);
  for (uint64_t i = 0; i < e_num; ++i) {
    rp[edges[i].src + 1]++;
    if (undirected) rp[edges[i].dst + 1]++;
  }
  for (uint32_t v = 0; v < v_num; ++v) rp[v + 1] += rp[v];
  std::vector<uint64_t> fill(rp.begin(), rp.end() - 1);
  std::vector<uint32_t> nbr(m);
  std::vector<uint32_t> cn;
  std::vector<float> weight;
  std::vector<uint64_t> rev;  // undirected: position of each entry's reverse entry
  const bool want_rev = undirected && cn_mode != DG_CN_FROM_DATA;
  if (want_rev) rev.resize(m);
  const bool from_data = cn_mode == DG_CN_FROM_DATA, weighted = cn_mode == DG_CN_COMPUTE_WEIGHTED;
  if (from_data) cn.resize(m);
  if (weighted) weight.resize(m);
  for (uint64_t i = 0; i < e_num; ++i) {
    const uint64_t p = fill[edges[i].src]++;
    nbr[p] = edges[i].dst;
    // walk.hpp:920 converts the float edge data to int
    const uint32_t c = from_data ? (uint32_t)(int)edges[i].data : 0u;
    if (from_data) cn[p] = c;
    if (weighted) weight[p] = edges[i].data;
    if (undirected) {
      const uint64_t q = fill[edges[i].dst]++;
      nbr[q] = edges[i].src;
      if (from_data) cn[q] = c;
      if (weighted) weight[q] = edges[i].data;
      if (want_rev) {
        rev[p] = q;
        rev[q] = p;
      }
    }
  }
  dg_walk_free(ctx);  // walker state belongs to the previous graph
  return finish_graph(ctx, v_num, rp, nbr, cn, weight, rev, undirected != 0, cn_mode);
}

extern "C" int dg_graph_load_bin(dg_ctx* ctx, const char* path, uint32_t v_num, int undirected,
                                 int cn_mode) {
  if (!ctx || !path) return DG_ERR_ARG;
  FILE* f = fopen(path, "rb");
  if (!f) return dg_fail(ctx, DG_ERR_IO, "cannot open %s", path);
  fseek(f, 0, SEEK_END);
  long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  if (sz < 0 || sz % (long)sizeof(dg_edge) != 0) {
    fclose(f);
    return dg_fail(ctx, DG_ERR_IO, "%s: size %ld is not a multiple of %zu", path, sz,
                   sizeof(dg_edge));
  }
  const uint64_t e_num = (uint64_t)sz / sizeof(dg_edge);
  std::vector<dg_edge> edges(e_num);
  const size_t got = e_num ? fread(edges.data(), sizeof(dg_edge), e_num, f) : 0;
  fclose(f);
  if (got != e_num) return dg_fail(ctx, DG_ERR_IO, "%s: short read", path);
  return dg_graph_from_edges(ctx, v_num, edges.data(), e_num, undirected, cn_mode);
}

extern "C" int dg_graph_info(dg_ctx* ctx, uint32_t* v_num, uint64_t* m) {
  if (!ctx) return DG_ERR_ARG;
  if (v_num) *v_num = ctx->v_num;
  if (m) *m = ctx->m;
  return DG_OK;
}

extern "C" int dg_graph_get_csr(dg_ctx* ctx, uint64_t* row_ptr, uint32_t* nbr, uint32_t* cn) {
  if (!ctx) return DG_ERR_ARG;
  if (!ctx->d_row_ptr) return dg_fail(ctx, DG_ERR_STATE, "no graph loaded");
  DG_HIP(ctx, hipSetDevice(ctx->device));
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (row_ptr)
    DG_HIP(ctx, hipMemcpy(row_ptr, ctx->d_row_ptr, sizeof(uint64_t) * (ctx->v_num + 1),
                          hipMemcpyDeviceToHost));
  if (nbr && ctx->m)
    DG_HIP(ctx, hipMemcpy(nbr, ctx->d_nbr, sizeof(uint32_t) * ctx->m, hipMemcpyDeviceToHost));
  if (cn && ctx->m)
    DG_HIP(ctx, hipMemcpy(cn, ctx->d_cn, sizeof(uint32_t) * ctx->m, hipMemcpyDeviceToHost));
  return DG_OK;
}

extern "C" int dg_graph_get_thresholds(dg_ctx* ctx, uint32_t* thr) {
  if (!ctx || !thr) return DG_ERR_ARG;
  if (!ctx->d_adj) return dg_fail(ctx, DG_ERR_STATE, "no graph loaded");
  DG_HIP(ctx, hipSetDevice(ctx->device));
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<uint2> adj(ctx->m);
  if (ctx->m)
    DG_HIP(ctx, hipMemcpy(adj.data(), ctx->d_adj, sizeof(uint2) * ctx->m, hipMemcpyDeviceToHost));
  for (uint64_t e = 0; e < ctx->m; ++e) thr[e] = adj[e].y;
  return DG_OK;
}
