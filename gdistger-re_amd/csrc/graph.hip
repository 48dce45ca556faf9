// graph.hip - device-resident CSR for the walker.
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
//   d_row_ptr : uint64[V+1]   row offsets (edge_id_t)
//   d_adj     : uint2[m]      {neighbour, common-neighbour count}; 8 bytes per
//                             entry like the reference AdjUnit<real_t>
//                             (include/graph.hpp:32-37), with the count stored
//                             as an integer instead of a float
//   d_deg     : uint32[V]     degree (= row length), read once per step
#include <algorithm>
#include <cstring>
#include <thread>

#include "dg_internal.h"

void dg_graph_free(dg_ctx* ctx) {
  if (ctx->d_row_ptr) hipFree(ctx->d_row_ptr);
  if (ctx->d_adj) hipFree(ctx->d_adj);
  if (ctx->d_deg) hipFree(ctx->d_deg);
  ctx->d_row_ptr = nullptr;
  ctx->d_adj = nullptr;
  ctx->d_deg = nullptr;
  ctx->v_num = 0;
  ctx->m = 0;
}

static unsigned host_threads() {
  unsigned n = std::thread::hardware_concurrency();
  if (n == 0) n = 1;
  return n > 16 ? 16 : n;
}

template <typename F>
static void parallel_for(uint64_t n, F f) {
  unsigned t = host_threads();
  if (n < 4096 || t == 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  for (unsigned i = 0; i < t; ++i) {
    uint64_t lo = n * i / t, hi = n * (i + 1) / t;
    th.emplace_back([=] { f(lo, hi); });
  }
  for (auto& x : th) x.join();
}

// Common-neighbour counts |(N(u) & N(v)) \ {u, v}| per CSR entry, over
// distinct neighbours (the quantity src/tools/gcn.cpp:304-357 writes into the
// edge data).  Sorted unique copies of every row, then a merge per entry.
static void host_common_neighbors(uint32_t v_num, const std::vector<uint64_t>& rp,
                                  const std::vector<uint32_t>& nbr, std::vector<uint32_t>& cn) {
  uint64_t m = rp[v_num];
  std::vector<uint32_t> srt(nbr);
  std::vector<uint32_t> ulen(v_num);
  parallel_for(v_num, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t v = lo; v < hi; ++v) {
      uint32_t* b = srt.data() + rp[v];
      uint32_t* e = srt.data() + rp[v + 1];
      std::sort(b, e);
      ulen[v] = (uint32_t)(std::unique(b, e) - b);
    }
  });
  cn.assign(m, 0);
  parallel_for(v_num, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t u = lo; u < hi; ++u) {
      const uint32_t* a = srt.data() + rp[u];
      uint32_t na = ulen[u];
      for (uint64_t e = rp[u]; e < rp[u + 1]; ++e) {
        uint32_t v = nbr[e];
        const uint32_t* b = srt.data() + rp[v];
        uint32_t nb = ulen[v];
        uint32_t i = 0, j = 0, c = 0;
        while (i < na && j < nb) {
          if (a[i] < b[j]) {
            ++i;
          } else if (a[i] > b[j]) {
            ++j;
          } else {
            c += (a[i] != (uint32_t)u && a[i] != v);
            ++i;
            ++j;
          }
        }
        cn[e] = c;
      }
    }
  });
}

static int upload_graph(dg_ctx* ctx, uint32_t v_num, const std::vector<uint64_t>& rp,
                        const std::vector<uint32_t>& nbr, const std::vector<uint32_t>& cn) {
  dg_graph_free(ctx);
  uint64_t m = rp[v_num];
  std::vector<uint2> adj(m);
  for (uint64_t e = 0; e < m; ++e) adj[e] = make_uint2(nbr[e], cn[e]);
  std::vector<uint32_t> deg(v_num);
  for (uint32_t v = 0; v < v_num; ++v) deg[v] = (uint32_t)(rp[v + 1] - rp[v]);
  DG_HIP(ctx, hipMalloc((void**)&ctx->d_row_ptr, sizeof(uint64_t) * (v_num + 1)));
  DG_HIP(ctx, hipMalloc((void**)&ctx->d_adj, sizeof(uint2) * (m ? m : 1)));
  DG_HIP(ctx, hipMalloc((void**)&ctx->d_deg, sizeof(uint32_t) * (v_num ? v_num : 1)));
  DG_HIP(ctx, hipMemcpy(ctx->d_row_ptr, rp.data(), sizeof(uint64_t) * (v_num + 1),
                        hipMemcpyHostToDevice));
  if (m)
    DG_HIP(ctx, hipMemcpy(ctx->d_adj, adj.data(), sizeof(uint2) * m, hipMemcpyHostToDevice));
  if (v_num)
    DG_HIP(ctx, hipMemcpy(ctx->d_deg, deg.data(), sizeof(uint32_t) * v_num, hipMemcpyHostToDevice));
  ctx->v_num = v_num;
  ctx->m = m;
  return DG_OK;
}

extern "C" int dg_graph_from_edges(dg_ctx* ctx, uint32_t v_num, const dg_edge* edges,
                                   uint64_t e_num, int undirected, int cn_mode) {
  if (!ctx || (!edges && e_num) || v_num == 0) return DG_ERR_ARG;
  if (cn_mode != DG_CN_FROM_DATA && cn_mode != DG_CN_COMPUTE)
    return dg_fail(ctx, DG_ERR_ARG, "unknown cn_mode %d", cn_mode);
  DG_HIP(ctx, hipSetDevice(ctx->device));
  for (uint64_t i = 0; i < e_num; ++i)
    if (edges[i].src >= v_num || edges[i].dst >= v_num)
      return dg_fail(ctx, DG_ERR_ARG, "edge %llu (%u, %u) out of range for %u vertices",
                     (unsigned long long)i, edges[i].src, edges[i].dst, v_num);
  uint64_t m = e_num * (undirected ? 2 : 1);
  std::vector<uint64_t> rp(v_num + 1, 0);
  for (uint64_t i = 0; i < e_num; ++i) {
    rp[edges[i].src + 1]++;
    if (undirected) rp[edges[i].dst + 1]++;
  }
  for (uint32_t v = 0; v < v_num; ++v) rp[v + 1] += rp[v];
  std::vector<uint64_t> fill(rp.begin(), rp.end() - 1);
  std::vector<uint32_t> nbr(m), cn(m);
  for (uint64_t i = 0; i < e_num; ++i) {
    // walk.hpp:920 converts the float edge data to int
    uint32_t c = (uint32_t)(int)edges[i].data;
    uint64_t p = fill[edges[i].src]++;
    nbr[p] = edges[i].dst;
    cn[p] = c;
    if (undirected) {
      uint64_t q = fill[edges[i].dst]++;
      nbr[q] = edges[i].src;
      cn[q] = c;
    }
  }
  if (cn_mode == DG_CN_COMPUTE) host_common_neighbors(v_num, rp, nbr, cn);
  // reset walker state bound to the previous graph
  dg_walk_free(ctx);
  return upload_graph(ctx, v_num, rp, nbr, cn);
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
  uint64_t e_num = (uint64_t)sz / sizeof(dg_edge);
  std::vector<dg_edge> edges(e_num);
  size_t got = e_num ? fread(edges.data(), sizeof(dg_edge), e_num, f) : 0;
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
  if ((nbr || cn) && ctx->m) {
    std::vector<uint2> adj(ctx->m);
    DG_HIP(ctx, hipMemcpy(adj.data(), ctx->d_adj, sizeof(uint2) * ctx->m, hipMemcpyDeviceToHost));
    for (uint64_t e = 0; e < ctx->m; ++e) {
      if (nbr) nbr[e] = adj[e].x;
      if (cn) cn[e] = adj[e].y;
    }
  }
  return DG_OK;
}
