// host_data.cpp - host-side data utilities of the library: seeded synthetic
// graph generators, the link-prediction split, AUC and the reference binary
// edge format.  Plain C++ (no device code); every random choice comes from a
// Philox stream so a given seed gives the same bytes on any machine and with
// any thread count.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <thread>
#include <unordered_set>
#include <vector>

#include "../../include/distger_hip.h"
#include "dg_philox.h"

namespace {

unsigned host_threads() {
  unsigned n = std::thread::hardware_concurrency();
  if (n == 0) n = 1;
  return n > 16 ? 16 : n;
}

template <typename F>
void parallel_for(uint64_t n, F f) {
  unsigned t = host_threads();
  if (n < 65536 || t == 1) {
    f(0, n, 0u);
    return;
  }
  std::vector<std::thread> th;
  for (unsigned i = 0; i < t; ++i) {
    uint64_t lo = n * i / t, hi = n * (i + 1) / t;
    th.emplace_back([=] { f(lo, hi, i); });
  }
  for (auto& x : th) x.join();
}

// Uniform double in [0, 1) with 53 random bits from two stream words.
double u01(dg_stream* s) {
  uint64_t a = dg_stream_next(s) >> 5;  // 27 bits
  uint64_t b = dg_stream_next(s) >> 6;  // 26 bits
  return (double)((a << 26) | b) * (1.0 / 9007199254740992.0);
}

// Parallel sort of 64-bit keys: sorted chunks merged pairwise.
void parallel_sort(std::vector<uint64_t>& v) {
  unsigned t = host_threads();
  size_t n = v.size();
  if (n < (1u << 20) || t == 1) {
    std::sort(v.begin(), v.end());
    return;
  }
  std::vector<size_t> cut(t + 1);
  for (unsigned i = 0; i <= t; ++i) cut[i] = n * i / t;
  {
    std::vector<std::thread> th;
    for (unsigned i = 0; i < t; ++i)
      th.emplace_back([&, i] { std::sort(v.begin() + cut[i], v.begin() + cut[i + 1]); });
    for (auto& x : th) x.join();
  }
  for (size_t width = 1; width < t; width *= 2) {
    std::vector<std::thread> th;
    for (size_t i = 0; i + width < t; i += 2 * width) {
      size_t lo = cut[i], mid = cut[i + width], hi = cut[std::min<size_t>(i + 2 * width, t)];
      th.emplace_back([&v, lo, mid, hi] {
        std::inplace_merge(v.begin() + lo, v.begin() + mid, v.begin() + hi);
      });
    }
    for (auto& x : th) x.join();
  }
}

inline uint64_t pair_key(uint32_t a, uint32_t b) {
  uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
  return ((uint64_t)lo << 32) | hi;
}

// Sorted, de-duplicated keys -> edges in a seeded random order with a random
// orientation, so the CSR built from them has no accidental row order.
int emit_edges(std::vector<uint64_t>& keys, uint64_t seed, uint32_t salt, dg_edge* edges,
               uint64_t e_cap, uint64_t* e_num, const float* data_or_null) {
  const uint64_t n = keys.size();
  if (n > e_cap || n >= (1ull << 32)) return DG_ERR_ARG;
  // order key = (random 31-bit tag, index): sorting it is a seeded shuffle
  std::vector<uint64_t> order(n);
  parallel_for(n, [&](uint64_t lo, uint64_t hi, unsigned) {
    for (uint64_t i = lo; i < hi; ++i) {
      dg_stream s;
      dg_stream_init(&s, seed, (uint32_t)i, salt, DG_TAG_GEN);
      uint64_t tag = (uint64_t)(dg_stream_next(&s) & 0x7fffffffu);
      order[i] = (tag << 32) | i;
    }
  });
  parallel_sort(order);
  parallel_for(n, [&](uint64_t lo, uint64_t hi, unsigned) {
    for (uint64_t j = lo; j < hi; ++j) {
      uint64_t i = order[j] & 0xffffffffu;
      uint32_t a = (uint32_t)(keys[i] >> 32), b = (uint32_t)keys[i];
      bool flip = ((order[j] >> 32) & 1u) != 0;
      edges[j].src = flip ? b : a;
      edges[j].dst = flip ? a : b;
      edges[j].data = data_or_null ? data_or_null[i] : 0.0f;
    }
  });
  *e_num = n;
  return DG_OK;
}

}  // namespace

extern "C" int dg_gen_chung_lu(uint32_t v_num, uint64_t e_target, double gamma,
                               uint32_t min_degree, uint64_t seed, dg_edge* edges, uint64_t e_cap,
                               uint64_t* e_num) {
  if (v_num < 2 || !(gamma > 1.0) || !edges || !e_num) return DG_ERR_ARG;
  const double alpha = 1.0 / (gamma - 1.0);
  const double total = 2.0 * (double)e_target;
  std::vector<double> w(v_num);
  for (uint32_t i = 0; i < v_num; ++i) w[i] = std::pow((double)(i + 1), -alpha);
  // scale to the target degree sum with a structural cutoff sqrt(total)
  const double cap = std::sqrt(total);
  std::vector<char> clipped(v_num, 0);
  for (int it = 0; it < 50; ++it) {
    double free_sum = 0.0, fixed = 0.0;
    for (uint32_t i = 0; i < v_num; ++i) {
      if (clipped[i])
        fixed += cap;
      else
        free_sum += w[i];
    }
    double scale = (total - fixed) / free_sum;
    bool changed = false;
    for (uint32_t i = 0; i < v_num; ++i) {
      if (clipped[i]) continue;
      w[i] *= scale;
      if (w[i] > cap) {
        clipped[i] = 1;
        changed = true;
      }
    }
    if (!changed) break;
  }
  for (uint32_t i = 0; i < v_num; ++i)
    if (clipped[i]) w[i] = cap;
  std::vector<double> cdf(v_num);
  double acc = 0.0;
  for (uint32_t i = 0; i < v_num; ++i) {
    acc += w[i];
    cdf[i] = acc;
  }
  auto draw_vertex = [&](dg_stream* s) -> uint32_t {
    double x = u01(s) * acc;
    uint32_t k = (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), x) - cdf.begin());
    return k < v_num ? k : v_num - 1;
  };
  // draw e_target endpoint pairs; drop self-loops; de-duplicate
  std::vector<uint64_t> keys(e_target);
  parallel_for(e_target, [&](uint64_t lo, uint64_t hi, unsigned) {
    for (uint64_t i = lo; i < hi; ++i) {
      dg_stream s;
      dg_stream_init(&s, seed, (uint32_t)i, (uint32_t)(i >> 32), DG_TAG_GEN);
      uint32_t a = draw_vertex(&s), b = draw_vertex(&s);
      keys[i] = (a == b) ? ~0ull : pair_key(a, b);
    }
  });
  parallel_sort(keys);
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  if (!keys.empty() && keys.back() == ~0ull) keys.pop_back();
  // top up vertices below min_degree with partners drawn ~ w
  if (min_degree > 0) {
    std::vector<uint32_t> deg(v_num, 0);
    for (uint64_t k : keys) {
      deg[k >> 32]++;
      deg[(uint32_t)k]++;
    }
    std::unordered_set<uint64_t> added;
    std::vector<uint64_t> extra;
    dg_stream s;
    dg_stream_init(&s, seed, 0xffffffffu, 0x70707070u, DG_TAG_GEN);
    for (uint32_t v = 0; v < v_num; ++v) {
      uint32_t tries = 0;
      while (deg[v] < min_degree && tries < 1000) {
        ++tries;
        uint32_t u = draw_vertex(&s);
        if (u == v) continue;
        uint64_t k = pair_key(u, v);
        if (std::binary_search(keys.begin(), keys.end(), k) || added.count(k)) continue;
        added.insert(k);
        extra.push_back(k);
        deg[v]++;
        deg[u]++;
      }
    }
    keys.insert(keys.end(), extra.begin(), extra.end());
    std::sort(keys.begin(), keys.end());
  }
  return emit_edges(keys, seed, 0x43484c55u, edges, e_cap, e_num, nullptr);
}

extern "C" int dg_gen_zipf(uint32_t v_num, double s, uint64_t n, uint64_t seed, uint32_t* out) {
  if (v_num == 0 || !(s > 0) || (n && !out)) return DG_ERR_ARG;
  std::vector<double> cdf(v_num);
  double acc = 0.0;
  for (uint32_t k = 0; k < v_num; ++k) {
    acc += std::pow((double)k + 1.0, -s);
    cdf[k] = acc;
  }
  parallel_for(n, [&](uint64_t lo, uint64_t hi, unsigned) {
    for (uint64_t i = lo; i < hi; ++i) {
      dg_stream st;
      dg_stream_init(&st, seed, (uint32_t)i, (uint32_t)(i >> 32), DG_TAG_TUPLE);
      const double u = u01(&st) * acc;
      const uint64_t k = (uint64_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
      out[i] = (uint32_t)(k < v_num ? k : v_num - 1);
    }
  });
  return DG_OK;
}

extern "C" int dg_gen_normal(uint64_t n, double sigma, uint64_t seed, float* out) {
  if (n && !out) return DG_ERR_ARG;
  const uint64_t pairs = (n + 1) / 2;
  parallel_for(pairs, [&](uint64_t lo, uint64_t hi, unsigned) {
    for (uint64_t i = lo; i < hi; ++i) {
      dg_stream st;
      dg_stream_init(&st, seed, (uint32_t)i, (uint32_t)(i >> 32), DG_TAG_TUPLE ^ 0x4e4f524du);
      const double u1 = u01(&st), u2 = u01(&st);
      const double r = std::sqrt(-2.0 * std::log(1.0 - u1));  // 1 - u1 in (0, 1]
      const double th = 6.283185307179586 * u2;
      out[2 * i] = (float)(sigma * r * std::cos(th));
      if (2 * i + 1 < n) out[2 * i + 1] = (float)(sigma * r * std::sin(th));
    }
  });
  return DG_OK;
}

extern "C" int dg_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                           uint64_t seed, dg_edge* edges, uint64_t e_cap, uint64_t* e_num) {
  if (scale == 0 || scale > 31 || !edges || !e_num) return DG_ERR_ARG;
  if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0)) return DG_ERR_ARG;
  const uint64_t draws = (uint64_t)edge_factor << scale;
  std::vector<uint64_t> keys(draws);
  parallel_for(draws, [&](uint64_t lo, uint64_t hi, unsigned) {
    for (uint64_t i = lo; i < hi; ++i) {
      dg_stream s;
      dg_stream_init(&s, seed, (uint32_t)i, (uint32_t)(i >> 32), DG_TAG_GEN);
      uint32_t src = 0, dst = 0;
      for (uint32_t bit = 0; bit < scale; ++bit) {
        double r = (double)dg_stream_next(&s) * (1.0 / 4294967296.0);
        uint32_t qs = r < a ? 0u : (r < a + b ? 1u : (r < a + b + c ? 2u : 3u));
        src = (src << 1) | (qs >> 1);
        dst = (dst << 1) | (qs & 1u);
      }
      keys[i] = (src == dst) ? ~0ull : pair_key(src, dst);
    }
  });
  parallel_sort(keys);
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  if (!keys.empty() && keys.back() == ~0ull) keys.pop_back();
  std::vector<float> wts(keys.size());
  parallel_for(keys.size(), [&](uint64_t lo, uint64_t hi, unsigned) {
    for (uint64_t i = lo; i < hi; ++i) {
      dg_stream s;
      dg_stream_init(&s, seed, (uint32_t)i, (uint32_t)(i >> 32), DG_TAG_GEN ^ 0x57u);
      // uniform(0, 1]: 24-bit grid 1/2^24 .. 1
      wts[i] = (float)((dg_stream_next(&s) >> 8) + 1u) * (1.0f / 16777216.0f);
    }
  });
  return emit_edges(keys, seed, 0x524d4154u, edges, e_cap, e_num, wts.data());
}

extern "C" int dg_lp_split(uint32_t v_num, dg_edge* edges, uint64_t e_num, double frac,
                           uint64_t seed, uint64_t* n_train, dg_edge* neg, uint64_t neg_cap,
                           uint64_t* n_neg) {
  if (!edges || !n_train || !n_neg || !(frac >= 0.0 && frac < 1.0)) return DG_ERR_ARG;
  // gtrain.cpp:283: npos = sample * edges.size() (truncating)
  const uint64_t npos = (uint64_t)(frac * (double)e_num);
  if (npos > neg_cap || (npos && !neg)) return DG_ERR_ARG;
  std::vector<uint32_t> deg(v_num, 0);
  for (uint64_t i = 0; i < e_num; ++i) {
    if (edges[i].src >= v_num || edges[i].dst >= v_num) return DG_ERR_ARG;
    deg[edges[i].src]++;
    deg[edges[i].dst]++;
  }
  // sorted keys of the full graph for the non-edge test (gtrain.cpp:323-336)
  std::vector<uint64_t> all(e_num);
  for (uint64_t i = 0; i < e_num; ++i) all[i] = pair_key(edges[i].src, edges[i].dst);
  parallel_sort(all);
  dg_stream s;
  dg_stream_init(&s, seed, 1u, 0u, DG_TAG_SPLIT);
  // gtrain.cpp:295-316: draw among the not-yet-removed edges, keep both
  // endpoints at degree > 1, swap the chosen edge to the tail
  uint64_t tail = e_num;  // edges[tail..e_num) are positives
  const uint64_t max_tries = 64 * (e_num + 1);
  uint64_t tries = 0;
  for (uint64_t k = 0; k < npos; ++k) {
    uint64_t ep;
    while (true) {
      if (++tries > max_tries) return DG_ERR_STATE;
      uint64_t r = ((uint64_t)dg_stream_next(&s) << 32) | dg_stream_next(&s);
      ep = r % tail;
      if (deg[edges[ep].src] > 1 && deg[edges[ep].dst] > 1) break;
    }
    deg[edges[ep].src]--;
    deg[edges[ep].dst]--;
    --tail;
    std::swap(edges[ep], edges[tail]);
  }
  std::reverse(edges + tail, edges + e_num);  // positives in draw order
  *n_train = tail;
  // uniform non-edges, distinct as unordered pairs, no self-pairs
  std::unordered_set<uint64_t> chosen;
  chosen.reserve((size_t)npos * 2 + 1);
  uint64_t nn = 0;
  tries = 0;
  while (nn < npos) {
    if (++tries > max_tries) return DG_ERR_STATE;
    uint32_t a = (uint32_t)(((uint64_t)dg_stream_next(&s) * v_num) >> 32);
    uint32_t b = (uint32_t)(((uint64_t)dg_stream_next(&s) * v_num) >> 32);
    if (a == b) continue;
    uint64_t key = pair_key(a, b);
    if (std::binary_search(all.begin(), all.end(), key) || chosen.count(key)) continue;
    chosen.insert(key);
    neg[nn].src = a;
    neg[nn].dst = b;
    neg[nn].data = 0.0f;
    ++nn;
  }
  *n_neg = nn;
  return DG_OK;
}

extern "C" double dg_auc(const float* pos, uint64_t n_pos, const float* negs, uint64_t n_neg) {
  if (n_pos == 0 || n_neg == 0) return 0.5;
  std::vector<std::pair<float, int>> all;
  all.reserve(n_pos + n_neg);
  for (uint64_t i = 0; i < n_pos; ++i) all.emplace_back(pos[i], 1);
  for (uint64_t i = 0; i < n_neg; ++i) all.emplace_back(negs[i], 0);
  std::sort(all.begin(), all.end(),
            [](const std::pair<float, int>& x, const std::pair<float, int>& y) {
              return x.first < y.first;
            });
  // rank sum of the positives with average ranks for ties (1-based ranks)
  double rank_sum = 0.0;
  size_t i = 0;
  while (i < all.size()) {
    size_t j = i;
    while (j < all.size() && all[j].first == all[i].first) ++j;
    double avg = 0.5 * ((double)(i + 1) + (double)j);
    for (size_t k = i; k < j; ++k)
      if (all[k].second) rank_sum += avg;
    i = j;
  }
  double np = (double)n_pos, nn = (double)n_neg;
  return (rank_sum - np * (np + 1.0) / 2.0) / (np * nn);
}

extern "C" int dg_write_edges_bin(const char* path, const dg_edge* edges, uint64_t e_num) {
  FILE* f = fopen(path, "wb");
  if (!f) return DG_ERR_IO;
  size_t w = e_num ? fwrite(edges, sizeof(dg_edge), e_num, f) : 0;
  int rc = fclose(f);
  return (w == e_num && rc == 0) ? DG_OK : DG_ERR_IO;
}

extern "C" int dg_read_edges_bin(const char* path, dg_edge* edges, uint64_t e_cap,
                                 uint64_t* e_num) {
  FILE* f = fopen(path, "rb");
  if (!f) return DG_ERR_IO;
  fseek(f, 0, SEEK_END);
  long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  if (sz < 0 || sz % (long)sizeof(dg_edge) != 0) {
    fclose(f);
    return DG_ERR_IO;
  }
  uint64_t n = (uint64_t)sz / sizeof(dg_edge);
  if (e_num) *e_num = n;
  if (!edges) {
    fclose(f);
    return DG_OK;
  }
  if (n > e_cap) {
    fclose(f);
    return DG_ERR_ARG;
  }
  size_t got = n ? fread(edges, sizeof(dg_edge), n, f) : 0;
  fclose(f);
  return got == n ? DG_OK : DG_ERR_IO;
}

extern "C" int dg_convert_text_edges(const char* in_path, const char* out_path, int weighted,
                                     uint64_t* e_num) {
  if (!in_path || !out_path) return DG_ERR_ARG;
  FILE* in = fopen(in_path, "r");
  if (!in) return DG_ERR_IO;
  FILE* out = fopen(out_path, "wb");
  if (!out) {
    fclose(in);
    return DG_ERR_IO;
  }
  // records are buffered and written in blocks; the reference collects them
  // in a vector and writes once (gconverter.cpp:70-76, storage.hpp:108-115)
  std::vector<char> buf;
  buf.reserve(1 << 22);
  uint64_t n = 0;
  bool ok = true;
  uint32_t rec[3];
  const size_t rsize = weighted ? 12 : 8;
  while (ok) {
    unsigned src = 0, dst = 0;
    float w = 0.0f;
    if (weighted) {
      if (fscanf(in, "%u %u %f", &src, &dst, &w) != 3) break;
    } else if (fscanf(in, "%u %u", &src, &dst) != 2) {
      break;
    }
    rec[0] = src;
    rec[1] = dst;
    memcpy(&rec[2], &w, sizeof(float));
    const char* p = reinterpret_cast<const char*>(rec);
    buf.insert(buf.end(), p, p + rsize);
    ++n;
    if (buf.size() >= (1u << 22)) {
      ok = fwrite(buf.data(), 1, buf.size(), out) == buf.size();
      buf.clear();
    }
  }
  if (ok && !buf.empty()) ok = fwrite(buf.data(), 1, buf.size(), out) == buf.size();
  fclose(in);
  if (fclose(out) != 0) ok = false;
  if (e_num) *e_num = n;
  return ok ? DG_OK : DG_ERR_IO;
}

extern "C" int dg_lp_write_files(const char* train_path, const char* test_edges_path,
                                 const char* test_labels_path, const dg_edge* train,
                                 uint64_t n_train, const dg_edge* pos, uint64_t n_pos,
                                 const dg_edge* neg, uint64_t n_neg) {
  if ((n_train && !train) || (n_pos && !pos) || (n_neg && !neg)) return DG_ERR_ARG;
  auto write_all = [](const char* path, auto&& emit) -> bool {
    if (!path) return true;
    FILE* f = fopen(path, "w");
    if (!f) return false;
    bool ok = emit(f);
    return fclose(f) == 0 && ok;
  };
  bool ok = write_all(train_path, [&](FILE* f) {
    for (uint64_t i = 0; i < n_train; ++i)
      if (fprintf(f, "%u %u %u\n", train[i].src, train[i].dst, (uint32_t)train[i].data) < 0) return false;
    return true;
  });
  ok = ok && write_all(test_edges_path, [&](FILE* f) {
    for (uint64_t i = 0; i < n_pos; ++i)
      if (fprintf(f, "%u %u\n", pos[i].src, pos[i].dst) < 0) return false;
    for (uint64_t i = 0; i < n_neg; ++i)
      if (fprintf(f, "%u %u\n", neg[i].src, neg[i].dst) < 0) return false;
    return true;
  });
  ok = ok && write_all(test_labels_path, [&](FILE* f) {
    for (uint64_t i = 0; i < n_pos; ++i)
      if (fputs("1\n", f) < 0) return false;
    for (uint64_t i = 0; i < n_neg; ++i)
      if (fputs("0\n", f) < 0) return false;
    return true;
  });
  return ok ? DG_OK : DG_ERR_IO;
}

extern "C" float dg_lr_value(int kind, float initial, float a, float b, int step) {
  switch (kind) {
    case DG_LR_FIXED:
      return initial;
    case DG_LR_EXPONENTIAL:  // lr_scheduler.hpp:23-31: float exp of a float argument
      return initial * std::exp(-a * (float)step);
    case DG_LR_COSINE: {  // lr_scheduler.hpp:33-47: min_lr held as double, cos in double
      const int t_max = (int)b;
      if (t_max <= 0) return std::nanf("");
      const float cosine = (float)std::cos(M_PI * step / t_max);
      const double min_lr = (double)a;
      return (float)(min_lr + 0.5 * (initial - min_lr) * (1 + cosine));
    }
    case DG_LR_STEP: {  // lr_scheduler.hpp:49-60: pow(float, int) in double
      const int step_size = (int)b;
      if (step_size <= 0) return std::nanf("");
      const int factor = step / step_size;
      return (float)(initial * std::pow((double)a, (double)factor));
    }
    default:
      return std::nanf("");
  }
}
