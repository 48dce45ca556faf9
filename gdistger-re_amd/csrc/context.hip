// context.hip - context lifetime, error reporting and kernel timing.
#include "dg_internal.h"
#include "dg_math.h"

#include <algorithm>
#include <cstring>

int dg_fail(dg_ctx* ctx, int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  return code;
}

extern "C" int dg_open(int device, dg_ctx** out) {
  if (!out) return DG_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return DG_ERR_HIP;
  if (device < 0 || device >= n) return DG_ERR_ARG;
  dg_ctx* ctx = new dg_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete ctx;
    return DG_ERR_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    delete ctx;
    return DG_ERR_HIP;
  }
  ctx->num_cus = prop.multiProcessorCount;
  // A blocking stream: work the library queues here is ordered with the
  // synchronous hipMemcpy/hipMemset calls it makes on the null stream.
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamDefault) != hipSuccess) {
    delete ctx;
    return DG_ERR_HIP;
  }
  if (hipMalloc((void**)&ctx->d_queue, sizeof(dg_walk_queue)) != hipSuccess ||
      hipHostMalloc((void**)&ctx->h_queue, sizeof(dg_walk_queue), hipHostMallocDefault) !=
          hipSuccess) {
    dg_close(ctx);
    return DG_ERR_NOMEM;
  }
  memset(ctx->h_queue, 0, sizeof(dg_walk_queue));
  if (dg_tuples_scratch_setup(ctx) != DG_OK) {
    dg_close(ctx);
    return DG_ERR_HIP;
  }
  *out = ctx;
  return DG_OK;
}

extern "C" void dg_close(dg_ctx* ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  dg_pipe_free(ctx);
  dg_dsgl_free(ctx);
  dg_tuples_free(ctx);
  dg_sgns_free(ctx);
  dg_walk_free(ctx);
  dg_graph_free(ctx);
  for (auto& p : ctx->pending) {
    hipEventDestroy(p.start);
    hipEventDestroy(p.stop);
  }
  for (auto e : ctx->event_pool) hipEventDestroy(e);
  if (ctx->d_stamps) hipFree(ctx->d_stamps);
  if (ctx->d_queue) hipFree(ctx->d_queue);
  if (ctx->h_queue) hipHostFree(ctx->h_queue);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" const char* dg_last_error(const dg_ctx* ctx) {
  return ctx ? ctx->err.c_str() : "no context";
}

extern "C" int dg_sync(dg_ctx* ctx) {
  if (!ctx) return DG_ERR_ARG;
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DG_OK;
}

static int take_event(dg_ctx* ctx, hipEvent_t* ev) {
  if (!ctx->event_pool.empty()) {
    *ev = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return DG_OK;
  }
  DG_HIP(ctx, hipEventCreate(ev));
  return DG_OK;
}

namespace {
constexpr uint32_t kStampIntervals = 4096;

// Writes the device's constant-rate wall clock to *slot.  Kernels of one
// captured stream run one after another, so two stamps bracket the work
// enqueued between them.
__global__ void stamp_kernel(unsigned long long* slot) {
  if (threadIdx.x == 0) *slot = (unsigned long long)wall_clock64();
}
}  // namespace

int dg_stamps_begin(dg_ctx* ctx) {
  if (!ctx->d_stamps) {
    DG_HIP(ctx, hipMalloc((void**)&ctx->d_stamps, sizeof(unsigned long long) * 2 * kStampIntervals));
    ctx->stamps_cap = kStampIntervals;
  }
  if (ctx->wall_khz <= 0.0) {
    int khz = 0;
    DG_HIP(ctx, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device));
    if (khz <= 0) return dg_fail(ctx, DG_ERR_HIP, "device reports no wall-clock rate");
    ctx->wall_khz = (double)khz;
  }
  ctx->stamp_names.clear();
  ctx->stamping = true;
  return DG_OK;
}

void dg_stamps_end(dg_ctx* ctx) { ctx->stamping = false; }

int dg_stamps_resolve(dg_ctx* ctx, double* walk_ms, double* sgns_ms) {
  *walk_ms = *sgns_ms = 0.0;
  const size_t n = ctx->stamp_names.size();
  if (n == 0) return DG_OK;
  std::vector<unsigned long long> h(2 * n);
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  DG_HIP(ctx, hipMemcpy(h.data(), ctx->d_stamps, sizeof(unsigned long long) * 2 * n,
                        hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; ++i) {
    const double ms = (double)(h[2 * i + 1] - h[2 * i]) / ctx->wall_khz;
    const std::string& name = ctx->stamp_names[i];
    auto& t = ctx->timing[name];
    t.launches += 1;
    t.total_ms += ms;
    if (name == "walk") *walk_ms += ms;
    if (name.rfind("sgns", 0) == 0) *sgns_ms += ms;
  }
  ctx->stamp_names.clear();
  return DG_OK;
}

int dg_time_begin(dg_ctx* ctx, const char* name, size_t* token) {
  if (ctx->stamping) {
    const size_t i = ctx->stamp_names.size();
    if (i >= ctx->stamps_cap) return dg_fail(ctx, DG_ERR_STATE, "too many timed intervals in one capture");
    ctx->stamp_names.push_back(name);
    stamp_kernel<<<1, 64, 0, ctx->stream>>>(ctx->d_stamps + 2 * i);
    DG_HIP(ctx, hipGetLastError());
    *token = i;
    return DG_OK;
  }
  dg_pending_event p;
  p.name = name;
  DG_CHECK(take_event(ctx, &p.start));
  DG_CHECK(take_event(ctx, &p.stop));
  DG_HIP(ctx, hipEventRecord(p.start, ctx->stream));
  ctx->pending.push_back(p);
  *token = ctx->pending.size() - 1;
  return DG_OK;
}

int dg_time_end(dg_ctx* ctx, size_t token) {
  if (ctx->stamping) {
    if (token >= ctx->stamp_names.size()) return dg_fail(ctx, DG_ERR_STATE, "bad timing token");
    stamp_kernel<<<1, 64, 0, ctx->stream>>>(ctx->d_stamps + 2 * token + 1);
    DG_HIP(ctx, hipGetLastError());
    return DG_OK;
  }
  if (token >= ctx->pending.size()) return dg_fail(ctx, DG_ERR_STATE, "bad timing token");
  DG_HIP(ctx, hipEventRecord(ctx->pending[token].stop, ctx->stream));
  return DG_OK;
}

int dg_time_resolve(dg_ctx* ctx) {
  if (ctx->pending.empty()) return DG_OK;
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (auto& p : ctx->pending) {
    float ms = 0.f;
    DG_HIP(ctx, hipEventElapsedTime(&ms, p.start, p.stop));
    auto& t = ctx->timing[p.name];
    t.launches += 1;
    t.total_ms += ms;
  }
  // back to the pool in reverse, so the next sequence of dg_time_begin calls
  // takes the same events in the same order (a recaptured graph then has the
  // same event nodes and updates in place)
  for (size_t i = ctx->pending.size(); i-- > 0;) {
    ctx->event_pool.push_back(ctx->pending[i].stop);
    ctx->event_pool.push_back(ctx->pending[i].start);
  }
  ctx->pending.clear();
  return DG_OK;
}

extern "C" int dg_timing_get(dg_ctx* ctx, const char* name, uint64_t* launches, double* total_ms) {
  if (!ctx || !name) return DG_ERR_ARG;
  DG_CHECK(dg_time_resolve(ctx));
  auto it = ctx->timing.find(name);
  if (launches) *launches = it == ctx->timing.end() ? 0 : it->second.launches;
  if (total_ms) *total_ms = it == ctx->timing.end() ? 0.0 : it->second.total_ms;
  return DG_OK;
}

extern "C" int dg_timing_reset(dg_ctx* ctx) {
  if (!ctx) return DG_ERR_ARG;
  DG_CHECK(dg_time_resolve(ctx));
  ctx->timing.clear();
  return DG_OK;
}

// ---- HBM copy calibration ----------------------------------------------------
// A plain streaming copy of known size; its measured bandwidth and its hardware
// byte counters calibrate the counters of the other kernels.
__global__ __launch_bounds__(256) void dg_copy_kernel(const uint4* __restrict__ src,
                                                      uint4* __restrict__ dst, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

extern "C" int dg_copy_bench(dg_ctx* ctx, uint64_t bytes, uint32_t iters, double* gbps) {
  if (!ctx || bytes < 16 || iters == 0) return DG_ERR_ARG;
  DG_HIP(ctx, hipSetDevice(ctx->device));
  const uint64_t n = bytes / 16;
  uint4 *a = nullptr, *b = nullptr;
  DG_HIP(ctx, hipMalloc((void**)&a, n * 16));
  hipError_t e = hipMalloc((void**)&b, n * 16);
  if (e != hipSuccess) {
    hipFree(a);
    return dg_fail(ctx, DG_ERR_NOMEM, "copy buffer: %s", hipGetErrorString(e));
  }
  hipMemsetAsync(a, 1, n * 16, ctx->stream);
  const unsigned grid = (unsigned)(ctx->num_cus * 8);
  size_t tok;
  int rc = DG_OK;
  for (uint32_t it = 0; it < iters && rc == DG_OK; ++it) {
    rc = dg_time_begin(ctx, "copy", &tok);
    if (rc == DG_OK) {
      dg_copy_kernel<<<grid, 256, 0, ctx->stream>>>(a, b, n);
      rc = dg_time_end(ctx, tok);
    }
  }
  uint64_t launches = 0;
  double ms = 0.0;
  if (rc == DG_OK) rc = dg_time_resolve(ctx);
  if (rc == DG_OK) {
    auto& t = ctx->timing["copy"];
    launches = t.launches;
    ms = t.total_ms;
  }
  hipFree(a);
  hipFree(b);
  if (rc != DG_OK) return rc;
  if (gbps) *gbps = launches ? (2.0 * (double)(n * 16) * (double)launches) / (ms * 1e-3) / 1e9 : 0.0;
  return DG_OK;
}

namespace {

__global__ void math_eval_kernel(int fn, const double* x, const double* y, uint64_t n, double* out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const double a = x[i];
    double r = 0.0;
    switch (fn) {
      case DG_MATH_SQRT: r = dg_sqrt(a); break;
      case DG_MATH_LOG2: r = dg_log2(a); break;
      case DG_MATH_LOG: r = dg_log(a); break;
      case DG_MATH_TANH: r = dg_tanh(a); break;
      default: r = dg_pow(a, y[i]); break;
    }
    out[i] = r;
  }
}

}  // namespace

extern "C" int dg_math_eval(dg_ctx* ctx, int fn, const double* x, const double* y, uint64_t n,
                            double* out) {
  if (!ctx || fn < DG_MATH_SQRT || fn > DG_MATH_POW || (n && (!x || !out)) ||
      (fn == DG_MATH_POW && n && !y))
    return DG_ERR_ARG;
  if (n == 0) return DG_OK;
  DG_HIP(ctx, hipSetDevice(ctx->device));
  double *dx = nullptr, *dy = nullptr, *dout = nullptr;
  int rc = DG_OK;
  hipError_t e = hipMalloc((void**)&dx, sizeof(double) * n);
  if (e == hipSuccess) e = hipMalloc((void**)&dout, sizeof(double) * n);
  if (e == hipSuccess && fn == DG_MATH_POW) e = hipMalloc((void**)&dy, sizeof(double) * n);
  if (e == hipSuccess) e = hipMemcpy(dx, x, sizeof(double) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess && dy) e = hipMemcpy(dy, y, sizeof(double) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, (uint64_t)ctx->num_cus * 16);
    math_eval_kernel<<<(unsigned)blocks, 256, 0, ctx->stream>>>(fn, dx, dy, n, dout);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = hipMemcpy(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost);
  if (e != hipSuccess) rc = dg_fail(ctx, DG_ERR_HIP, "dg_math_eval: %s", hipGetErrorString(e));
  hipFree(dx);
  hipFree(dy);
  hipFree(dout);
  return rc;
}
