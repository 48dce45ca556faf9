// context.hip - context lifetime, error reporting and kernel timing.
#include "dg_internal.h"

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
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
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
  *out = ctx;
  return DG_OK;
}

extern "C" void dg_close(dg_ctx* ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  dg_sgns_free(ctx);
  dg_walk_free(ctx);
  dg_graph_free(ctx);
  for (auto& p : ctx->pending) {
    hipEventDestroy(p.start);
    hipEventDestroy(p.stop);
  }
  for (auto e : ctx->event_pool) hipEventDestroy(e);
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

int dg_time_begin(dg_ctx* ctx, const char* name, size_t* token) {
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
    ctx->event_pool.push_back(p.start);
    ctx->event_pool.push_back(p.stop);
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
