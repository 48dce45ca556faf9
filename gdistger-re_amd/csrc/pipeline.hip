// pipeline.hip - one walk round and the trainer's pass over its corpus as a
// single HIP graph.
//
// In the reference the walker writes each round to a walk file and a trainer
// thread picks the file up from a SyncQueue (include/type.hpp:29-61,
// include/walk.hpp:686, src/word2vec.cu:1285-1300).  Here the round's corpus
// never leaves the device, and the whole step (walk, packing and visit
// counts, relative entropy, subsampling, draws and the sentence kernels of
// every launch) is recorded by stream capture and launched as one graph: no
// host wait between its kernels.
//
// The host does not know the round's token count while the step is captured,
// so every buffer is sized for any outcome and each kernel reads its extent
// from the device (offsets, queue counters).  A round that overflows the
// walker's token buffer, or hits the length limit, is emptied on the device
// by round_guard_kernel; the rest of the step then changes nothing, the host
// grows the buffer and runs the step again (the walks depend only on seed,
// round and walker id, so the second run walks the same walks).
//
// Kernel arguments change every call (round, epoch, learning rate), so the
// step is captured on every call and the instantiated graph is updated in
// place (hipGraphExecUpdate) when the topology is unchanged; otherwise it is
// instantiated again.
#include "dg_internal.h"

struct dg_pipe_state {
  hipGraphExec_t exec = nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  uint64_t instantiations = 0;
  uint64_t updates = 0;
  uint64_t launches = 0;
};

void dg_pipe_free(dg_ctx* ctx) {
  dg_pipe_state* ps = ctx->pipe;
  if (!ps) return;
  if (ps->exec) hipGraphExecDestroy(ps->exec);
  if (ps->t0) hipEventDestroy(ps->t0);
  if (ps->t1) hipEventDestroy(ps->t1);
  delete ps;
  ctx->pipe = nullptr;
}

// Capture the step into a graph and make pipe->exec run it.
// Timing inside the graph uses device wall-clock stamps (events recorded in
// a capture cannot be timed).
static int capture_step(dg_ctx* ctx, dg_pipe_state* ps) {
  hipGraph_t g = nullptr;
  DG_CHECK(dg_stamps_begin(ctx));
  const hipError_t eb = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
  if (eb != hipSuccess) {
    dg_stamps_end(ctx);
    return dg_fail(ctx, DG_ERR_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(eb));
  }
  int rc = dg_walk_graph_enqueue(ctx, true);
  if (rc == DG_OK) rc = dg_sgns_graph_enqueue(ctx, ctx->d_tokens, ctx->d_offsets, true);
  const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
  dg_stamps_end(ctx);
  if (rc != DG_OK || ec != hipSuccess) {
    if (g) hipGraphDestroy(g);
    ctx->stamp_names.clear();
    if (rc != DG_OK) return rc;
    return dg_fail(ctx, DG_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ec));
  }
  if (ps->exec) {
    hipGraphNode_t err_node = nullptr;
    hipGraphExecUpdateResult res = hipGraphExecUpdateError;
    const hipError_t eu = hipGraphExecUpdate(ps->exec, g, &err_node, &res);
    if (eu == hipSuccess && res == hipGraphExecUpdateSuccess) {
      ps->updates++;
    } else {
      (void)hipGetLastError();
      hipGraphExecDestroy(ps->exec);
      ps->exec = nullptr;
    }
  }
  if (!ps->exec) {
    const hipError_t ei = hipGraphInstantiate(&ps->exec, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
      ps->exec = nullptr;
      hipGraphDestroy(g);
      return dg_fail(ctx, DG_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
    }
    ps->instantiations++;
  }
  DG_HIP(ctx, hipGraphDestroy(g));
  return DG_OK;
}

extern "C" int dg_pipeline_step(dg_ctx* ctx, const dg_walk_params* wp, uint32_t round,
                                uint32_t walker_begin, uint32_t walker_num,
                                const dg_sgns_params* sp, float alpha, dg_round_stats* rs,
                                dg_sgns_stats* ss) {
  if (!ctx || !wp || !sp) return DG_ERR_ARG;
  if (walker_num == 0) return dg_fail(ctx, DG_ERR_ARG, "walker_num must be positive");
  DG_HIP(ctx, hipSetDevice(ctx->device));
  // the trainer's side first: a step it cannot run must not touch the round
  DG_CHECK(dg_sgns_graph_check(ctx, sp));
  if (!ctx->pipe) {
    dg_pipe_state* ps = new dg_pipe_state();
    if (hipEventCreate(&ps->t0) != hipSuccess || hipEventCreate(&ps->t1) != hipSuccess) {
      ctx->pipe = ps;
      dg_pipe_free(ctx);
      return dg_fail(ctx, DG_ERR_HIP, "hipEventCreate failed");
    }
    ctx->pipe = ps;
  }
  dg_pipe_state* ps = ctx->pipe;
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int attempt = 0; attempt < 3; ++attempt) {
    uint64_t tok_bound = 0, slot_tokens = 0, overflow_tokens = 0;
    DG_CHECK(dg_walk_graph_plan(ctx, wp, round, walker_begin, walker_num, &tok_bound, &slot_tokens,
                                &overflow_tokens));
    DG_CHECK(dg_sgns_graph_plan(ctx, sp, alpha, walker_num, tok_bound, slot_tokens,
                                overflow_tokens));
    DG_CHECK(capture_step(ctx, ps));
    DG_HIP(ctx, hipEventRecord(ps->t0, ctx->stream));
    DG_HIP(ctx, hipGraphLaunch(ps->exec, ctx->stream));
    DG_HIP(ctx, hipEventRecord(ps->t1, ctx->stream));
    DG_HIP(ctx, hipEventSynchronize(ps->t1));
    ps->launches++;
    float step_ms = 0.f;
    DG_HIP(ctx, hipEventElapsedTime(&step_ms, ps->t0, ps->t1));
    double walk_ms = 0.0, sgns_ms = 0.0;
    DG_CHECK(dg_stamps_resolve(ctx, &walk_ms, &sgns_ms));
    bool retry = false;
    const int rc = dg_walk_graph_collect(ctx, walk_ms, step_ms, rs, &retry);
    if (rc == DG_OK && retry) continue;  // nothing changed but the grown token buffer
    // an undefined relative entropy (unvisited vertices) still completes the
    // round and its training, as dg_walk_round completes it; the other errors
    // (the length limit) left the round empty
    if (rc != DG_OK && rc != DG_ERR_RE_UNDEFINED) return rc;
    DG_CHECK(dg_sgns_graph_collect(ctx, true, sgns_ms, ss));
    return rc;
  }
  return dg_fail(ctx, DG_ERR_STATE, "token buffer sizing did not converge");
}

extern "C" int dg_pipeline_info(dg_ctx* ctx, uint64_t* instantiations, uint64_t* updates,
                                uint64_t* launches) {
  if (!ctx) return DG_ERR_ARG;
  const dg_pipe_state* ps = ctx->pipe;
  if (instantiations) *instantiations = ps ? ps->instantiations : 0;
  if (updates) *updates = ps ? ps->updates : 0;
  if (launches) *launches = ps ? ps->launches : 0;
  return DG_OK;
}
