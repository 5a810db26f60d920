// Captured solves: the whole eigensolver of one (n, b, nb, jobz) on fixed
// buffers recorded once as a hipGraph and replayed.  The reference launches
// every kernel from the host on every call (src/EVD/myEVD.cu:251-706); for
// small orders that launch stream is the cost, and a replay submits it as one
// unit.  The solve makes no host round trip between its kernels (merge plans
// and occupancy are resolved by the eager solve that precedes the capture,
// launch epochs of the panel exchange are kept on the device), so the
// captured work is the eager call's work exactly.
#include "common.hpp"

struct evd_graph {
  evd_ctx *ctx = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  char *desc = nullptr;  // device descriptors the graph reads (mixed-size graphs)
};

using namespace evd;

namespace {

// Capture `solve` (which enqueues the work on ctx->stream) as a graph.  The
// capture runs on a private non-blocking stream (the context's stream may be
// the null stream, which cannot be captured); replays go to the context's
// stream.
template <typename F>
int capture(evd_ctx *ctx, F solve, evd_graph **out) {
  hipStream_t cap = nullptr;
  EVD_HIP(ctx, hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
  hipStream_t saved_stream = ctx->stream;
  const bool saved_timing = ctx->timing;
  ctx->stream = cap;
  ctx->timing = false;
  ctx->capturing = true;
  int status = 0;
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeRelaxed);
  if (e != hipSuccess) {
    status = set_error(ctx, EVD_E_HIP, "hipStreamBeginCapture failed: %s", hipGetErrorString(e));
  } else {
    status = solve();
    e = hipStreamEndCapture(cap, &graph);
    if (status == 0 && e != hipSuccess)
      status = set_error(ctx, EVD_E_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
  }
  ctx->capturing = false;
  ctx->timing = saved_timing;
  ctx->stream = saved_stream;
  hipStreamDestroy(cap);
  hipGetLastError();  // a failed capture leaves its error here; it is reported above
  hipGraphExec_t exec = nullptr;
  if (status == 0) {
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) status = set_error(ctx, EVD_E_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
  }
  if (status != 0) {
    if (graph) hipGraphDestroy(graph);
    return status;
  }
  evd_graph *g = new evd_graph;
  g->ctx = ctx;
  g->graph = graph;
  g->exec = exec;
  ctx->graphs++;
  *out = g;
  return 0;
}

}  // namespace

extern "C" int evd_graph_create(evd_ctx *ctx, char jobz, long n, long b, long nb, float *A, long lda, double *w,
                                float *Z, long ldz, evd_graph **out) {
  if (!ctx) return -1;
  EVD_ARG(ctx, out != nullptr, 11);
  EVD_ARG(ctx, n >= 1, 3);
  *out = nullptr;
  // eager solve: checks the arguments, sizes the arena, builds the merge plan
  EVD_TRY(evd_syevd_f32(ctx, jobz, n, b, nb, A, lda, w, Z, ldz));
  return capture(ctx, [&] { return syevd_call(ctx, jobz, n, b, nb, A, lda, w, Z, ldz); }, out);
}

extern "C" int evd_graph_create_batched(evd_ctx *ctx, char jobz, int count, long n, long b, long nb, float *A,
                                        long lda, long a_stride, double *w, long w_stride, float *Z, long ldz,
                                        long z_stride, evd_graph **out) {
  if (!ctx) return -1;
  EVD_ARG(ctx, out != nullptr, 15);
  EVD_ARG(ctx, count >= 1, 3);
  EVD_ARG(ctx, n >= 1, 4);
  *out = nullptr;
  EVD_TRY(evd_syevd_batched_f32(ctx, jobz, count, n, b, nb, A, lda, a_stride, w, w_stride, Z, ldz, z_stride));
  return capture(
      ctx,
      [&] {
        return syevd_batched_call(ctx, jobz, count, n, b, nb, A, lda, a_stride, w, w_stride, Z, ldz, z_stride);
      },
      out);
}

extern "C" int evd_graph_create_mixed(evd_ctx *ctx, int count, const long *n, float *const *A, const long *lda,
                                      double *const *w, float *const *Z, const long *ldz, const char *jobz, long b,
                                      evd_graph **out) {
  if (!ctx) return -1;
  EVD_ARG(ctx, out != nullptr, 11);
  EVD_ARG(ctx, count >= 1, 2);
  *out = nullptr;
  // eager solve: checks the arguments, sizes the arena, builds every
  // bucket's merge plan
  EVD_TRY(evd_syevd_mixed_f32(ctx, count, n, A, lda, w, Z, ldz, jobz, b));
  const MixedArgs a{count, n, A, lda, w, Z, ldz, jobz, b};
  // the descriptors live as long as the graph, outside the arena
  const size_t bytes = mixed_desc_bytes(a);
  char *desc = nullptr;
  if (bytes) {
    hipError_t e = hipMalloc(&desc, bytes);
    if (e != hipSuccess) return set_error(ctx, EVD_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  int s = mixed_upload(ctx, a, desc);
  if (s == 0) s = capture(ctx, [&] { return mixed_enqueue(ctx, a, desc); }, out);
  if (s != 0) {
    if (desc) hipFree(desc);
    return s;
  }
  (*out)->desc = desc;
  return 0;
}

extern "C" int evd_graph_launch(evd_graph *g) {
  if (!g) return -1;
  EVD_HIP(g->ctx, hipGraphLaunch(g->exec, g->ctx->stream));
  return 0;
}

extern "C" int evd_graph_wait(evd_graph *g) {
  if (!g) return -1;
  return check_device_error(g->ctx, "evd_graph_launch");
}

extern "C" int evd_graph_destroy(evd_graph *g) {
  if (!g) return 0;
  hipStreamSynchronize(g->ctx->stream);
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->graph) hipGraphDestroy(g->graph);
  if (g->desc) hipFree(g->desc);
  g->ctx->graphs--;
  delete g;
  return 0;
}
