// myevd: command-line driver with the reference's interface
// (`myEVD <n> <b> <nb>`, src/EVD/myEVD.cu:251-266): builds a seeded symmetric
// matrix on the GPU, runs the full eigen-decomposition through the C ABI and
// prints per-stage times, TFLOP/s and the accuracy metrics.
//
//   myevd n b nb [seed] [--values-only]
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "evd4sc.h"

#define CK(x)                                                                              \
  do {                                                                                     \
    int s_ = (x);                                                                          \
    if (s_) {                                                                              \
      std::fprintf(stderr, "%s failed (%d): %s\n", #x, s_, evd_ctx_last_error(ctx));       \
      return 1;                                                                            \
    }                                                                                      \
  } while (0)

int main(int argc, char **argv) {
  if (argc < 4) {
    std::printf("Usage: myevd <n> <b> <nb> [seed] [--values-only]\n");
    return 0;
  }
  const long n = std::atol(argv[1]), b = std::atol(argv[2]), nb = std::atol(argv[3]);
  unsigned long long seed = 0;
  bool values_only = false;
  for (int i = 4; i < argc; i++) {
    if (!std::strcmp(argv[i], "--values-only"))
      values_only = true;
    else
      seed = std::strtoull(argv[i], nullptr, 10);
  }
  const char jobz = values_only ? 'N' : 'V';
  std::printf("myevd: n=%ld, b=%ld, nb=%ld, seed=%llu, jobz=%c\n", n, b, nb, seed, jobz);
  evd_ctx *ctx = nullptr;
  if (evd_ctx_create(0, nullptr, &ctx) != 0) {
    std::fprintf(stderr, "no usable GPU\n");
    return 1;
  }
  float *A0 = nullptr, *A = nullptr, *Z = nullptr;
  double *w = nullptr;
  if (hipMalloc(&A0, sizeof(float) * n * n) != hipSuccess || hipMalloc(&A, sizeof(float) * n * n) != hipSuccess ||
      hipMalloc(&w, sizeof(double) * n) != hipSuccess ||
      (!values_only && hipMalloc(&Z, sizeof(float) * n * n) != hipSuccess)) {
    std::fprintf(stderr, "hipMalloc failed\n");
    return 1;
  }
  CK(evd_gen_symmetric_f32(ctx, n, seed, 0, A0, n));
  CK(evd_ctx_reserve(ctx, evd_syevd_workspace_f32(n, b, nb, jobz)));
  CK(evd_ctx_enable_timing(ctx, 1));
  hipStream_t st = static_cast<hipStream_t>(evd_ctx_get_stream(ctx));
  hipMemcpyAsync(A, A0, sizeof(float) * n * n, hipMemcpyDeviceToDevice, st);
  CK(evd_ctx_synchronize(ctx));
  auto t0 = std::chrono::high_resolution_clock::now();
  CK(evd_syevd_f32(ctx, jobz, n, b, nb, A, n, w, Z, n));
  CK(evd_ctx_synchronize(ctx));
  auto t1 = std::chrono::high_resolution_clock::now();
  const double sec = std::chrono::duration<double>(t1 - t0).count();
  float ms[EVD_STAGE_COUNT];
  evd_ctx_stage_ms(ctx, ms, EVD_STAGE_COUNT);
  const double n3 = (double)n * n * n;
  const char *names[EVD_STAGE_COUNT] = {"sy2sb", "Bugle chasing", "DC (GPU)", "BC back", "SBR back"};
  const double fl[EVD_STAGE_COUNT] = {4.0 / 3.0 * n3, 0.0, 0.0, 2.0 * n3, 2.0 * n3};
  for (int i = 0; i < EVD_STAGE_COUNT; i++) {
    if (fl[i] > 0 && ms[i] > 0)
      std::printf("%s %ldx%ld takes %f ms, tflops is %f\n", names[i], n, n, ms[i], fl[i] / (ms[i] * 1e9));
    else
      std::printf("%s %ldx%ld takes %f ms\n", names[i], n, n, ms[i]);
  }
  const double sy2tr = ms[0] + ms[1];
  std::printf("sy2tr %ldx%ld takes %f ms, tflops is %f\n", n, n, sy2tr, (4.0 * n3 / 3.0) / (sy2tr * 1e9));
  std::printf("EVD total %ldx%ld takes %f ms, tflops is %f (10/3 n^3)\n", n, n, sec * 1e3, (10.0 / 3.0) * n3 / sec / 1e12);
  if (!values_only) {
    double orth = 0, be = 0;
    CK(evd_orthogonality_f32(ctx, n, n, Z, n, &orth));
    CK(evd_backward_error_f32(ctx, n, A0, n, w, Z, n, &be));
    std::printf("The Orthogonality metrix is %g\n", orth);
    std::printf("Backforward err: %g\n", be);
  }
  std::vector<double> hw(n);
  hipMemcpy(hw.data(), w, sizeof(double) * n, hipMemcpyDeviceToHost);
  std::printf("eigenvalues: min %.10g max %.10g\n", hw[0], hw[n - 1]);
  hipFree(A0);
  hipFree(A);
  hipFree(w);
  if (Z) hipFree(Z);
  evd_ctx_destroy(ctx);
  return 0;
}
