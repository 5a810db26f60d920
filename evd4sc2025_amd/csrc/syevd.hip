// Whole pipeline SBR -> bulge chasing -> tridiagonal solve -> back
// transformation (the stage sequence of main in src/EVD/myEVD.cu:251-571),
// plus the accuracy metrics of util/tool/checkMetric.cu:5-25 and
// src/EVD/myEVD.cu:598-639 computed in FP64 on the GPU.
//
// Differences from the reference driver, by design:
//  * the tridiagonal stage runs on the GPU (FP64 divide and conquer) instead
//    of host MKL, for eigenvalues-only requests too (their tridiagonal
//    eigenvectors stay internal: the D&C's merges need them, and it runs
//    every eigenvalue in parallel where a Sturm-count bisection is a serial
//    chain of n steps per eigenvalue);
//  * eigenvectors are formed as Z = Q_SBR (Q_BC Z_T) by applying the
//    reflectors to Z_T directly (4 n^3 flops) instead of forming Q and a
//    final GEMM (6 n^3 in the reference);
//  * every buffer comes from the zero-initialised context arena.
#include <cfloat>

#include "common.hpp"

namespace evd {

int sy2sb(evd_ctx *ctx, long n, long b, long nb, float *A, long lda, float *W, long ldw, float *Y, long ldy,
          long wcols);
size_t sy2sb_workspace(long n, long b, long nb);
int band_pack(evd_ctx *ctx, long n, long b, const float *A, long lda, double *band, long ldband);
int sb2st(evd_ctx *ctx, long n, long b, double *band, long ldband, float *U, long ldu, double *D, double *E);
int stedc(evd_ctx *ctx, long n, double *D, double *E, double *Z, long ldz);
size_t stedc_workspace(long n);
int stedc_batched(evd_ctx *ctx, int count, long n, double *D, double *E, double *Z, long ldz, long zs);
size_t stedc_batched_workspace(int count, long n);
size_t bc_back_pack_bytes(long n, long b);
size_t bc_back_wy_bytes(long n);
long sbr_back_blk(long knob, long b, long wcols);
bool tune_lookup(long n, long b, int precision, int *nb, Tuning *t);
int sb2st_batched(evd_ctx *ctx, int count, long n, long b, double *band, long ldband, long band_stride, float *U,
                  long ldu, long u_stride, double *D, long d_stride, double *E, long e_stride, int *ext_com);
size_t sb2st_batched_workspace(int count, long n);
bool bc_back_wy_fits(long n, long ldq);
int bc_back_wy_batched(evd_ctx *ctx, int count, long n, long ncols, float *Q, long ldq, long q_stride, const float *U,
                       long ldu, long u_stride, bool split16);
size_t bc_back_wy64_bytes(long n);
bool bc_back_wy64_fits(long n, long ldq);
int bc_back_wy64_batched(evd_ctx *ctx, int count, long n, long ncols, float *Q, long ldq, long q_stride,
                         const float *U, long ldu, long u_stride);
int bc_back(evd_ctx *ctx, long n, long b, long ncols, float *Q, long ldq, const float *U, long ldu, bool rev);
int bc_back_small_batched(evd_ctx *ctx, int count, long n, long b, long ncols, float *Q, long ldq, long q_stride,
                          const float *U, long ldu, long u_stride, bool rev);
int sbr_back(evd_ctx *ctx, long n, long b, long ncols, float *Z, long ldz, float *W, long ldw, const float *Y,
             long ldy, long wcols);
size_t sy2sb_batched_workspace(int count, long n, long b, long nb);
int sy2sb_batched(evd_ctx *ctx, int count, long n, long b, long nb, float *A, long lda, long sA, float *W, long ldw,
                  long sW, float *Y, long ldy, long sY, long wcols);
int sbr_back_batched(evd_ctx *ctx, int count, long n, long b, long ncols, float *Z, long ldz, long sZ, float *W,
                     long ldw, const float *Y, long ldy, long wcols);

static long wy_cols(long n) { return (n + 1023) / 1024 * 1024; }

// ------------------------------------------------------------- helpers
__global__ void f64_to_f32_kernel(long rows, long cols, const double *__restrict__ s, long lds, float *__restrict__ d,
                                  long ldd) {
  for (long c = blockIdx.y; c < cols; c += gridDim.y)
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < rows; r += (long)gridDim.x * blockDim.x)
      d[r + c * ldd] = (float)s[r + c * lds];
}
__global__ void f32_to_f64_kernel(long rows, long cols, const float *__restrict__ s, long lds, double *__restrict__ d,
                                  long ldd) {
  for (long c = blockIdx.y; c < cols; c += gridDim.y)
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < rows; r += (long)gridDim.x * blockDim.x)
      d[r + c * ldd] = (double)s[r + c * lds];
}
static dim3 grid2d(long rows, long cols) {
  return dim3((unsigned)std::min<long>((rows + 255) / 256, 64), (unsigned)std::min<long>(cols, 8192));
}

// Symmetric fill of the upper triangle from the lower one + non-finite
// screen + max |a_ij| (as the bits of a non-negative float, which order like
// unsigned integers) into *amax.  A non-finite entry raises EVD_E_NONFINITE
// in the error word (read when the call ends) and is replaced by 0, so the
// rest of the solve runs on finite data without a host round trip after the
// screen.
// Entries at rows (and columns) >= nmax do not count toward the maximum.
__global__ void symmetrize_check_kernel(long n, float *__restrict__ A, long lda, unsigned *__restrict__ err,
                                        unsigned *__restrict__ amax, long nmax) {
  unsigned mx = 0;
  for (long c = blockIdx.y; c < n; c += gridDim.y)
    for (long r = c + blockIdx.x * (long)blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x) {
      float v = A[r + c * lda];
      if (!isfinite(v)) {
        atomicOr(err, (unsigned)EVD_E_NONFINITE);
        v = 0.f;
        A[r + c * lda] = v;
      }
      A[c + r * lda] = v;
      if (r < nmax) mx = max(mx, __float_as_uint(fabsf(v)));
    }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = max(mx, (unsigned)__shfl_xor((int)mx, o, 64));
  // one atomic per wave at most, and only while it can still raise the
  // maximum (a stale read only lets an unneeded atomic through)
  if ((threadIdx.x & 63) == 0 && mx > __hip_atomic_load(amax, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    atomicMax(amax, mx);
}

// Scaling into the safe range (the step of LAPACK ssyevd): with
// anrm = max |a_ij|, a matrix with 0 < anrm < rmin is scaled by rmin / anrm and
// one with anrm > rmax by rmax / anrm (rmin = sqrt(smlnum), rmax =
// sqrt(1 / smlnum), smlnum = 2^-126 / 2^-23), so that no FP32 product of the
// solve underflows or overflows; the eigenvalues are divided by the factor at
// the end and the eigenvectors need nothing.  Problem blockIdx.z (matrices at
// a_stride); sigma[q] receives the factor (1 when no scaling is needed).
//
// The split-FP16 form (unit16 = true) instead scales every nonzero matrix by
// the power of two that brings max |a_ij| into [0.5, 1): FP16 holds 5
// exponent bits, so an unscaled matrix with small entries would put the lo
// halves of the split operands into FP16 subnormals (max |a_ij| = 1e-3 keeps
// only ~3e-5 relative accuracy) and a large one would overflow the products
// (|A W| reaches ~0.6 sqrt(n) max |a_ij|, above 65504 for max |a_ij| = 1e3
// at n = 16384).  A power of two changes no mantissa bit.
__device__ inline float safe_scale(unsigned amax_bits, bool unit16) {
  const float anrm = __uint_as_float(amax_bits);
  if (unit16) {
    if (!(anrm > 0.f)) return 1.f;
    int e = 0;
    frexpf(anrm, &e);  // anrm = f 2^e, f in [0.5, 1)
    return ldexpf(1.f, min(-e, 126));  // subnormal anrm: 2^126 still lifts it above 2^-23
  }
  const float rmin = 3.14018491736755e-16f;  // 2^-51.5
  const float rmax = 3.18452583626289e+15f;  // 2^51.5
  if (anrm > 0.f && anrm < rmin) return rmin / anrm;
  if (anrm > rmax) return rmax / anrm;
  return 1.f;
}

__global__ void scale_input_kernel(long n, float *__restrict__ A, long lda, long a_stride,
                                   const unsigned *__restrict__ amax, double *__restrict__ sigma, int unit16) {
  const int q = blockIdx.z;
  const float sg = safe_scale(amax[q], unit16 != 0);
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) sigma[q] = (double)sg;
  if (sg == 1.f) return;
  float *Aq = A + q * a_stride;
  for (long c = blockIdx.y; c < n; c += gridDim.y)
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x)
      Aq[r + c * lda] *= sg;
}

__global__ void unscale_values_kernel(long n, double *__restrict__ w, long w_stride, const double *__restrict__ sigma) {
  const int q = blockIdx.y;
  const double sg = sigma[q];
  if (sg == 1.0) return;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    w[q * w_stride + i] /= sg;
}

// screen, symmetrise and, where needed, scale `count` matrices (A at a_stride)
static int prepare_inputs(evd_ctx *ctx, int count, long n, float *A, long lda, long a_stride, double *sigma,
                          unsigned *amax) {
  EVD_TRY(fill_u32(ctx, amax, (size_t)count, 0u));
  for (int q = 0; q < count; q++) {
    const long nmax = ctx->amax_rows ? std::min(n, ctx->amax_rows[q]) : n;
    symmetrize_check_kernel<<<grid2d(n, n), 256, 0, ctx->stream>>>(n, A + q * a_stride, lda, ctx->ctrl + kCtrlError,
                                                                     amax + q, nmax);
    EVD_LAUNCHED(ctx);
  }
  const dim3 g = grid2d(n, n);
  scale_input_kernel<<<dim3(g.x, std::min(g.y, 1024u), (unsigned)count), 256, 0, ctx->stream>>>(
      n, A, lda, a_stride, amax, sigma, ctx->precision == EVD_PREC_FP16 ? 1 : 0);
  EVD_LAUNCHED(ctx);
  return 0;
}

static int unscale_values(evd_ctx *ctx, int count, long n, double *w, long w_stride, const double *sigma) {
  unscale_values_kernel<<<dim3((unsigned)std::min<long>((n + 255) / 256, 64), (unsigned)count), 256, 0,
                          ctx->stream>>>(n, w, w_stride, sigma);
  EVD_LAUNCHED(ctx);
  return 0;
}

__global__ void copy_f64_kernel(long n, const double *__restrict__ s, double *__restrict__ d) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) d[i] = s[i];
}

size_t syevd_workspace(long n, long b, long nb, char jobz, int precision, long sbr_knob) {
  const long wc = wy_cols(n);
  const long blk = sbr_back_blk(sbr_knob, b, wc);
  size_t s = 0;
  s += align256((size_t)n * wc * sizeof(float)) * 2;          // W, Y
  s += align256((size_t)2 * b * n * sizeof(double));          // band
  s += align256((size_t)(n + b) * n * sizeof(float));         // U
  s += align256((size_t)n * sizeof(double)) * 2;              // D, E
  size_t later = sy2sb_workspace(n, b, nb);
  // the tridiagonal stage is the D&C for both jobz (eigenvalues only: its
  // vectors are internal)
  s += align256((size_t)n * n * sizeof(double));  // Z_T
  later = std::max(later, stedc_workspace(n));
  if (jobz == 'V' || jobz == 'v') {
    size_t dc = stedc_workspace(n);
    size_t back = align256((size_t)blk * n * sizeof(float)) + align256((size_t)(wc / 2) * blk * sizeof(float)) +
                  ((size_t)64 << 20);
    back = std::max(back, bc_back_pack_bytes(n, b) + ((size_t)1 << 20));
    if (b == 32) back = std::max(back, bc_back_wy_bytes(n) + ((size_t)1 << 20));
    if (b == 64) back = std::max(back, bc_back_wy64_bytes(n) + ((size_t)1 << 20));
    later = std::max(later, std::max(dc, back));
  }
  return s + later + ((size_t)4 << 20);
}

static void stage_mark(evd_ctx *ctx, int i) {
  if (ctx->timing && !ctx->capturing) hipEventRecord(ctx->ev[i], ctx->stream);
}

int syevd(evd_ctx *ctx, char jobz, long n, long b, long nb, float *A, long lda, double *w, float *Z, long ldz) {
  const bool vec = (jobz == 'V' || jobz == 'v');
  if (n == 0) return 0;
  ArenaFrame frame(ctx);
  const long wc = wy_cols(n);
  float *W = arena_alloc_n<float>(ctx, (size_t)n * wc);
  float *Y = arena_alloc_n<float>(ctx, (size_t)n * wc);
  double *band = arena_alloc_n<double>(ctx, (size_t)2 * b * n);
  const long ldu = n + b;
  float *U = arena_alloc_n<float>(ctx, (size_t)ldu * n);
  double *D = arena_alloc_n<double>(ctx, (size_t)n);
  double *E = arena_alloc_n<double>(ctx, (size_t)std::max<long>(n, 1));
  double *Zt = arena_alloc_n<double>(ctx, (size_t)n * n);
  double *sigma = arena_alloc_n<double>(ctx, 1);
  unsigned *amax = arena_alloc_n<unsigned>(ctx, 1);
  if (!W || !Y || !band || !U || !D || !E || !Zt || !sigma || !amax) return EVD_E_NOMEM;

  // full symmetric storage from the lower triangle, NaN/Inf screen, scaling
  EVD_TRY(prepare_inputs(ctx, 1, n, A, lda, 0, sigma, amax));

  stage_mark(ctx, 0);
  EVD_TRY(sy2sb(ctx, n, b, nb, A, lda, W, n, Y, n, wc));
  EVD_TRY(band_pack(ctx, n, b, A, lda, band, 2 * b));
  stage_mark(ctx, 1);
  EVD_TRY(sb2st(ctx, n, b, band, 2 * b, U, ldu, D, E));
  stage_mark(ctx, 2);
  EVD_TRY(stedc(ctx, n, D, E, Zt, n));
  copy_f64_kernel<<<(unsigned)std::min<long>((n + 255) / 256, 1024), 256, 0, ctx->stream>>>(n, D, w);
  EVD_LAUNCHED(ctx);
  EVD_TRY(unscale_values(ctx, 1, n, w, 0, sigma));
  if (vec) {
    f64_to_f32_kernel<<<grid2d(n, n), 256, 0, ctx->stream>>>(n, n, Zt, n, Z, ldz);
    EVD_LAUNCHED(ctx);
    stage_mark(ctx, 3);
    EVD_TRY(bc_back(ctx, n, b, n, Z, ldz, U, ldu, true));
    stage_mark(ctx, 4);
    EVD_TRY(sbr_back(ctx, n, b, n, Z, ldz, W, n, Y, n, wc));
    stage_mark(ctx, 5);
  } else {
    stage_mark(ctx, 3);
    stage_mark(ctx, 4);
    stage_mark(ctx, 5);
  }
  if (ctx->capturing) return 0;  // the graph's owner reads the error word after a replay
  EVD_TRY(check_device_error(ctx, "evd_syevd_f32"));
  if (ctx->timing) {
    for (int i = 0; i < EVD_STAGE_COUNT; i++) {
      float ms = 0.f;
      hipEventElapsedTime(&ms, ctx->ev[i], ctx->ev[i + 1]);
      ctx->stage_ms[i] = ms;
    }
    float km = 0.f, cm = 0.f;
    if (vec) hipEventElapsedTime(&km, ctx->kev[0], ctx->kev[1]);
    ctx->stage_ms[EVD_TIMER_BCBACK_KERNEL] = km;
    if (n > 2) hipEventElapsedTime(&cm, ctx->cev[0], ctx->cev[1]);
    ctx->stage_ms[EVD_TIMER_CHASE_KERNEL] = cm;
  }
  return 0;
}

// ------------------------------------------------------------- batched
// `count` same-size problems (A + q a_stride, w + q w_stride, Z + q z_stride).
// The bulge chase of all problems is one persistent launch (tickets
// interleave the problems' sweeps) and so is the BC back transformation (one
// WY prep and one WY kernel launch, problem = grid y); SBR, the D&C and the
// SBR back transformation run per problem in order on the context stream.
size_t syevd_batched_workspace(int count, long n, long b, long nb, char jobz, int precision, long sbr_knob) {
  const long wc = wy_cols(n);
  const long blk = sbr_back_blk(sbr_knob, b, wc);
  const bool vec = (jobz == 'V' || jobz == 'v');
  size_t per = align256((size_t)n * wc * sizeof(float)) * 2 + align256((size_t)2 * b * n * sizeof(double)) +
               align256((size_t)(n + b) * n * sizeof(float)) + align256((size_t)n * sizeof(double)) * 2;
  size_t s = per * count + 4 * 256;
  size_t later = std::max(sy2sb_batched_workspace(count, n, b, nb), sb2st_batched_workspace(count, n) + ((size_t)1 << 20));
  s += align256((size_t)n * n * count * sizeof(double));  // tridiagonal eigenvectors (internal for jobz = N)
  later = std::max(later, stedc_batched_workspace(count, n) + ((size_t)1 << 20));
  if (vec) {
    size_t back = syevd_workspace(n, b, nb, jobz, precision, sbr_knob);  // per-problem back transformation and D&C bound
    // batched SBR back transformation: merge products, block products, partials
    back = std::max(back, sy2sb_batched_workspace(count, n, b, nb) +
                              (align256((size_t)(wc / 2) * blk * sizeof(float)) +
                               align256((size_t)blk * n * sizeof(float))) * (size_t)count);
    if (b == 32) back = std::max(back, bc_back_wy_bytes(n) * count + ((size_t)1 << 20));
    if (b == 64) back = std::max(back, bc_back_wy64_bytes(n) * count + ((size_t)1 << 20));
    later = std::max(later, back);
  }
  return s + later + ((size_t)4 << 20);
}

int syevd_batched(evd_ctx *ctx, char jobz, int count, long n, long b, long nb, float *A, long lda, long a_stride,
                  double *w, long w_stride, float *Z, long ldz, long z_stride) {
  const bool vec = (jobz == 'V' || jobz == 'v');
  if (n == 0 || count == 0) return 0;
  ArenaFrame frame(ctx);
  const long wc = wy_cols(n);
  // W, Y of all problems contiguous (problem q at q n wc)
  const long sWY = n * wc;
  float *W = arena_alloc_n<float>(ctx, (size_t)sWY * count);
  float *Y = arena_alloc_n<float>(ctx, (size_t)sWY * count);
  if (!W || !Y) return EVD_E_NOMEM;
  const long ldband = 2 * b, band_stride = ldband * n, ldu = n + b, u_stride = ldu * n;
  const long e_stride = std::max<long>(n, 1);
  double *band = arena_alloc_n<double>(ctx, (size_t)band_stride * count);
  float *U = arena_alloc_n<float>(ctx, (size_t)u_stride * count);
  double *D = arena_alloc_n<double>(ctx, (size_t)n * count);
  double *E = arena_alloc_n<double>(ctx, (size_t)e_stride * count);
  double *Zt = arena_alloc_n<double>(ctx, (size_t)n * n * count);
  double *sigma = arena_alloc_n<double>(ctx, (size_t)count);
  unsigned *amax = arena_alloc_n<unsigned>(ctx, (size_t)count);
  if (!band || !U || !D || !E || !Zt || !sigma || !amax) return EVD_E_NOMEM;

  EVD_TRY(prepare_inputs(ctx, count, n, A, lda, a_stride, sigma, amax));

  stage_mark(ctx, 0);
  EVD_TRY(sy2sb_batched(ctx, count, n, b, nb, A, lda, a_stride, W, n, sWY, Y, n, sWY, wc));
  for (int q = 0; q < count; q++) EVD_TRY(band_pack(ctx, n, b, A + q * a_stride, lda, band + q * band_stride, ldband));
  stage_mark(ctx, 1);
  EVD_TRY(sb2st_batched(ctx, count, n, b, band, ldband, band_stride, U, ldu, u_stride, D, n, E, e_stride, nullptr));
  stage_mark(ctx, 2);
  EVD_TRY(stedc_batched(ctx, count, n, D, E, Zt, n, n * n));
  for (int q = 0; q < count; q++) {
    copy_f64_kernel<<<(unsigned)std::min<long>((n + 255) / 256, 1024), 256, 0, ctx->stream>>>(n, D + q * n,
                                                                                             w + q * w_stride);
    EVD_LAUNCHED(ctx);
  }
  EVD_TRY(unscale_values(ctx, count, n, w, w_stride, sigma));
  if (vec) {
    for (int q = 0; q < count; q++) {
      f64_to_f32_kernel<<<grid2d(n, n), 256, 0, ctx->stream>>>(n, n, Zt + q * n * n, n, Z + q * z_stride, ldz);
      EVD_LAUNCHED(ctx);
    }
    stage_mark(ctx, 3);
    if (b != 32 && b != 64) {
      EVD_TRY(bc_back_small_batched(ctx, count, n, b, n, Z, ldz, z_stride, U, ldu, u_stride, true));
    } else if (b == 32 && !ctx->tune.bc_valu && bc_back_wy_fits(n, ldz)) {
      EVD_TRY(bc_back_wy_batched(ctx, count, n, n, Z, ldz, z_stride, U, ldu, u_stride,
                                 ctx->precision == EVD_PREC_FP16));
    } else if (b == 64 && !ctx->tune.bc_valu && bc_back_wy64_fits(n, ldz)) {
      EVD_TRY(bc_back_wy64_batched(ctx, count, n, n, Z, ldz, z_stride, U, ldu, u_stride));
    } else {
      for (int q = 0; q < count; q++)
        EVD_TRY(bc_back(ctx, n, b, n, Z + q * z_stride, ldz, U + q * u_stride, ldu, true));
    }
    stage_mark(ctx, 4);
    EVD_TRY(sbr_back_batched(ctx, count, n, b, n, Z, ldz, z_stride, W, n, Y, n, wc));
    stage_mark(ctx, 5);
  } else {
    stage_mark(ctx, 3);
    stage_mark(ctx, 4);
    stage_mark(ctx, 5);
  }
  if (ctx->capturing) return 0;  // the graph's owner reads the error word after a replay
  EVD_TRY(check_device_error(ctx, "evd_syevd_batched_f32"));
  if (ctx->timing) {
    for (int i = 0; i < EVD_STAGE_COUNT; i++) {
      float ms = 0.f;
      hipEventElapsedTime(&ms, ctx->ev[i], ctx->ev[i + 1]);
      ctx->stage_ms[i] = ms;
    }
    float km = 0.f, cm = 0.f;
    if (vec) hipEventElapsedTime(&km, ctx->kev[0], ctx->kev[1]);
    ctx->stage_ms[EVD_TIMER_BCBACK_KERNEL] = km;
    if (n > 2) hipEventElapsedTime(&cm, ctx->cev[0], ctx->cev[1]);
    ctx->stage_ms[EVD_TIMER_CHASE_KERNEL] = cm;
  }
  return 0;
}

// ------------------------------------------------------------- metrics
__global__ void frob_sq_kernel(long m, long n, const double *__restrict__ X, long ldx, int minus_identity,
                               double *__restrict__ out) {
  __shared__ double red[4];
  double s = 0.0;
  for (long c = blockIdx.y; c < n; c += gridDim.y)
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < m; r += (long)gridDim.x * blockDim.x) {
      double v = X[r + c * ldx];
      if (minus_identity && r == c) v -= 1.0;
      s += v * v;
    }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(out, red[0] + red[1] + red[2] + red[3]);
}

__global__ void scale_cols_sub_kernel(long n, double *__restrict__ R, long ldr, const double *__restrict__ Z, long ldz,
                                      const double *__restrict__ w) {
  for (long c = blockIdx.y; c < n; c += gridDim.y)
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x)
      R[r + c * ldr] -= Z[r + c * ldz] * w[c];
}

static int frob_sq(evd_ctx *ctx, long m, long n, const double *X, long ldx, bool minus_identity, double *dacc,
                   double *host) {
  double zero = 0.0;
  EVD_HIP(ctx, hipMemcpyAsync(dacc, &zero, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  frob_sq_kernel<<<grid2d(m, n), 256, 0, ctx->stream>>>(m, n, X, ldx, minus_identity ? 1 : 0, dacc);
  EVD_LAUNCHED(ctx);
  EVD_HIP(ctx, hipMemcpyAsync(host, dacc, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  EVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

}  // namespace evd

using namespace evd;

extern "C" size_t evd_syevd_workspace_f32(long n, long b, long nb, char jobz) {
  if (n <= 0) return 0;
  return syevd_workspace(n, b, nb, jobz, EVD_PREC_FP32);
}

extern "C" size_t evd_syevd_workspace_ex(long n, long b, long nb, char jobz, int precision) {
  if (n <= 0) return 0;
  return syevd_workspace(n, b, nb, jobz, precision);
}

extern "C" int evd_syevd_f32(evd_ctx *ctx, char jobz, long n, long b, long nb, float *A, long lda, double *w, float *Z,
                             long ldz) {
  if (!ctx) return -1;
  const bool vec = (jobz == 'V' || jobz == 'v');
  EVD_ARG(ctx, vec || jobz == 'N' || jobz == 'n', 2);
  EVD_ARG(ctx, n >= 0, 3);
  // eigenvectors with b other than 32 / 64: orders up to 128 (small-order back transformation)
  EVD_ARG(ctx, b == 32 || b == 64 || (b >= 1 && b <= 64 && (!vec || (b >= 2 && n <= 128))), 4);
  EVD_ARG(ctx, nb == 0 || (nb >= b && nb % b == 0), 5);
  EVD_ARG(ctx, A != nullptr || n == 0, 6);
  EVD_ARG(ctx, lda >= std::max<long>(1, n), 7);
  EVD_ARG(ctx, w != nullptr || n == 0, 8);
  EVD_ARG(ctx, !vec || Z != nullptr || n == 0, 9);
  EVD_ARG(ctx, !vec || ldz >= std::max<long>(1, n), 10);
  if (ctx->arena_top != 0) return set_error(ctx, EVD_E_UNSUPPORTED, "evd_syevd_f32 cannot nest");
  return syevd_call(ctx, jobz, n, b, nb, A, lda, w, Z, ldz);
}

// knobs the caller left at 0 from a tuning-cache entry (chase workgroups
// scale with the problems of a batch)
static void fill_knobs(Tuning *dst, const Tuning &t, int count) {
  if (!dst->chase_groups) dst->chase_groups = t.chase_groups * std::max(1, count);
  if (!dst->bc_valu) dst->bc_valu = t.bc_valu;
  if (!dst->sbr_back_block) dst->sbr_back_block = t.sbr_back_block;
  if (!dst->chase_depth) dst->chase_depth = t.chase_depth;
  if (!dst->panel_rows) dst->panel_rows = t.panel_rows;
  if (!dst->gemm_tile) dst->gemm_tile = t.gemm_tile;
  if (!dst->bc_sweeps) dst->bc_sweeps = t.bc_sweeps;
  if (!dst->bc_waves) dst->bc_waves = t.bc_waves;
}

// The arena bytes syevd_call / syevd_batched_call will ask for on this
// context (the same knob and nb resolution as the calls).
static void resolve_knobs(const evd_ctx *ctx, long n, long b, long *nb, long *sbr_knob) {
  int tnb = 0;
  Tuning t;
  const bool hit = ctx->tune_cache && tune_lookup(n, b, ctx->precision, &tnb, &t);
  *sbr_knob = ctx->tune.sbr_back_block ? ctx->tune.sbr_back_block : (hit ? t.sbr_back_block : 0);
  if (*nb == 0) *nb = (hit && tnb > 0) ? tnb : b * std::max<long>(1, 1024 / b);
}
size_t evd::syevd_call_workspace(const evd_ctx *ctx, long n, long b, long nb, char jobz) {
  long sbr = 0;
  resolve_knobs(ctx, n, b, &nb, &sbr);
  return syevd_workspace(n, b, nb, jobz, ctx->precision, sbr);
}
size_t evd::syevd_batched_call_workspace(const evd_ctx *ctx, int count, long n, long b, long nb, char jobz) {
  long sbr = 0;
  resolve_knobs(ctx, n, b, &nb, &sbr);
  return syevd_batched_workspace(count, n, b, nb, jobz, ctx->precision, sbr);
}

// evd_syevd_f32 after argument checks: tuning knobs, workspace, solve.
int evd::syevd_call(evd_ctx *ctx, char jobz, long n, long b, long nb, float *A, long lda, double *w, float *Z,
                    long ldz) {
  // knobs left at 0 (and nb = 0) come from the tuning cache, else defaults
  const Tuning user = ctx->tune;
  int tnb = 0;
  Tuning t;
  if (ctx->tune_cache && tune_lookup(n, b, ctx->precision, &tnb, &t)) {
    fill_knobs(&ctx->tune, t, 1);
  }
  if (nb == 0) nb = tnb > 0 ? tnb : b * std::max<long>(1, 1024 / b);
  int s = arena_ensure(ctx, syevd_workspace(n, b, nb, jobz, ctx->precision, ctx->tune.sbr_back_block));
  if (s == 0) s = syevd(ctx, jobz, n, b, nb, A, lda, w, Z, ldz);
  ctx->tune = user;
  return s;
}

extern "C" size_t evd_syevd_batched_workspace_f32(int count, long n, long b, long nb, char jobz, int precision) {
  if (n <= 0 || count <= 0) return 0;
  return syevd_batched_workspace(count, n, b, nb == 0 ? b * std::max<long>(1, 1024 / b) : nb, jobz, precision);
}

extern "C" int evd_syevd_batched_f32(evd_ctx *ctx, char jobz, int count, long n, long b, long nb, float *A, long lda,
                                     long a_stride, double *w, long w_stride, float *Z, long ldz, long z_stride) {
  if (!ctx) return -1;
  const bool vec = (jobz == 'V' || jobz == 'v');
  EVD_ARG(ctx, vec || jobz == 'N' || jobz == 'n', 2);
  // count <= 65535: per-problem kernels carry the problem index in grid y / z
  EVD_ARG(ctx, count >= 0 && count <= 65535, 3);
  EVD_ARG(ctx, n >= 0 && (long)count * n < (1L << 30), 4);
  EVD_ARG(ctx, b == 32 || b == 64 || (b >= 1 && b <= 64 && (!vec || (b >= 2 && n <= 128))), 5);
  EVD_ARG(ctx, nb == 0 || (nb >= b && nb % b == 0), 6);
  EVD_ARG(ctx, A != nullptr || n == 0 || count == 0, 7);
  EVD_ARG(ctx, lda >= std::max<long>(1, n), 8);
  EVD_ARG(ctx, count <= 1 || a_stride >= lda * n, 9);
  EVD_ARG(ctx, w != nullptr || n == 0 || count == 0, 10);
  EVD_ARG(ctx, count <= 1 || w_stride >= n, 11);
  EVD_ARG(ctx, !vec || Z != nullptr || n == 0 || count == 0, 12);
  EVD_ARG(ctx, !vec || ldz >= std::max<long>(1, n), 13);
  EVD_ARG(ctx, !vec || count <= 1 || z_stride >= ldz * n, 14);
  if (ctx->arena_top != 0) return set_error(ctx, EVD_E_UNSUPPORTED, "evd_syevd_batched_f32 cannot nest");
  return syevd_batched_call(ctx, jobz, count, n, b, nb, A, lda, a_stride, w, w_stride, Z, ldz, z_stride);
}

// evd_syevd_batched_f32 after argument checks: tuning knobs, workspace, solve.
int evd::syevd_batched_call(evd_ctx *ctx, char jobz, int count, long n, long b, long nb, float *A, long lda,
                            long a_stride, double *w, long w_stride, float *Z, long ldz, long z_stride) {
  const Tuning user = ctx->tune;
  int tnb = 0;
  Tuning t;
  if (ctx->tune_cache && tune_lookup(n, b, ctx->precision, &tnb, &t)) {
    fill_knobs(&ctx->tune, t, count);
  }
  if (nb == 0) nb = tnb > 0 ? tnb : b * std::max<long>(1, 1024 / b);
  int s = arena_ensure(ctx, syevd_batched_workspace(count, n, b, nb, jobz, ctx->precision, ctx->tune.sbr_back_block));
  if (s == 0) s = syevd_batched(ctx, jobz, count, n, b, nb, A, lda, a_stride, w, w_stride, Z, ldz, z_stride);
  ctx->tune = user;
  return s;
}

extern "C" int evd_orthogonality_f32(evd_ctx *ctx, long m, long n, const float *Q, long ldq, double *out) {
  if (!ctx) return -1;
  EVD_ARG(ctx, m >= 0, 2);
  EVD_ARG(ctx, n >= 0, 3);
  EVD_ARG(ctx, ldq >= std::max<long>(1, m), 5);
  EVD_ARG(ctx, out != nullptr, 6);
  if (ctx->arena_top != 0) return set_error(ctx, EVD_E_UNSUPPORTED, "metrics cannot nest");
  EVD_TRY(arena_ensure(ctx, align256((size_t)m * n * 8) + align256((size_t)n * n * 8) + 4096));
  ArenaFrame frame(ctx);
  double *Q64 = arena_alloc_n<double>(ctx, (size_t)m * n);
  double *G = arena_alloc_n<double>(ctx, (size_t)n * n);
  double *acc = arena_alloc_n<double>(ctx, 1);
  if (!Q64 || !G || !acc) return EVD_E_NOMEM;
  f32_to_f64_kernel<<<grid2d(m, n), 256, 0, ctx->stream>>>(m, n, Q, ldq, Q64, m);
  EVD_LAUNCHED(ctx);
  EVD_TRY(dgemm(ctx, 'T', 'N', n, n, m, 1.0, Q64, m, Q64, m, 0.0, G, n));
  double s = 0.0;
  EVD_TRY(frob_sq(ctx, n, n, G, n, true, acc, &s));
  *out = sqrt(s) / (double)(n > 0 ? n : 1);
  return 0;
}

extern "C" int evd_orthogonality_f64(evd_ctx *ctx, long m, long n, const double *Q, long ldq, double *out) {
  if (!ctx) return -1;
  EVD_ARG(ctx, m >= 0, 2);
  EVD_ARG(ctx, n >= 0, 3);
  EVD_ARG(ctx, ldq >= std::max<long>(1, m), 5);
  EVD_ARG(ctx, out != nullptr, 6);
  if (ctx->arena_top != 0) return set_error(ctx, EVD_E_UNSUPPORTED, "metrics cannot nest");
  EVD_TRY(arena_ensure(ctx, align256((size_t)n * n * 8) + 4096));
  ArenaFrame frame(ctx);
  double *G = arena_alloc_n<double>(ctx, (size_t)n * n);
  double *acc = arena_alloc_n<double>(ctx, 1);
  if (!G || !acc) return EVD_E_NOMEM;
  EVD_TRY(dgemm(ctx, 'T', 'N', n, n, m, 1.0, Q, ldq, Q, ldq, 0.0, G, n));
  double s = 0.0;
  EVD_TRY(frob_sq(ctx, n, n, G, n, true, acc, &s));
  *out = sqrt(s) / (double)(n > 0 ? n : 1);
  return 0;
}

extern "C" int evd_backward_error_f32(evd_ctx *ctx, long n, const float *A, long lda, const double *w, const float *Z,
                                      long ldz, double *out) {
  if (!ctx) return -1;
  EVD_ARG(ctx, n >= 0, 2);
  EVD_ARG(ctx, lda >= std::max<long>(1, n), 4);
  EVD_ARG(ctx, ldz >= std::max<long>(1, n), 7);
  EVD_ARG(ctx, out != nullptr, 8);
  if (ctx->arena_top != 0) return set_error(ctx, EVD_E_UNSUPPORTED, "metrics cannot nest");
  EVD_TRY(arena_ensure(ctx, 3 * align256((size_t)n * n * 8) + 4096));
  ArenaFrame frame(ctx);
  double *A64 = arena_alloc_n<double>(ctx, (size_t)n * n);
  double *Z64 = arena_alloc_n<double>(ctx, (size_t)n * n);
  double *R = arena_alloc_n<double>(ctx, (size_t)n * n);
  double *acc = arena_alloc_n<double>(ctx, 1);
  if (!A64 || !Z64 || !R || !acc) return EVD_E_NOMEM;
  f32_to_f64_kernel<<<grid2d(n, n), 256, 0, ctx->stream>>>(n, n, A, lda, A64, n);
  EVD_LAUNCHED(ctx);
  f32_to_f64_kernel<<<grid2d(n, n), 256, 0, ctx->stream>>>(n, n, Z, ldz, Z64, n);
  EVD_LAUNCHED(ctx);
  EVD_TRY(dgemm(ctx, 'N', 'N', n, n, n, 1.0, A64, n, Z64, n, 0.0, R, n));
  scale_cols_sub_kernel<<<grid2d(n, n), 256, 0, ctx->stream>>>(n, R, n, Z64, n, w);
  EVD_LAUNCHED(ctx);
  double rr = 0.0, aa = 0.0;
  EVD_TRY(frob_sq(ctx, n, n, R, n, false, acc, &rr));
  EVD_TRY(frob_sq(ctx, n, n, A64, n, false, acc, &aa));
  *out = (aa > 0.0) ? sqrt(rr) / ((double)n * sqrt(aa)) : sqrt(rr);
  return 0;
}
