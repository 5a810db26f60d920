// evd4sc_api.hpp - C++ drop-ins with the reference's names and parameter
// order (hansheng1001/EVD4SC2025 api/inc/*.h), implemented over the C ABI of
// evd4sc.h.  The cuBLAS / cuSOLVER handle parameters are replaced by one
// evd_ctx*; buffers stay caller-owned device pointers, column-major with
// 64-bit leading dimensions.  Scratch arguments the reference takes (work,
// dZ, dR, dOriA) are accepted and ignored: scratch comes from the context
// arena.  `info` is a HOST pointer (nullable) that receives the status
// (0 = ok, -k = bad argument k, > 0 = runtime failure); the reference never
// reads its device-side info.  T = double (the reference's own instantiation)
// and T = float are provided; another T is a compile-time error, not a silent
// no-op.
#pragma once

#include <cstdio>
#include <type_traits>

#include "evd4sc.h"

namespace evd4sc {

namespace detail {
template <typename T>
constexpr bool is_f32 = std::is_same<T, float>::value;
template <typename T>
constexpr void require_f32_f64() {
  static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value,
                "evd4sc drop-ins are provided for T = float and T = double");
}
inline void set_info(int *info, int s) {
  if (info) *info = s;
}
inline int sy2sb(evd_ctx *c, long n, long b, long nb, float *A, long lda, float *W, long ldw, float *Y, long ldy) {
  return evd_sy2sb_f32(c, n, b, nb, A, lda, W, ldw, Y, ldy);
}
inline int sy2sb(evd_ctx *c, long n, long b, long nb, double *A, long lda, double *W, long ldw, double *Y, long ldy) {
  return evd_sy2sb_f64(c, n, b, nb, A, lda, W, ldw, Y, ldy);
}
inline int panel(evd_ctx *c, long m, long n, float *A, long lda, float *W, long ldw, float *R, long ldr) {
  return evd_panel_qr_f32(c, m, n, A, lda, W, ldw, R, ldr);
}
inline int panel(evd_ctx *c, long m, long n, double *A, long lda, double *W, long ldw, double *R, long ldr) {
  return evd_panel_qr_f64(c, m, n, A, lda, W, ldw, R, ldr);
}
inline int syr2k(evd_ctx *c, long n, long k, float alpha, float *A, long lda, float *B, long ldb, float beta, float *C,
                 long ldc) {
  return evd_syr2k_f32(c, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
inline int syr2k(evd_ctx *c, long n, long k, double alpha, double *A, long lda, double *B, long ldb, double beta,
                 double *C, long ldc) {
  return evd_syr2k_f64(c, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
}  // namespace detail

// api/inc/zy_zy_sy2sb.h:5-25 (my_ZY_ZY_SBR_Vector<T>): full -> band.  dA holds
// the symmetric matrix on entry and the band (lower, bandwidth b) on exit;
// dW, dY receive the panel WY pairs.
template <typename T>
void my_ZY_ZY_SBR_Vector(evd_ctx *ctx, long M, long N, long b, long nb, T *dOriA, long ldOriA, T *dA, long ldA, T *dW,
                         long ldW, T *dY, long ldY, T *dZ, long ldZ, T *dR, long ldR, T *work, int *info) {
  detail::require_f32_f64<T>();
  (void)dOriA, (void)ldOriA, (void)dZ, (void)ldZ, (void)dR, (void)ldR, (void)work;
  if (M != N) return detail::set_info(info, -2);
  detail::set_info(info, detail::sy2sb(ctx, N, b, nb, dA, ldA, dW, ldW, dY, ldY));
}

// api/inc/PanelQR.h:7-19 (panelQR<T>): A (m x n) <- Y (unit lower), W = V T, R.
template <typename T>
void panelQR(evd_ctx *ctx, long m, long n, T *A, long lda, T *W, long ldw, T *R, long ldr, T *work, int *info) {
  detail::require_f32_f64<T>();
  (void)work;
  detail::set_info(info, detail::panel(ctx, m, n, A, lda, W, ldw, R, ldr));
}

// api/inc/TallShinnyQR.h:9-17 (hou_tsqr_panel<T, M, N>): the explicit thin Q
// (m x n) of the panel replaces A, R receives the triangular factor.  `work`
// must hold m*n + n*n elements (the WY pair of the factorisation).  The
// reference returns void; the status is returned here (0 = ok).
template <typename T, long M, long N>
int hou_tsqr_panel(evd_ctx *ctx, long m, long n, T *A, long lda, T *R, long ldr, T *work) {
  detail::require_f32_f64<T>();
  static_assert(N <= 64, "panel width above 64 is not supported");
  T *W = work;
  int s = detail::panel(ctx, m, n, A, lda, W, m, R, ldr);
  if (s) return s;
  if constexpr (detail::is_f32<T>)
    return evd_orgqr_f32(ctx, m, n, A, lda, W, m, work + (size_t)m * n);
  else
    return evd_orgqr_f64(ctx, m, n, A, lda, W, m, work + (size_t)m * n);
}

// api/inc/tc_ozimmu_syr2k.h:6-18: C = alpha (A B^T + B A^T) + beta C; the
// lower triangle is computed on MFMA and mirrored (nb chunking is internal).
template <typename T>
void tc_ozimmu_syr2k(evd_ctx *ctx, long n, long k, T alpha, T *A, long lda, T *B, long ldb, T beta, T *C, long ldc,
                     long nb) {
  detail::require_f32_f64<T>();
  (void)nb;
  detail::syr2k(ctx, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

// api/inc/computerWYFromSlide.h:4-5 (double, as in the reference) and the FP32 form
inline void computerWYFromSlide(evd_ctx *ctx, long M, long N, long slideWidth, double *dW, long ldW, double *dY,
                                long ldY, double *work) {
  (void)work;
  evd_wy_from_slide_f64(ctx, M, N, slideWidth, dW, ldW, dY, ldY);
}
inline void computerWYFromSlide(evd_ctx *ctx, long M, long N, long slideWidth, float *dW, long ldW, float *dY,
                                long ldY, float *work) {
  (void)work;
  evd_wy_from_slide_f32(ctx, M, N, slideWidth, dW, ldW, dY, ldY);
}

// api/inc/computerQFromWY.h:4-5 (double, as in the reference) and the FP32 form
inline void computerQFromWY(evd_ctx *ctx, long M, long N, double *dQ, long ldQ, double *dW, long ldW, double *dY,
                            long ldY) {
  evd_q_from_wy_f64(ctx, M, N, dQ, ldQ, dW, ldW, dY, ldY);
}
inline void computerQFromWY(evd_ctx *ctx, long M, long N, float *dQ, long ldQ, float *dW, long ldW, float *dY,
                            long ldY) {
  evd_q_from_wy_f32(ctx, M, N, dQ, ldQ, dW, ldW, dY, ldY);
}

// api/inc/BC_backTrans.h:3: Q <- H_K ... H_1 Q for the noBandU reflector
// layout.  The double form (the reference's) needs b = 32; the FP32 form
// takes b = 32 or 64.  Returns 0 or the status.
inline int my_BC_back_trans_v8_10_noBandU(evd_ctx *ctx, double *Q, long ldQ, double *U, long ldU, long n, int b) {
  return evd_bc_back_f64(ctx, n, b, n, Q, ldQ, U, ldU, 0);
}
inline int my_BC_back_trans_v8_10_noBandU(evd_ctx *ctx, float *Q, long ldQ, float *U, long ldU, long n, int b) {
  return evd_bc_back_f32(ctx, n, b, n, Q, ldQ, U, ldU, 0);
}
// The reference's argument order exactly (api/inc/BC_backTrans.h:3): the
// context takes the place of its trailing stream argument.
inline int my_BC_back_trans_v8_10_noBandU(double *Q, long ldQ, double *U, long ldU, long n, int b, evd_ctx *ctx) {
  return evd_bc_back_f64(ctx, n, b, n, Q, ldQ, U, ldU, 0);
}
inline int my_BC_back_trans_v8_10_noBandU(float *Q, long ldQ, float *U, long ldU, long n, int b, evd_ctx *ctx) {
  return evd_bc_back_f32(ctx, n, b, n, Q, ldQ, U, ldU, 0);
}

// util/inc/checkMetric.h: prints ||A^T A - I||_F / M as the reference does
// (A == B in every reference call).
inline void checkOrthogonality(evd_ctx *ctx, long M, long N, double *dA, long ldA, double *dB, long ldB,
                               double *work) {
  (void)dB, (void)ldB, (void)work;
  double o = -1.0;
  if (evd_orthogonality_f64(ctx, M, N, dA, ldA, &o) == 0) std::printf("The Orthogonality metrix is %g\n", o);
}
inline void checkOrthogonality(evd_ctx *ctx, long M, long N, float *dA, long ldA, float *dB, long ldB, float *work) {
  (void)dB, (void)ldB, (void)work;
  double o = -1.0;
  if (evd_orthogonality_f32(ctx, M, N, dA, ldA, &o) == 0) std::printf("The Orthogonality metrix is %g\n", o);
}

// LAPACKE_dstedc(LAPACK_COL_MAJOR, 'I', n, D, E, Z, n) as used at
// src/EVD/myEVD.cu:529, on device pointers.
inline int dstedc(evd_ctx *ctx, long n, double *D, double *E, double *Z, long ldz) {
  return evd_stedc_f64(ctx, n, D, E, Z, ldz);
}

}  // namespace evd4sc
