// FP64 forms of the reference's api/ entry points.  The reference's library
// is instantiated for double only (api/zy_zy_sy2sb.cu:488-507,
// api/inc/computerWYFromSlide.h:4-5, api/inc/computerQFromWY.h:4-5,
// api/inc/BC_backTrans.h:3), so a caller that binds to it passes double
// buffers; these entry points take them as they are.  The arithmetic runs in
// FP64 throughout: GEMMs on the FP64 matrix cores (v_mfma_f64_16x16x4f64,
// gemm.hip), the panel factorisation in the FP64 panel kernel
// (panel_qr.hip, FP64 storage), and the back transformation of the chase
// reflectors in a register-window kernel below.
#include "common.hpp"

namespace evd {

int panel_qr_f64(evd_ctx *ctx, long m, long nc, const double *A, long lda, double *Y, long ldy, double *W, long ldw,
                 double *R, long ldr, double *Aout, long ldo, double *tau_out);

namespace {

// upper triangle of C (n x n) <- its lower triangle
__global__ void mirror_lower_f64_kernel(long n, double *__restrict__ C, long ldc) {
  for (long c = blockIdx.y; c < n; c += gridDim.y)
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < c; r += (long)gridDim.x * blockDim.x)
      C[r + c * ldc] = C[c + r * ldc];
}

__global__ void add_identity_f64_kernel(long m, double *Q, long ldq) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < m; i += (long)gridDim.x * blockDim.x)
    Q[i + i * ldq] += 1.0;
}

__global__ void fill_f64_2d_kernel(double *p, long rows, long cols, long ld) {
  for (long c = blockIdx.y; c < cols; c += gridDim.y)
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < rows; r += (long)gridDim.x * blockDim.x)
      p[r + c * ld] = 0.0;
}

dim3 grid2(long rows, long cols) {
  return dim3((unsigned)std::max<long>(1, std::min<long>((rows + 255) / 256, 64)),
              (unsigned)std::max<long>(1, std::min<long>(cols, 8192)));
}

}  // namespace

// C = alpha (A B^T + B A^T) + beta C, lower triangle computed as two FP64
// MFMA products and mirrored to the upper (the semantics of
// tc_ozimmu_syr2k, api/TCblas/tc_ozimmu_syr2k.cu:209-341).
int syr2k_f64(evd_ctx *ctx, long n, long k, double alpha, const double *A, long lda, const double *B, long ldb,
              double beta, double *C, long ldc) {
  if (n <= 0) return 0;
  EVD_TRY(dgemm(ctx, 'N', 'T', n, n, k, alpha, A, lda, B, ldb, beta, C, ldc));
  EVD_TRY(dgemm(ctx, 'N', 'T', n, n, k, alpha, B, ldb, A, lda, 1.0, C, ldc));
  mirror_lower_f64_kernel<<<grid2(n, n), 256, 0, ctx->stream>>>(n, C, ldc);
  EVD_LAUNCHED(ctx);
  return 0;
}

// FP64 SBR: the ZY panel loop of sy2sb (sbr.hip) on FP64 storage.
size_t sy2sb_f64_workspace(long n, long b, long nb) {
  return align256((size_t)n * nb * sizeof(double)) + align256((size_t)nb * b * sizeof(double)) * 2 +
         ((size_t)64 << 20);
}

int sy2sb_f64(evd_ctx *ctx, long n, long b, long nb, double *A, long lda, double *W, long ldw, double *Y, long ldy) {
  if (n <= 0) return 0;
  fill_f64_2d_kernel<<<grid2(n, n), 256, 0, ctx->stream>>>(W, n, n, ldw);
  EVD_LAUNCHED(ctx);
  fill_f64_2d_kernel<<<grid2(n, n), 256, 0, ctx->stream>>>(Y, n, n, ldy);
  EVD_LAUNCHED(ctx);
  ArenaFrame frame(ctx);
  const long ldz = n;
  double *Z = arena_alloc_n<double>(ctx, (size_t)n * nb);
  double *t = arena_alloc_n<double>(ctx, (size_t)nb * b);
  double *t3 = arena_alloc_n<double>(ctx, (size_t)b * b);
  if (!Z || !t || !t3) return EVD_E_NOMEM;
  const long P = nb / b;
  long c0 = 0;
  while (n - c0 - b > 1) {
    long p = 0;
    for (;; p++) {
      const long c = c0 + p * b, r0 = c + b, m = n - r0;
      double *Ap = A + r0 + c * lda;
      double *Wp = W + r0 + c * ldw;
      double *Yp = Y + r0 + c * ldy;
      double *Zp = Z + r0 + p * b * ldz;
      EVD_TRY(panel_qr_f64(ctx, m, b, Ap, lda, Yp, ldy, Wp, ldw, nullptr, 0, Ap, lda, nullptr));
      EVD_TRY(dgemm(ctx, 'N', 'N', m, b, m, 1.0, A + r0 + r0 * lda, lda, Wp, ldw, 0.0, Zp, ldz));
      if (p > 0) {
        const long pb = p * b;
        const double *Yq = Y + r0 + c0 * ldy;
        const double *Zq = Z + r0;
        EVD_TRY(dgemm(ctx, 'T', 'N', pb, b, m, 1.0, Zq, ldz, Wp, ldw, 0.0, t, pb));
        EVD_TRY(dgemm(ctx, 'N', 'N', m, b, pb, -1.0, Yq, ldy, t, pb, 1.0, Zp, ldz));
        EVD_TRY(dgemm(ctx, 'T', 'N', pb, b, m, 1.0, Yq, ldy, Wp, ldw, 0.0, t, pb));
        EVD_TRY(dgemm(ctx, 'N', 'N', m, b, pb, -1.0, Zq, ldz, t, pb, 1.0, Zp, ldz));
      }
      EVD_TRY(dgemm(ctx, 'T', 'N', b, b, m, 1.0, Wp, ldw, Zp, ldz, 0.0, t3, b));
      EVD_TRY(dgemm(ctx, 'N', 'N', m, b, b, -0.5, Yp, ldy, t3, b, 1.0, Zp, ldz));
      const long kk = (p + 1) * b;
      const double *Ya = Y + r0 + c0 * ldy;
      const double *Za = Z + r0;
      const bool more = (p + 1 < P) && (n - (r0 + b) > 1);
      if (more) {
        double *An = A + r0 + r0 * lda;
        EVD_TRY(dgemm(ctx, 'N', 'T', m, b, kk, -1.0, Ya, ldy, Za, ldz, 1.0, An, lda));
        EVD_TRY(dgemm(ctx, 'N', 'T', m, b, kk, -1.0, Za, ldz, Ya, ldy, 1.0, An, lda));
        continue;
      }
      EVD_TRY(syr2k_f64(ctx, m, kk, -1.0, Ya, ldy, Za, ldz, 1.0, A + r0 + r0 * lda, lda));
      break;
    }
    c0 += (p + 1) * b;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// FP64 back transformation of the chase reflectors, b = 32 (the reference's
// BC_kernel_computerQ_1Col_V8_10_noBandU, api/BC_backTransV8_10_noBandU.cu:
// 60-313, applies exactly these in FP64).  Block (S, j) holds the 32
// reflectors of sweeps 32S .. 32S+31 at step j; reflector t acts on rows
// R0 + t .. R0 + t + 31 with R0 = 32 (S + j) + 1, so the block lives on the
// 64-row window R0 .. R0 + 63.  Reflectors that share rows are applied in
// sweep order: for Q <- H_K .. H_1 Q (order 0) blocks run S ascending, j
// descending, t ascending; for Q <- H_1 .. H_K Q (order 1) the reverse.
// One wave owns 64 columns of Q (lane = column) and keeps the window in
// registers as two 32-row halves; consecutive blocks of one S share 32 rows,
// so each block loads and stores 32 rows.  The block's reflectors are staged
// in LDS and read as broadcasts.  Missing sweeps and rows count as zero
// reflectors (the identity).
namespace {

constexpr int kBT = 32;

__device__ __attribute__((always_inline)) inline void load_half(double (&h)[kBT], const double *__restrict__ Q,
                                                                long ldq, long col, long r0, long n, bool colok) {
#pragma unroll
  for (int i = 0; i < kBT; i++) {
    const long r = r0 + i;
    h[i] = (colok && r < n) ? Q[r + col * ldq] : 0.0;
  }
}
__device__ __attribute__((always_inline)) inline void store_half(const double (&h)[kBT], double *__restrict__ Q,
                                                                 long ldq, long col, long r0, long n, bool colok) {
#pragma unroll
  for (int i = 0; i < kBT; i++) {
    const long r = r0 + i;
    if (colok && r < n) Q[r + col * ldq] = h[i];
  }
}

// apply reflector t of the staged block to the window (lo = rows 0..31,
// hi = rows 32..63 of the window)
template <int t>
__device__ __attribute__((always_inline)) inline void apply_t(double (&lo)[kBT], double (&hi)[kBT],
                                                              const double *__restrict__ u) {
  double d = 0.0;
#pragma unroll
  for (int i = 0; i < kBT; i++) {
    const int r = t + i;
    d += u[t * kBT + i] * (r < kBT ? lo[r < kBT ? r : 0] : hi[r >= kBT ? r - kBT : 0]);
  }
#pragma unroll
  for (int i = 0; i < kBT; i++) {
    const int r = t + i;
    if (r < kBT)
      lo[r < kBT ? r : 0] -= u[t * kBT + i] * d;
    else
      hi[r >= kBT ? r - kBT : 0] -= u[t * kBT + i] * d;
  }
}

template <int... T>
struct TSeq {};
template <int N, int... T>
struct MakeT : MakeT<N - 1, N - 1, T...> {};
template <int... T>
struct MakeT<0, T...> {
  using type = TSeq<T...>;
};

template <int... T>
__device__ __attribute__((always_inline)) inline void apply_fwd(double (&lo)[kBT], double (&hi)[kBT],
                                                                const double *u, TSeq<T...>) {
  (apply_t<T>(lo, hi, u), ...);
}
template <int... T>
__device__ __attribute__((always_inline)) inline void apply_rev(double (&lo)[kBT], double (&hi)[kBT],
                                                                const double *u, TSeq<T...>) {
  (apply_t<kBT - 1 - T>(lo, hi, u), ...);
}

__global__ __launch_bounds__(64) void bcback_f64_kernel(int n, int ncols, double *__restrict__ Q, long ldq,
                                                        const double *__restrict__ U, long ldu, int order) {
  __shared__ double us[kBT * kBT];
  const int lane = threadIdx.x;
  const long col = (long)blockIdx.x * 64 + lane;
  const bool colok = col < ncols;
  const long nsweeps = n - 2;
  if (nsweeps <= 0) return;
  const long nS = (nsweeps + kBT - 1) / kBT;
  double lo[kBT], hi[kBT];
  // stage block (S, j): us[t][i] = U[R0 + t + i, 32 S + t]
  auto stage = [&](long S, long R0) {
    __syncthreads();
    for (int e = lane; e < kBT * kBT; e += 64) {
      const int t = e / kBT, i = e % kBT;
      const long s = S * kBT + t, r = R0 + t + i;
      us[e] = (s < nsweeps && r < n) ? U[r + s * ldu] : 0.0;
    }
    __syncthreads();
  };
  for (long si = 0; si < nS; si++) {
    const long S = order == 0 ? si : nS - 1 - si;
    // steps of this sweep block: windows start below row n - 1
    const long J = std::max<long>(0, (n - 2) / kBT - S + 1);
    if (J <= 0) continue;
    if (order == 0) {
      // j descending: the window moves up by 32 rows per block
      long R0 = kBT * (S + J - 1) + 1;
      load_half(lo, Q, ldq, col, R0, n, colok);
      load_half(hi, Q, ldq, col, R0 + kBT, n, colok);
      for (long j = J - 1; j >= 0; j--) {
        stage(S, R0);
        apply_fwd(lo, hi, us, typename MakeT<kBT>::type{});
        store_half(hi, Q, ldq, col, R0 + kBT, n, colok);
        if (j > 0) {
#pragma unroll
          for (int i = 0; i < kBT; i++) hi[i] = lo[i];
          R0 -= kBT;
          load_half(lo, Q, ldq, col, R0, n, colok);
        } else {
          store_half(lo, Q, ldq, col, R0, n, colok);
        }
      }
    } else {
      // j ascending: the window moves down by 32 rows per block
      long R0 = kBT * S + 1;
      load_half(lo, Q, ldq, col, R0, n, colok);
      load_half(hi, Q, ldq, col, R0 + kBT, n, colok);
      for (long j = 0; j < J; j++) {
        stage(S, R0);
        apply_rev(lo, hi, us, typename MakeT<kBT>::type{});
        store_half(lo, Q, ldq, col, R0, n, colok);
        if (j + 1 < J) {
#pragma unroll
          for (int i = 0; i < kBT; i++) lo[i] = hi[i];
          R0 += kBT;
          load_half(hi, Q, ldq, col, R0 + kBT, n, colok);
        } else {
          store_half(hi, Q, ldq, col, R0 + kBT, n, colok);
        }
      }
    }
  }
}

}  // namespace

int bc_back_f64(evd_ctx *ctx, long n, long b, long ncols, double *Q, long ldq, const double *U, long ldu, int order) {
  if (n <= 2 || ncols <= 0) return 0;
  if (b != kBT) return set_error(ctx, EVD_E_UNSUPPORTED, "the FP64 back transformation supports b = 32 (got %ld)", b);
  bcback_f64_kernel<<<(unsigned)((ncols + 63) / 64), 64, 0, ctx->stream>>>((int)n, (int)ncols, Q, ldq, U, ldu, order);
  EVD_LAUNCHED(ctx);
  return 0;
}

}  // namespace evd

using namespace evd;

extern "C" int evd_syr2k_f64(evd_ctx *ctx, long n, long k, double alpha, const double *A, long lda, const double *B,
                             long ldb, double beta, double *C, long ldc) {
  if (!ctx) return -1;
  EVD_ARG(ctx, n >= 0, 2);
  EVD_ARG(ctx, k >= 0, 3);
  EVD_ARG(ctx, lda >= std::max<long>(1, n), 6);
  EVD_ARG(ctx, ldb >= std::max<long>(1, n), 8);
  EVD_ARG(ctx, ldc >= std::max<long>(1, n), 11);
  return syr2k_f64(ctx, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

extern "C" int evd_panel_qr_f64(evd_ctx *ctx, long m, long nb, double *A, long lda, double *W, long ldw, double *R,
                                long ldr) {
  if (!ctx) return -1;
  EVD_ARG(ctx, m >= 0, 2);
  EVD_ARG(ctx, nb >= 0 && nb <= 64, 3);
  EVD_ARG(ctx, A != nullptr || m == 0, 4);
  EVD_ARG(ctx, lda >= std::max<long>(1, m), 5);
  EVD_ARG(ctx, W != nullptr || m == 0, 6);
  EVD_ARG(ctx, ldw >= std::max<long>(1, m), 7);
  EVD_ARG(ctx, ldr >= std::max<long>(1, nb) || R == nullptr, 9);
  if (ctx->arena_top == 0) EVD_TRY(arena_ensure(ctx, (size_t)1 << 20));
  return panel_qr_f64(ctx, m, nb, A, lda, A, lda, W, ldw, R, ldr, nullptr, 0, nullptr);
}

extern "C" int evd_sy2sb_f64(evd_ctx *ctx, long n, long b, long nb, double *A, long lda, double *W, long ldw,
                             double *Y, long ldy) {
  if (!ctx) return -1;
  EVD_ARG(ctx, n >= 0, 2);
  EVD_ARG(ctx, b >= 1 && b <= 64, 3);
  EVD_ARG(ctx, nb >= b && nb % b == 0, 4);
  EVD_ARG(ctx, lda >= std::max<long>(1, n), 6);
  EVD_ARG(ctx, ldw >= std::max<long>(1, n), 8);
  EVD_ARG(ctx, ldy >= std::max<long>(1, n), 10);
  if (ctx->arena_top == 0) EVD_TRY(arena_ensure(ctx, sy2sb_f64_workspace(n, b, nb)));
  return sy2sb_f64(ctx, n, b, nb, A, lda, W, ldw, Y, ldy);
}

extern "C" int evd_wy_from_slide_f64(evd_ctx *ctx, long M, long N, long slide, double *W, long ldw, const double *Y,
                                     long ldy) {
  if (!ctx) return -1;
  EVD_ARG(ctx, M >= 0, 2);
  EVD_ARG(ctx, N >= 0, 3);
  EVD_ARG(ctx, slide >= 1, 4);
  EVD_ARG(ctx, ldw >= std::max<long>(1, M), 6);
  EVD_ARG(ctx, ldy >= std::max<long>(1, M), 8);
  if (ctx->arena_top == 0) EVD_TRY(arena_ensure(ctx, (size_t)N * slide * sizeof(double) + ((size_t)1 << 20)));
  ArenaFrame frame(ctx);
  double *work = arena_alloc_n<double>(ctx, (size_t)std::max<long>(N, 1) * slide);
  if (!work) return EVD_E_NOMEM;
  for (long i = 2 * slide; i <= N; i += slide) {
    const long w = slide;
    EVD_TRY(dgemm(ctx, 'T', 'N', i - w, w, M, 1.0, Y, ldy, W + (i - w) * ldw, ldw, 0.0, work, i - w));
    EVD_TRY(dgemm(ctx, 'N', 'N', M, w, i - w, -1.0, W, ldw, work, i - w, 1.0, W + (i - w) * ldw, ldw));
  }
  return 0;
}

extern "C" int evd_q_from_wy_f64(evd_ctx *ctx, long M, long N, double *Q, long ldq, const double *W, long ldw,
                                 const double *Y, long ldy) {
  if (!ctx) return -1;
  EVD_ARG(ctx, M >= 0, 2);
  EVD_ARG(ctx, N >= 0, 3);
  EVD_ARG(ctx, ldq >= std::max<long>(1, M), 5);
  EVD_ARG(ctx, ldw >= std::max<long>(1, M), 7);
  EVD_ARG(ctx, ldy >= std::max<long>(1, M), 9);
  EVD_TRY(dgemm(ctx, 'N', 'T', M, M, N, -1.0, W, ldw, Y, ldy, 0.0, Q, ldq));
  if (M > 0) {
    add_identity_f64_kernel<<<(unsigned)std::min<long>((M + 255) / 256, 1024), 256, 0, ctx->stream>>>(M, Q, ldq);
    EVD_LAUNCHED(ctx);
  }
  return 0;
}

__global__ void copy_f64_2d_kernel(const double *s, long lds, double *d, long ldd, long rows, long cols) {
  for (long c = blockIdx.y; c < cols; c += gridDim.y)
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < rows; r += (long)gridDim.x * blockDim.x)
      d[r + c * ldd] = s[r + c * lds];
}

extern "C" int evd_orgqr_f64(evd_ctx *ctx, long m, long n, double *Q, long ldq, const double *W, long ldw,
                             double *work) {
  if (!ctx) return -1;
  EVD_ARG(ctx, m >= 0, 2);
  EVD_ARG(ctx, n >= 0 && n <= m, 3);
  EVD_ARG(ctx, ldq >= std::max<long>(1, m), 5);
  EVD_ARG(ctx, ldw >= std::max<long>(1, m), 7);
  EVD_ARG(ctx, work != nullptr || n == 0, 8);
  if (n == 0) return 0;
  copy_f64_2d_kernel<<<grid2(n, n), 256, 0, ctx->stream>>>(Q, ldq, work, n, n, n);
  EVD_LAUNCHED(ctx);
  EVD_TRY(dgemm(ctx, 'N', 'T', m, n, n, -1.0, W, ldw, work, n, 0.0, Q, ldq));
  add_identity_f64_kernel<<<(unsigned)std::min<long>((n + 255) / 256, 1024), 256, 0, ctx->stream>>>(n, Q, ldq);
  EVD_LAUNCHED(ctx);
  return 0;
}

extern "C" int evd_bc_back_f64(evd_ctx *ctx, long n, long b, long ncols, double *Q, long ldq, const double *U,
                               long ldu, int order) {
  if (!ctx) return -1;
  EVD_ARG(ctx, n >= 0, 2);
  EVD_ARG(ctx, b == 32, 3);
  EVD_ARG(ctx, ncols >= 0, 4);
  EVD_ARG(ctx, ldq >= std::max<long>(1, n), 6);
  EVD_ARG(ctx, ldu >= n + b, 8);
  EVD_ARG(ctx, order == 0 || order == 1, 9);
  return bc_back_f64(ctx, n, b, ncols, Q, ldq, U, ldu, order);
}
