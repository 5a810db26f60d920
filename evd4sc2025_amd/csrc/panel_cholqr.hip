// Panel QR by Cholesky QR and Householder reconstruction: the role of the
// reference's TSQR panel (api/inc/TallShinnyQR.h:63-98: tile QRs, a QR of
// the stacked R factors, then Householder vectors rebuilt with getrf), with
// the same outputs as panel_qr_kernel: Y = V (unit lower), W = V T, R (and
// [R; 0] into the panel storage), T and R those of LAPACK dgeqrf's reflectors.
//
// Algorithm (FP64 throughout, nc <= 32 columns, m >= 2 nc rows):
//   G1 = A^T A, R1 = chol(G1), Q1 = A R1^-1          (Cholesky QR)
//   G2 = Q1^T Q1, R2 = chol(G2), Q = Q1 R2^-1         (second pass: Q
//                                                       orthonormal to rounding)
//   R = R2 R1, A = Q R.
// Householder reconstruction (Ballard et al., "Reconstructing Householder
// vectors from tall-skinny QR", 2015): the LU factorisation without pivoting
// of Q - [S; 0], with the diagonal sign matrix S chosen step by step as
// S_jj = -sign(pivot) (the pivot then has magnitude >= 1), gives
//   Q - [S; 0] = V U',  V = [V1; Q2 U'^-1] unit lower trapezoidal,
//   T = -U' S V1^-T,   R_H = S R,
// so that (I - V T V^T) [R_H; 0] = A.  The sign rule is LAPACK's (beta =
// -sign(alpha) ||x||), so V, T and R_H are dgeqrf's reflectors.
//
// MI355X mapping: three stream-ordered launches instead of one exchange per
// column.  The Gram matrices are 32 x 32 FP64 products over the panel's rows
// on v_mfma_f64_16x16x4f64, per workgroup of 256 rows, summed in a fixed
// order inside the launch; the workgroup that forms a problem's sum factors it
// (Cholesky, triangular inverses, the sign-chosen LU on one wave); the second
// pass does the same when the first asked for it; the output kernel forms V,
// W and [R; 0] row by row.  Same-size panels of several problems run as one
// launch each (grid y = problem).
//
// A panel whose Gram matrix is not safely positive definite (rank deficient or
// condition number above about 1e6: Cholesky QR would lose orthogonality) is
// flagged on the device; the later launches skip that problem and the
// column-exchange kernel (panel_qr.hip), launched after them, factors it (and
// exits at once for the problems that succeeded).
#include "common.hpp"

namespace evd {

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kQ = 32;          // columns of a panel (padded)
constexpr int kRows = 256;      // rows per workgroup of the row kernels
constexpr int kGroup = 8;       // partial Gram matrices summed per reduction step
constexpr int kQQ = kQ * kQ;    // one 32 x 32 matrix
constexpr int kLD = kQ + 1;     // LDS pitch of a 32 x 32 matrix or of a row

// per-problem factors, kFactorWords doubles: R1, R1^-1, T, M = R^-1 U'^-1,
// V1, R_H, tau
constexpr int kF_R1 = 0, kF_R1I = kQQ, kF_T = 2 * kQQ, kF_M = 3 * kQQ, kF_V1 = 4 * kQQ, kF_RH = 5 * kQQ,
              kF_TAU = 6 * kQQ;
constexpr int kFactorWords = 6 * kQQ + kQ;

// per-problem state word (the caller's `bad` array)
constexpr int kNeedPass2 = 0;  // first factor done, second Gram pass pending
constexpr int kFallback = 1;   // not factored: the column-exchange kernel takes it
constexpr int kDone = 2;       // V, T, R_H ready for the output kernel

// Completes this wave's LDS accesses before its next ones (and is a compiler
// barrier): used where lanes of one wave hand rows to each other through LDS
// without a workgroup barrier
__device__ __attribute__((always_inline)) inline void lds_order() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

__device__ __forceinline__ f64x4 mfma64(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// the value of v in lane src (a compile-time or wave-uniform index)
__device__ __forceinline__ double lane_value(double v, int src) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, src);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), src);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Row r of the m x nc panel A (column-major) as 32 doubles, zero past m and
// nc.  The loads use clamped (in-bounds) addresses and the zeroing is a
// select, so all 32 are in flight together.
template <typename T>
__device__ __forceinline__ void load_row(const T *__restrict__ A, long lda, int m, int nc, long r, double (&a)[kQ]) {
  const bool rin = r < m;
  const long rr = rin ? r : 0;
  T v[kQ];
#pragma unroll
  for (int c = 0; c < kQ; c++) v[c] = A[rr + (long)(c < nc ? c : 0) * lda];
#pragma unroll
  for (int c = 0; c < kQ; c++) a[c] = (rin && c < nc) ? (double)v[c] : 0.0;
}

// The 32 x 32 row-major matrix src (global) into LDS (pitch kLD): four loads
// per thread, all in flight together.  Call from all 256 threads.
__device__ __forceinline__ void load_mat(const double *__restrict__ src, double *dst, int tid) {
  double v[4];
#pragma unroll
  for (int r = 0; r < 4; r++) v[r] = src[tid + 256 * r];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int e = tid + 256 * r;
    dst[(e / kQ) * kLD + e % kQ] = v[r];
  }
}

// dst = sum of the n 32 x 32 matrices src[0 .. n) (kQQ doubles apart), added
// in index order; written by other workgroups of this launch, so read and
// written with device-coherent accesses.  Call from all 256 threads.
__device__ __forceinline__ void sum_slots(const double *src, int n, double *dst, int tid, double (&sum)[4]) {
#pragma unroll
  for (int r = 0; r < 4; r++) sum[r] = 0.0;
  for (int g0 = 0; g0 < n; g0 += 8) {
    double v[8][4];
#pragma unroll
    for (int u = 0; u < 8; u++)
#pragma unroll
      for (int r = 0; r < 4; r++) v[u][r] = ld_dev(&src[(long)min(g0 + u, n - 1) * kQQ + tid + 256 * r]);
#pragma unroll
    for (int u = 0; u < 8; u++)
#pragma unroll
      for (int r = 0; r < 4; r++) sum[r] += (g0 + u < n) ? v[u][r] : 0.0;
  }
#pragma unroll
  for (int r = 0; r < 4; r++) st_dev(&dst[tid + 256 * r], sum[r]);
}

// Counts this workgroup into *cnt (n members) after its coherent stores are
// acknowledged; true in the last one to arrive, which also resets the count
// for the next launch.  Call from all 256 threads.
__device__ __forceinline__ bool last_arrival(unsigned *cnt, unsigned n) {
  __shared__ int s_last;
  wait_mem();
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned t = __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (t == n - 1) ? 1 : 0;
    if (t == n - 1) st_dev(cnt, 0u);
  }
  __syncthreads();
  return s_last != 0;
}

// Rows 64w .. 64w+63 of xs (LDS, pitch kLD) times the 32 x 32 matrix B (LDS,
// pitch kLD) on wave w: acc[p][t][r] = C[64w + 16p + (lane & 15)][16t +
// (lane >> 4) + 4r].  The MFMA takes B as its first operand, so the
// accumulator lane index runs down the rows.
__device__ __forceinline__ void wave_rows_times(const double *xs, const double *B, int w, int lane,
                                                f64x4 (&acc)[4][2]) {
#pragma unroll
  for (int p = 0; p < 4; p++)
#pragma unroll
    for (int t = 0; t < 2; t++) acc[p][t] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int ks = 0; ks < kQ / 4; ks++) {
    const int k = 4 * ks + (lane >> 4);
    double a[4], b[2];
#pragma unroll
    for (int p = 0; p < 4; p++) a[p] = xs[(64 * w + 16 * p + (lane & 15)) * kLD + k];
#pragma unroll
    for (int t = 0; t < 2; t++) b[t] = B[k * kLD + 16 * t + (lane & 15)];
#pragma unroll
    for (int p = 0; p < 4; p++)
#pragma unroll
      for (int t = 0; t < 2; t++) acc[p][t] = mfma64(b[t], a[p], acc[p][t]);
  }
}

__device__ __forceinline__ void wave_store_rows(double *xs, int w, int lane, const f64x4 (&acc)[4][2]) {
#pragma unroll
  for (int p = 0; p < 4; p++)
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
      for (int r = 0; r < 4; r++) xs[(64 * w + 16 * p + (lane & 15)) * kLD + 16 * t + (lane >> 4) + 4 * r] = acc[p][t][r];
}

// Gram matrix of the panel's rows: rows of A (MODE 0) or of Q1 = A R1^-1
// (MODE 1, R1^-1 from the factors, product on MFMA), 256 rows per workgroup.
// Thread t stages row t in LDS and wave w reads only its own 64 rows.  The
// FP64 MFMA forms the three distinct 16 x 16 blocks of the symmetric partial
// sum.  Reduction inside the launch: the last workgroup to arrive in each
// group of kGroup sums the group's partials into a level-2 slot, and the
// last group to finish sums those into the final slot (always in index
// order, so the sum does not depend on arrival order).  part per problem:
// [ng partials][ngroups level-2 sums][final]; cnt: ngroups + 1 counters per
// problem, zero between launches.  Returns true in the workgroup that formed
// the final sum, whose thread t then holds its elements t + 256 r in gv[r].
// A, fac, part and cnt point at this problem's data; xs has kRows * kLD
// doubles of LDS (the rows, then the waves' partial sums) and ri (MODE 1)
// kQ * kLD.
template <int MODE, typename T>
__device__ __forceinline__ bool gram_pass(int m, int nc, const T *__restrict__ A, long lda,
                                          const double *__restrict__ fac, double *part, unsigned *cnt, int g, int ng,
                                          double *xs, double *ri, double (&gv)[4]) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int ngroups = (ng + kGroup - 1) / kGroup;
  double a[kQ];
  load_row(A, lda, m, nc, (long)g * kRows + tid, a);
  if (MODE == 1) load_mat(fac + kF_R1I, ri, tid);
#pragma unroll
  for (int c = 0; c < kQ; c++) xs[tid * kLD + c] = a[c];
  if (MODE == 1) {
    __syncthreads();  // ri
    f64x4 qa[4][2];
    wave_rows_times(xs, ri, w, lane, qa);
    lds_order();
    wave_store_rows(xs, w, lane, qa);
  }
  lds_order();
  // acc[0] = G[0:16][0:16], acc[1] = G[0:16][16:32], acc[2] = G[16:32][16:32];
  // element e of a lane: row (lane & 15), column (lane >> 4) + 4e of the block
  f64x4 acc[3];
#pragma unroll
  for (int t = 0; t < 3; t++) acc[t] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int ks = 0; ks < 16; ks++) {
    const int row = 64 * w + 4 * ks + (lane >> 4);
    const double x0 = xs[row * kLD + (lane & 15)], x1 = xs[row * kLD + 16 + (lane & 15)];
    acc[0] = mfma64(x0, x0, acc[0]);
    acc[1] = mfma64(x1, x0, acc[1]);
    acc[2] = mfma64(x1, x1, acc[2]);
  }
  __syncthreads();  // xs is free: the waves' partial sums go there
#pragma unroll
  for (int e = 0; e < 4; e++) {
    const int i = lane & 15, j = (lane >> 4) + 4 * e;
    double *xw = xs + w * kQQ;
    xw[i * kQ + j] = acc[0][e];
    xw[i * kQ + 16 + j] = acc[1][e];
    xw[(16 + j) * kQ + i] = acc[1][e];
    xw[(16 + i) * kQ + 16 + j] = acc[2][e];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int e = tid + 256 * r;
    st_dev(&part[(long)g * kQQ + e], ((xs[e] + xs[kQQ + e]) + xs[2 * kQQ + e]) + xs[3 * kQQ + e]);
  }
  const int grp = g / kGroup, gsize = min(kGroup, ng - grp * kGroup);
  if (!last_arrival(cnt + grp, (unsigned)gsize)) return false;
  sum_slots(part + (long)grp * kGroup * kQQ, gsize, part + (long)(ng + grp) * kQQ, tid, gv);
  if (!last_arrival(cnt + ngroups, (unsigned)ngroups)) return false;
  sum_slots(part + (long)ng * kQQ, ngroups, part + (long)(ng + ngroups) * kQQ, tid, gv);
  return true;
}

// 1 / sqrt(x) and 1 / x for x > 0 (finite) to FP64 rounding level: the
// hardware approximation refined by two Newton steps (each doubles the
// correct bits), a short dependent chain next to the IEEE square root and
// division sequences.
__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
#pragma unroll
  for (int it = 0; it < 2; it++) {
    const double e = fma(-x * y, y, 1.0);  // 1 - x y^2
    y = fma(0.5 * y, e, y);
  }
  return y;
}
__device__ __forceinline__ double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
#pragma unroll
  for (int it = 0; it < 2; it++) {
    const double e = fma(-x, y, 1.0);  // 1 - x y
    y = fma(y, e, y);
  }
  return y;
}

// Upper Cholesky factor of the 32 x 32 matrix G (LDS, pitch kLD) into R (LDS),
// on one wave in left-looking (Crout) order: lane c keeps column c of R in
// registers and reads the finished rows k < j back from LDS as broadcasts; the
// pivot comes from lane j (readlane) and its reciprocal square root from
// rsqrt_nr (lanes 32..63 repeat lanes 0..31 and store nothing).  Returns false
// when a pivot is not above tol_rel times the largest diagonal entry (or is
// not finite).
__device__ __forceinline__ bool wave_cholesky(const double *G, double *R, int lane, double tol_rel) {
  const int c = lane & (kQ - 1);
  double dmax = G[c * kLD + c];
#pragma unroll
  for (int o = 16; o > 0; o >>= 1) dmax = fmax(dmax, __shfl_xor(dmax, o, 64));
  const double tol = tol_rel * dmax;
  bool ok = dmax > 0.0 && dmax < 1e300;
  double r[kQ];
#pragma unroll
  for (int j = 0; j < kQ; j++) {
    double s0 = G[j * kLD + c], s1 = 0.0;  // two chains of the dot product
#pragma unroll
    for (int k = 0; k < j; k++) {
      if (k & 1)
        s1 -= R[k * kLD + j] * r[k];
      else
        s0 -= R[k * kLD + j] * r[k];
    }
    const double s = s0 + s1;
    const double d = lane_value(s, j);  // G[j][j] - sum_k R[k][j]^2
    if (!(d > tol)) ok = false;
    const double y = ok ? rsqrt_nr(d) : 1.0;
    r[j] = (c >= j) ? s * y : 0.0;
    if (lane < kQ) R[j * kLD + c] = r[j];
    lds_order();
  }
  return ok;
}

// X = U^-1 for an upper triangular 32 x 32 U on one wave: lane c forms column
// c by right-looking back substitution (x_k is final once the rows below k
// are applied, and the updates of the rows above it are independent, so the
// dependent chain is one multiply-add per step; the diagonal's reciprocals
// are formed up front).  UNIT_LT: U is the transpose of the unit lower
// triangle stored in M (U[i][k] = M[k][i], unit diagonal); otherwise U is
// M's upper triangle.
template <bool UNIT_LT>
__device__ __forceinline__ void wave_upper_inverse(const double *M, double *X, int lane) {
  const int c = lane & (kQ - 1);
  double s[kQ], rd[kQ];
#pragma unroll
  for (int i = 0; i < kQ; i++) {
    s[i] = (i == c) ? 1.0 : 0.0;
    rd[i] = UNIT_LT ? 1.0 : 1.0 / M[i * kLD + i];
  }
#pragma unroll
  for (int k = kQ - 1; k >= 0; k--) {
    const double xk = s[k] * rd[k];
    s[k] = xk;
#pragma unroll
    for (int i = 0; i < k; i++) s[i] -= (UNIT_LT ? M[k * kLD + i] : M[i * kLD + k]) * xk;
  }
  if (lane < kQ) {
#pragma unroll
    for (int i = 0; i < kQ; i++) X[i * kLD + c] = s[i];
  }
}

// C = A B for 32 x 32 matrices in LDS (pitch kLD), by the whole workgroup:
// thread t forms C[4 (t / 32) + r][t % 32], r < 4.  C may not alias A or B.
// Call from all 256 threads; the caller synchronises before and after.
__device__ __forceinline__ void block_matmul(const double *A, const double *B, double *C, int tid,
                                             double scale = 1.0) {
  const int j = tid % kQ, i0 = 4 * (tid / kQ);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
  for (int k = 0; k < kQ; k++) {
    const double bkj = B[k * kLD + j];
#pragma unroll
    for (int r = 0; r < 4; r++) acc[r] += A[(i0 + r) * kLD + k] * bkj;
  }
#pragma unroll
  for (int r = 0; r < 4; r++) C[(i0 + r) * kLD + j] = scale * acc[r];
}

// Sum of v over the workgroup (all 256 threads call; ends synchronised).
__device__ __forceinline__ double block_sum(double v, double *red, int tid) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((tid & 63) == 0) red[tid >> 6] = v;
  __syncthreads();
  return ((red[0] + red[1]) + red[2]) + red[3];
}

// Factorisation of one problem's panel from its Gram matrix, by the workgroup
// that formed the sum.  MODE 0: G1 -> R1, R1^-1; when
// ||R1||_F ||R1^-1||_F <= one_pass_max (a bound on the condition number of
// the panel small enough that Q1 = A R1^-1 is orthonormal to well below the
// output precision) the reconstruction follows at once with R = R1, else R1
// and R1^-1 are stored for the second pass.  MODE 1: G2 -> R2, R = R2 R1.
// Reconstruction: Q_top = A_top R^-1, sign-chosen LU of Q_top - S, T,
// M = R^-1 U'^-1, V1, R_H = S R, tau.
// gv: the Gram matrix, elements tid + 256 r; atop: the panel's top block,
// elements (i, k) = (e % 32, e / 32) of e = tid + 256 r; fac, state: this
// problem's; sm: 6 kQ * kLD + kQ + 4 doubles of LDS.
template <int MODE, typename T>
__device__ __forceinline__ void factor_body(int m, int nc, const double (&gv)[4], const T (&atop)[4],
                                            double *__restrict__ fac, int *state, double one_pass_max, double *sm) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  double *G = sm, *R = sm + kQ * kLD, *X = sm + 2 * kQ * kLD, *Y = sm + 3 * kQ * kLD, *Z = sm + 4 * kQ * kLD,
         *U = sm + 5 * kQ * kLD, *sgn = sm + 6 * kQ * kLD, *red = sgn + kQ;
  __shared__ int s_ok;
  // padded columns (c >= nc) get an identity block so the factorisation
  // proceeds
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int e = tid + 256 * r, i = e / kQ, j = e % kQ;
    G[i * kLD + j] = (i >= nc || j >= nc) ? ((i == j) ? 1.0 : 0.0) : gv[r];
  }
  __syncthreads();
  if (w == 0) {
    // a first pass accepts condition numbers up to about 1e6 (pivots >= 1e-12
    // of the largest diagonal entry); the second Gram matrix is near I
    const bool ok = wave_cholesky(G, R, lane, MODE == 0 ? 1e-12 : 0.5);
    if (lane == 0) s_ok = ok ? 1 : 0;
  }
  __syncthreads();
  if (!s_ok) {
    if (tid == 0) *state = kFallback;
    return;
  }
  // ---- Y <- R (the triangular factor of the panel), G <- R^-1
  if (MODE == 0) {
    if (w == 0) wave_upper_inverse<false>(R, X, lane);
    __syncthreads();
    double sr = 0.0, sx = 0.0;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int e = tid + 256 * r, i = e / kQ, j = e % kQ;
      const double a = R[i * kLD + j], b = X[i * kLD + j];
      sr += a * a;
      sx += b * b;
    }
    sr = block_sum(sr, red, tid);
    sx = block_sum(sx, red, tid);
    if (!(sqrt(sr * sx) <= one_pass_max)) {
      for (int e = tid; e < kQQ; e += 256) {
        fac[kF_R1 + e] = R[(e / kQ) * kLD + e % kQ];
        fac[kF_R1I + e] = X[(e / kQ) * kLD + e % kQ];
      }
      if (tid == 0) *state = kNeedPass2;
      return;
    }
    for (int e = tid; e < kQQ; e += 256) {
      const int i = e / kQ, j = e % kQ;
      Y[i * kLD + j] = R[i * kLD + j];
      G[i * kLD + j] = X[i * kLD + j];
    }
  } else {
    // R = R2 (from the Cholesky); R1 into G, R1^-1 into Z
    load_mat(fac + kF_R1, G, tid);
    load_mat(fac + kF_R1I, Z, tid);
    __syncthreads();
    block_matmul(R, G, Y, tid);  // Y = R2 R1
    if (w == 0) wave_upper_inverse<false>(R, X, lane);
    __syncthreads();
    block_matmul(Z, X, G, tid);  // G = R1^-1 R2^-1
  }
  __syncthreads();
  // ---- Q_top = A_top R^-1 (rows past m and padded columns are zero): A_top
  // into X, the product into Z
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int e = tid + 256 * r, i = e % kQ, k = e / kQ;
    X[i * kLD + k] = (k < nc && i < m) ? (double)atop[r] : 0.0;
  }
  __syncthreads();
  block_matmul(X, G, Z, tid);
  __syncthreads();
  // ---- sign-chosen LU of Q_top - S on wave 0, lane i holding row i (the
  // pivot row is read from lane j's registers): Z ends as U' (upper, diagonal
  // included) and L (strictly lower; unit diagonal)
  if (w == 0) {
    const int i = lane & (kQ - 1);
    double row[kQ];
#pragma unroll
    for (int c = 0; c < kQ; c++) row[c] = Z[i * kLD + c];
#pragma unroll
    for (int j = 0; j < kQ; j++) {
      const double piv = lane_value(row[j], j);
      const double sj = (piv >= 0.0) ? -1.0 : 1.0;
      const double ujj = piv - sj;  // |ujj| >= 1
      const double uinv = rcp_nr(ujj);
      double u[kQ];  // row j of U' (wave-uniform)
#pragma unroll
      for (int c = j + 1; c < kQ; c++) u[c] = lane_value(row[c], j);
      if (i > j) {
        const double l = row[j] * uinv;
        row[j] = l;
#pragma unroll
        for (int c = j + 1; c < kQ; c++) row[c] -= l * u[c];
      } else if (i == j) {
        row[j] = ujj;
      }
      if (lane == j) sgn[j] = sj;
    }
    if (lane < kQ) {
#pragma unroll
      for (int c = 0; c < kQ; c++) Z[i * kLD + c] = row[c];
    }
  }
  __syncthreads();
  // ---- R_H = S R and V1 = L (unit lower) out; U' into R, U' S into U
  for (int e = tid; e < kQQ; e += 256) {
    const int i = e / kQ, j = e % kQ;
    const double z = Z[i * kLD + j];
    fac[kF_RH + e] = sgn[i] * Y[i * kLD + j];
    fac[kF_V1 + e] = (i > j) ? z : (i == j ? 1.0 : 0.0);
    R[i * kLD + j] = (i <= j) ? z : 0.0;
    U[i * kLD + j] = (i <= j) ? z * sgn[j] : 0.0;
  }
  __syncthreads();
  // ---- wave 0: X = U'^-1; wave 1: Y = V1^-T (R is out, Y is free)
  if (w == 0) wave_upper_inverse<false>(R, X, lane);
  if (w == 1) wave_upper_inverse<true>(Z, Y, lane);
  __syncthreads();
  // ---- M = R^-1 U'^-1 into R; T = -U' S V1^-T into Z
  block_matmul(G, X, R, tid);
  block_matmul(U, Y, Z, tid, -1.0);
  __syncthreads();
  for (int e = tid; e < kQQ; e += 256) {
    const int i = e / kQ, j = e % kQ;
    fac[kF_M + e] = R[i * kLD + j];
    fac[kF_T + e] = (i <= j) ? Z[i * kLD + j] : 0.0;
  }
  if (tid < kQ) fac[kF_TAU + tid] = Z[tid * kLD + tid];
  if (tid == 0) *state = kDone;
}

// One Gram pass and, in the workgroup that forms a problem's Gram sum, its
// factorisation.  MODE 1 (the second pass) runs only for the problems whose
// first pass asked for it.  Grid (ng, count), 256 rows per workgroup.
template <int MODE, typename T>
__global__ __launch_bounds__(256) void cq_pass_kernel(int m, int nc, const T *__restrict__ A, long lda, long s_in,
                                                      double *__restrict__ fac, double *__restrict__ part,
                                                      long s_part, unsigned *__restrict__ cnt,
                                                      int *__restrict__ state, double one_pass_max) {
  const int g = blockIdx.x, q = blockIdx.y, ng = gridDim.x, tid = threadIdx.x;
  const int ngroups = (ng + kGroup - 1) / kGroup;
  if (MODE == 1 && state[q] != kNeedPass2) return;
  A += q * s_in;
  fac += (long)q * kFactorWords;
  // the panel's top 32 x 32 block, for the factorisation (loaded by every
  // workgroup: which one forms the sum is known only at the end)
  T atop[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int e = tid + 256 * r, i = e % kQ, k = e / kQ;
    atop[r] = A[min(i, m - 1) + (long)min(k, nc - 1) * lda];
  }
  // LDS: the Gram pass's rows (+ R1^-1), then the factorisation's matrices
  __shared__ double sm[kRows * kLD + (MODE == 1 ? kQ * kLD : 0)];
  double gv[4];
  if (!gram_pass<MODE, T>(m, nc, A, lda, fac, part + q * s_part, cnt + q * (ngroups + 1), g, ng, sm,
                          sm + kRows * kLD, gv))
    return;
  factor_body<MODE, T>(m, nc, gv, atop, fac, state + q, one_pass_max, sm);
}


// Outputs, 256 rows per workgroup: V = A M (rows < 32: V1), W = V T on MFMA
// through LDS (each wave on its own 64 rows), then Y = V, W, [R_H; 0] (Aout),
// R and tau with one thread per row.
template <typename T>
__global__ __launch_bounds__(256) void cq_out_kernel(int m, int nc, const T *__restrict__ A, long lda, long s_in,
                                                     const double *__restrict__ fac, T *__restrict__ Y, long ldy,
                                                     long s_y, T *__restrict__ W, long ldw, long s_w,
                                                     T *__restrict__ Aout, long ldo, long s_out, T *__restrict__ R,
                                                     long ldr, double *__restrict__ tau_out,
                                                     const int *__restrict__ state) {
  const int g = blockIdx.x, q = blockIdx.y, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  // every global load of the kernel is issued before the first wait: the
  // state word, this group's rows, M and T, and (group 0) V1 and R_H
  const int st = state[q];
  A += q * s_in;
  Y += q * s_y;
  W += q * s_w;
  if (Aout) Aout += q * s_out;
  fac += (long)q * kFactorWords;
  __shared__ double xs[kRows * kLD];
  __shared__ double Ms[kQ * kLD], Ts[kQ * kLD], V1s[kQ * kLD], RHs[kQ * kLD];
  const long r = (long)g * kRows + tid;
  double a[kQ], fv[4][4];
  load_row(A, lda, m, nc, r, a);
#pragma unroll
  for (int u = 0; u < 4; u++) {
    fv[0][u] = fac[kF_M + tid + 256 * u];
    fv[1][u] = fac[kF_T + tid + 256 * u];
  }
  if (g == 0) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      fv[2][u] = fac[kF_V1 + tid + 256 * u];
      fv[3][u] = fac[kF_RH + tid + 256 * u];
    }
  }
  if (st != kDone) return;
#pragma unroll
  for (int c = 0; c < kQ; c++) xs[tid * kLD + c] = a[c];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int e = tid + 256 * u, o = (e / kQ) * kLD + e % kQ;
    Ms[o] = fv[0][u];
    Ts[o] = fv[1][u];
    if (g == 0) {
      V1s[o] = fv[2][u];
      RHs[o] = fv[3][u];
    }
  }
  __syncthreads();
  f64x4 acc[4][2];
  wave_rows_times(xs, Ms, w, lane, acc);
  lds_order();
  wave_store_rows(xs, w, lane, acc);  // V (wave-local rows)
  if (r < kQ) {
#pragma unroll
    for (int c = 0; c < kQ; c++) xs[tid * kLD + c] = V1s[r * kLD + c];
  }
  lds_order();
  wave_rows_times(xs, Ts, w, lane, acc);  // W = V T
  if (r < m && Y) {
#pragma unroll
    for (int c = 0; c < kQ; c++)
      if (c < nc) Y[r + (long)c * ldy] = (T)xs[tid * kLD + c];
  }
  lds_order();
  wave_store_rows(xs, w, lane, acc);
  lds_order();
  if (r < m) {
#pragma unroll
    for (int c = 0; c < kQ; c++) {
      if (c >= nc) continue;
      if (W) W[r + (long)c * ldw] = (T)xs[tid * kLD + c];
      if (r < kQ) {
        const double rv = (r <= c) ? RHs[r * kLD + c] : 0.0;
        if (Aout) Aout[r + (long)c * ldo] = (T)rv;
        if (R && r < nc) R[r + (long)c * ldr] = (T)rv;
      } else if (Aout) {
        Aout[r + (long)c * ldo] = (T)0;
      }
    }
  }
  if (tau_out && r < nc) tau_out[r] = fac[kF_TAU + r];
}

}  // namespace

// Panels of `count` problems (A of problem q at q s_in, ...; R and tau only
// for count = 1).  Returns 0 after enqueueing; the state words `bad` (count
// ints, device) end as 1 for the problems the caller's fallback must factor
// (and 2 for the others).
template <typename T>
int panel_cholqr(evd_ctx *ctx, int count, long m, long nc, const T *A, long lda, long s_in, T *Y, long ldy, long s_y,
                 T *W, long ldw, long s_w, T *R, long ldr, T *Aout, long ldo, long s_out, double *tau_out,
                 int *bad) {
  const int ng = (int)((m + kRows - 1) / kRows);  // row workgroups
  const int ngroups = (ng + kGroup - 1) / kGroup;
  const long s_part = (long)(ng + ngroups + 1) * kQQ;
  // problems per launch: the reduction counters of one launch must fit
  const int per = std::max(1, std::min(count, kCqCountWords / (ngroups + 1)));
  if (ngroups + 1 > kCqCountWords) return set_error(ctx, EVD_E_UNSUPPORTED, "panel of %ld rows", m);
  double *part = arena_alloc_n<double>(ctx, (size_t)s_part * per);
  double *fac = arena_alloc_n<double>(ctx, (size_t)kFactorWords * per);
  if (!part || !fac) return EVD_E_NOMEM;
  unsigned *cnt = ctx->tags + kTagWords;
  const int mi = (int)m, ni = (int)nc;
  // One Cholesky pass leaves ||Q1^T Q1 - I|| near 1e-16 kappa^2 (FP64 Gram
  // sums): FP32 panels with ||R1||_F ||R1^-1||_F <= 1e3 (kappa_2 <= 1e3) take
  // it; FP64 panels always take the second pass.
  const double one_pass_max = sizeof(T) == 4 ? 1e3 : 0.0;
  for (int q0 = 0; q0 < count; q0 += per) {
    const int cq = std::min(per, count - q0);
    const T *Aq = A + q0 * s_in;
    T *Yq = Y + q0 * s_y, *Wq = W + q0 * s_w, *Oq = Aout ? Aout + q0 * s_out : nullptr;
    int *st = bad + q0;
    const dim3 rows_grid((unsigned)ng, (unsigned)cq);
    cq_pass_kernel<0, T><<<rows_grid, 256, 0, ctx->stream>>>(mi, ni, Aq, lda, s_in, fac, part, s_part, cnt, st,
                                                             one_pass_max);
    EVD_LAUNCHED(ctx);
    cq_pass_kernel<1, T><<<rows_grid, 256, 0, ctx->stream>>>(mi, ni, Aq, lda, s_in, fac, part, s_part, cnt, st,
                                                             one_pass_max);
    EVD_LAUNCHED(ctx);
    cq_out_kernel<T><<<rows_grid, 256, 0, ctx->stream>>>(mi, ni, Aq, lda, s_in, fac, Yq, ldy, s_y, Wq, ldw, s_w, Oq,
                                                        ldo, s_out, R, ldr, tau_out, st);
    EVD_LAUNCHED(ctx);
  }
  return 0;
}

template int panel_cholqr<float>(evd_ctx *, int, long, long, const float *, long, long, float *, long, long, float *,
                                 long, long, float *, long, float *, long, long, double *, int *);
template int panel_cholqr<double>(evd_ctx *, int, long, long, const double *, long, long, double *, long, long,
                                  double *, long, long, double *, long, double *, long, long, double *, int *);

}  // namespace evd
