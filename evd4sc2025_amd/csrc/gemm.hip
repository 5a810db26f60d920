// Dense MFMA kernels for gfx950.
//
//  * sgemm: C = alpha op(A) op(B) + beta C, FP32 in/out, v_mfma_f32_16x16x4f32
//    with FP32 accumulation; tile shapes for square, tall-skinny (n <= 32) and
//    short-wide (m <= 32) products; split-K with a deterministic two-pass
//    reduction for the long-K / small-MN reductions of the ZY updates.
//  * syr2k_lower_mirror: C = alpha (A B^T + B A^T) + beta C on lower tiles
//    only, the result mirrored to the upper triangle in the epilogue
//    (the role of tc_ozimmu_syr2k + CpyMatrixL2U in the reference).
//  * split 16-bit variants of sgemm / syr2k for the mixed-precision mode
//    (evd_ctx_set_precision): FP32 operands split into FP16 hi+lo or BF16
//    hi+mid+lo terms on the way into LDS, FP32 accumulation.
//  * dgemm / dgemm_grouped: FP64 on v_mfma_f64_16x16x4f64 for the
//    divide-and-conquer merges and the FP64 metrics.
//
// Operand staging: global -> registers (prefetch of tile k+1) -> LDS (k-major,
// the MFMA lane index running along the contiguous LDS dimension, rows padded
// by 16 elements) -> MFMA.  Operands are passed to the MFMA as (B-fragment,
// A-fragment) so that the accumulator lane index runs down a column of the
// column-major C, which makes the epilogue stores contiguous.
#include <rocblas/rocblas.h>

#include "common.hpp"

namespace evd {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kBK = 16;

// Batch count that tile-shape and split-K choices see: a launch that carries
// the same product for ctx->shape_div independent problems is shaped like the
// launch of one problem's products, so each problem's arithmetic (the K
// split, hence the summation order) is that of its single solve.
static inline long shape_batch(const evd_ctx *ctx, long batch) {
  return (ctx->shape_div > 1 && batch % ctx->shape_div == 0) ? batch / ctx->shape_div : batch;
}

template <bool TA, bool TB, int BM, int BN, int WM, int WN, typename T, int BK = kBK>
struct TileLoader {
  static constexpr int NA = BM * BK / 256;
  static constexpr int NB = BN * BK / 256;
  // LDS row pitch: an FP32 operand fragment reads 16 consecutive elements
  // from each of 4 k-rows; a pitch of 16 (mod 64) dwords puts the 4 rows on
  // disjoint banks.  FP64 keeps a 4-element pad (the FP64 GEMM measured
  // 47.6 TFLOP/s with it at 8192^3, 44.5 with the 16-element pad).
  static constexpr int LDSA = BM + (sizeof(T) == 4 ? 16 : 4);
  static constexpr int LDSB = BN + (sizeof(T) == 4 ? 16 : 4);
  T ra[NA], rb[NB];

  // 16-byte operand loads.  An operand goes through them when its base is
  // 16-byte aligned and its leading dimension a multiple of VW elements (the
  // caller's check, va / vb); each thread then moves VW consecutive elements
  // of the operand's contiguous dimension (M or N for A / B^T, K for A^T / B)
  // per load.  A vector that runs past the matrix edge or the K range falls
  // back to single elements.  ra / rb hold the elements in the vector
  // mapping, which store(..., va, vb) undoes; the element values (hence the
  // products) are those of the scalar path.
  // (FP32 only: for FP64 the two-element vectors slowed the grouped D&C
  // GEMM from 81 to 97 ms at n = 16384)
  static constexpr int VW = 16 / sizeof(T);
  static constexpr bool CAN_VA = sizeof(T) == 4 && (NA % VW == 0) && ((!TA ? BM : BK) % VW == 0);
  static constexpr bool CAN_VB = sizeof(T) == 4 && (NB % VW == 0) && ((TB ? BN : BK) % VW == 0);
  typedef T vecT __attribute__((ext_vector_type(16 / sizeof(T))));

  __device__ inline void load_a(const T *__restrict__ A, long lda, int M, int m0, int k0, int ke, int tid, bool va) {
    if (CAN_VA && va) {
#pragma unroll
      for (int v = 0; v < NA / VW; v++) {
        const int vidx = tid + 256 * v;
        int i, kk;
        if (!TA) {  // VW consecutive rows of one column
          i = (vidx % (BM / VW)) * VW;
          kk = vidx / (BM / VW);
        } else {  // VW consecutive k of one row of op(A)
          kk = (vidx % (BK / VW)) * VW;
          i = vidx / (BK / VW);
        }
        const int gi = m0 + i, gk = k0 + kk;
        const bool full = !TA ? (gi + VW <= M && gk < ke) : (gi < M && gk + VW <= ke);
        if (full) {
          const vecT x = *reinterpret_cast<const vecT *>(TA ? &A[gk + (long)gi * lda] : &A[gi + (long)gk * lda]);
#pragma unroll
          for (int e = 0; e < VW; e++) ra[v * VW + e] = x[e];
        } else {
#pragma unroll
          for (int e = 0; e < VW; e++) {
            const int ri = !TA ? gi + e : gi, rk = !TA ? gk : gk + e;
            ra[v * VW + e] = (ri < M && rk < ke) ? (TA ? A[rk + (long)ri * lda] : A[ri + (long)rk * lda]) : T(0);
          }
        }
      }
      return;
    }
#pragma unroll
    for (int q = 0; q < NA; q++) {
      int idx = tid + 256 * q;
      int i, kk;
      if (!TA) {
        i = idx % BM;
        kk = idx / BM;
      } else {
        kk = idx % BK;
        i = idx / BK;
      }
      int gi = m0 + i, gk = k0 + kk;
      T v = T(0);
      if (gi < M && gk < ke) v = TA ? A[gk + (long)gi * lda] : A[gi + (long)gk * lda];
      ra[q] = v;
    }
  }

  __device__ inline void load_b(const T *__restrict__ B, long ldb, int N, int n0, int k0, int ke, int tid, bool vb) {
    if (CAN_VB && vb) {
#pragma unroll
      for (int v = 0; v < NB / VW; v++) {
        const int vidx = tid + 256 * v;
        int j, kk;
        if (TB) {  // VW consecutive columns j of op(B) = rows of B
          j = (vidx % (BN / VW)) * VW;
          kk = vidx / (BN / VW);
        } else {  // VW consecutive k of one column
          kk = (vidx % (BK / VW)) * VW;
          j = vidx / (BK / VW);
        }
        const int gj = n0 + j, gk = k0 + kk;
        const bool full = TB ? (gj + VW <= N && gk < ke) : (gj < N && gk + VW <= ke);
        if (full) {
          const vecT x = *reinterpret_cast<const vecT *>(TB ? &B[gj + (long)gk * ldb] : &B[gk + (long)gj * ldb]);
#pragma unroll
          for (int e = 0; e < VW; e++) rb[v * VW + e] = x[e];
        } else {
#pragma unroll
          for (int e = 0; e < VW; e++) {
            const int rj = TB ? gj + e : gj, rk = TB ? gk : gk + e;
            rb[v * VW + e] = (rj < N && rk < ke) ? (TB ? B[rj + (long)rk * ldb] : B[rk + (long)rj * ldb]) : T(0);
          }
        }
      }
      return;
    }
#pragma unroll
    for (int q = 0; q < NB; q++) {
      int idx = tid + 256 * q;
      int j, kk;
      if (TB) {
        j = idx % BN;
        kk = idx / BN;
      } else {
        kk = idx % BK;
        j = idx / BK;
      }
      int gj = n0 + j, gk = k0 + kk;
      T v = T(0);
      if (gj < N && gk < ke) v = TB ? B[gj + (long)gk * ldb] : B[gk + (long)gj * ldb];
      rb[q] = v;
    }
  }

  __device__ inline void store_a(T *As, int tid, bool va) {
    if (CAN_VA && va) {
#pragma unroll
      for (int v = 0; v < NA / VW; v++) {
        const int vidx = tid + 256 * v;
        if (!TA) {
          const int i = (vidx % (BM / VW)) * VW, kk = vidx / (BM / VW);
          vecT x;
#pragma unroll
          for (int e = 0; e < VW; e++) x[e] = ra[v * VW + e];
          *reinterpret_cast<vecT *>(&As[kk * LDSA + i]) = x;
        } else {
          const int kk = (vidx % (BK / VW)) * VW, i = vidx / (BK / VW);
#pragma unroll
          for (int e = 0; e < VW; e++) As[(kk + e) * LDSA + i] = ra[v * VW + e];
        }
      }
      return;
    }
#pragma unroll
    for (int q = 0; q < NA; q++) {
      int idx = tid + 256 * q;
      int i, kk;
      if (!TA) {
        i = idx % BM;
        kk = idx / BM;
      } else {
        kk = idx % BK;
        i = idx / BK;
      }
      As[kk * LDSA + i] = ra[q];
    }
  }

  __device__ inline void store_b(T *Bs, int tid, bool vb) {
    if (CAN_VB && vb) {
#pragma unroll
      for (int v = 0; v < NB / VW; v++) {
        const int vidx = tid + 256 * v;
        if (TB) {
          const int j = (vidx % (BN / VW)) * VW, kk = vidx / (BN / VW);
          vecT x;
#pragma unroll
          for (int e = 0; e < VW; e++) x[e] = rb[v * VW + e];
          *reinterpret_cast<vecT *>(&Bs[kk * LDSB + j]) = x;
        } else {
          const int kk = (vidx % (BK / VW)) * VW, j = vidx / (BK / VW);
#pragma unroll
          for (int e = 0; e < VW; e++) Bs[(kk + e) * LDSB + j] = rb[v * VW + e];
        }
      }
      return;
    }
#pragma unroll
    for (int q = 0; q < NB; q++) {
      int idx = tid + 256 * q;
      int j, kk;
      if (TB) {
        j = idx % BN;
        kk = idx / BN;
      } else {
        kk = idx % BK;
        j = idx / BK;
      }
      Bs[kk * LDSB + j] = rb[q];
    }
  }

  // A slab that lies wholly inside both operands (the caller checks the tile,
  // the K range and the 16-byte alignment of A): the same register mapping as
  // load(..., va = true, vb) with every bounds test and branch left out, so a
  // step issues a fixed number of memory operations
  __device__ inline void load_full(const T *__restrict__ A, long lda, const T *__restrict__ B, long ldb, int m0, int n0,
                                   int k0, int tid, bool vb) {
    static_assert(CAN_VA, "full-slab loads need the vector A mapping");
#pragma unroll
    for (int v = 0; v < NA / VW; v++) {
      const int vidx = tid + 256 * v;
      const int i = !TA ? (vidx % (BM / VW)) * VW : vidx / (BK / VW);
      const int kk = !TA ? vidx / (BM / VW) : (vidx % (BK / VW)) * VW;
      const int gi = m0 + i, gk = k0 + kk;
      const vecT x = *reinterpret_cast<const vecT *>(TA ? &A[gk + (long)gi * lda] : &A[gi + (long)gk * lda]);
#pragma unroll
      for (int e = 0; e < VW; e++) ra[v * VW + e] = x[e];
    }
    if (CAN_VB && vb) {
#pragma unroll
      for (int v = 0; v < NB / VW; v++) {
        const int vidx = tid + 256 * v;
        const int j = TB ? (vidx % (BN / VW)) * VW : vidx / (BK / VW);
        const int kk = TB ? vidx / (BN / VW) : (vidx % (BK / VW)) * VW;
        const int gj = n0 + j, gk = k0 + kk;
        const vecT x = *reinterpret_cast<const vecT *>(TB ? &B[gj + (long)gk * ldb] : &B[gk + (long)gj * ldb]);
#pragma unroll
        for (int e = 0; e < VW; e++) rb[v * VW + e] = x[e];
      }
    } else {
#pragma unroll
      for (int q = 0; q < NB; q++) {
        const int idx = tid + 256 * q;
        const int j = TB ? idx % BN : idx / BK, kk = TB ? idx / BN : idx % BK;
        const int gj = n0 + j, gk = k0 + kk;
        rb[q] = TB ? B[gj + (long)gk * ldb] : B[gk + (long)gj * ldb];
      }
    }
  }

  // element (scalar) mapping: the split-precision staging reads ra / rb in it
  __device__ inline void load(const T *__restrict__ A, long lda, const T *__restrict__ B, long ldb, int M, int N,
                              int m0, int n0, int k0, int ke, int tid, bool va = false, bool vb = false) {
    load_a(A, lda, M, m0, k0, ke, tid, va);
    load_b(B, ldb, N, n0, k0, ke, tid, vb);
  }
  __device__ inline void store(T *As, T *Bs, int tid, bool va = false, bool vb = false) {
    store_a(As, tid, va);
    store_b(Bs, tid, vb);
  }
};

// Index (0..15) of the MFMA D-matrix row held in accumulator element r of a
// lane: v_mfma_f32_16x16x4f32 keeps rows 4*(lane/16)+r, v_mfma_f64_16x16x4f64
// keeps rows lane/16 + 4*r.  With the (B, A) operand order used here the D row
// is the C column inside the 16-wide tile.
__device__ inline int acc_row(float, int lane, int r) { return 4 * (lane >> 4) + r; }
__device__ inline int acc_row(double, int lane, int r) { return (lane >> 4) + 4 * r; }

__device__ inline f32x4 mfma16(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ inline f64x4 mfma16(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Core: accumulate op(A)[m0:m0+BM, kb:ke] * op(B)[kb:ke, n0:n0+BN] into acc.
// Optional second operand pair of a K-concatenated product
// op(A) op(B) + op(A2) op(B2): K indices from K1 on (a multiple of BK) read
// A2 / B2 at k - K1.
template <typename T>
struct KPair {
  const T *A2 = nullptr;
  long lda2 = 0;
  const T *B2 = nullptr;
  long ldb2 = 0;
  int K1 = 0x7fffffff;
};

template <bool TA, bool TB, int BM, int BN, int WM, int WN, typename T, typename VT, int BK = kBK, bool DEEP = false,
          bool KCAT = false>
__device__ inline void tile_mainloop(const T *__restrict__ A, long lda, const T *__restrict__ B, long ldb, int M,
                                     int N, int m0, int n0, int kb, int ke, T *As0, T *Bs0, VT (&acc)[BM / WM / 16][BN / WN / 16],
                                     KPair<T> kp = KPair<T>()) {
  using L = TileLoader<TA, TB, BM, BN, WM, WN, T, BK>;
  constexpr int MI = BM / WM / 16, NI = BN / WN / 16;
  constexpr int SA = BK * L::LDSA, SB = BK * L::LDSB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  L ld;
  // 16-byte loads where the operands' alignment allows them
  constexpr int VW = L::VW;
  auto vec_ok = [](const T *p, long ld) { return ((reinterpret_cast<unsigned long>(p) & 15) == 0) && (ld % VW == 0); };
  const bool va = vec_ok(A, lda) && (!KCAT || vec_ok(kp.A2, kp.lda2));
  const bool vb = vec_ok(B, ldb) && (!KCAT || vec_ok(kp.B2, kp.ldb2));
  // the K slab at k0 from the pair it lies in
  auto load_slab = [&](int k0) {
    if (!KCAT)
      ld.load(A, lda, B, ldb, M, N, m0, n0, k0, ke, tid, va, vb);
    else if (k0 < kp.K1)
      ld.load(A, lda, B, ldb, M, N, m0, n0, k0, min(ke, kp.K1), tid, va, vb);
    else
      ld.load(kp.A2, kp.lda2, kp.B2, kp.ldb2, M, N, m0, n0, k0 - kp.K1, ke - kp.K1, tid, va, vb);
  };
  auto compute = [&](const T *As, const T *Bs) {
#pragma unroll
    for (int ks = 0; ks < BK / 4; ks++) {
      T a[MI], b[NI];
      const int kr = ks * 4 + (lane >> 4);
#pragma unroll
      for (int mi = 0; mi < MI; mi++) a[mi] = As[kr * L::LDSA + wm * (BM / WM) + mi * 16 + (lane & 15)];
#pragma unroll
      for (int ni = 0; ni < NI; ni++) b[ni] = Bs[kr * L::LDSB + wn * (BN / WN) + ni * 16 + (lane & 15)];
#pragma unroll
      for (int mi = 0; mi < MI; mi++)
#pragma unroll
        for (int ni = 0; ni < NI; ni++) acc[mi][ni] = mfma16(b[ni], a[mi], acc[mi][ni]);
    }
  };
  if constexpr (DEEP && !KCAT && L::CAN_VA) {
    // Interior tiles of a single-pair product: slabs two ahead of the one in
    // the MFMAs, in two register sets that alternate (the loop is unrolled
    // by two), and a step body without branches around its loads, so the
    // wait before an LDS store covers only the older set.  The HBM-bound
    // long-K products need that much in flight.
    const int nfull = (ke - kb) / BK;
    if (va && m0 + BM <= M && n0 + BN <= N && nfull >= 2) {
      L ld1;
      ld.load_full(A, lda, B, ldb, m0, n0, kb, tid, vb);
      ld1.load_full(A, lda, B, ldb, m0, n0, kb + BK, tid, vb);
      ld.store(As0, Bs0, tid, true, vb);
      __syncthreads();
      // stage 0 holds slab t, ld1 carries slab t + 1
      int t = 0;
      for (; t + 3 < nfull; t += 2) {
        ld.load_full(A, lda, B, ldb, m0, n0, kb + (t + 2) * BK, tid, vb);
        compute(As0, Bs0);
        ld1.store(As0 + SA, Bs0 + SB, tid, true, vb);
        __syncthreads();
        ld1.load_full(A, lda, B, ldb, m0, n0, kb + (t + 3) * BK, tid, vb);
        compute(As0 + SA, Bs0 + SB);
        ld.store(As0, Bs0, tid, true, vb);
        __syncthreads();
      }
      const bool three = t + 2 < nfull;
      if (three) ld.load_full(A, lda, B, ldb, m0, n0, kb + (t + 2) * BK, tid, vb);
      compute(As0, Bs0);
      ld1.store(As0 + SA, Bs0 + SB, tid, true, vb);
      __syncthreads();
      compute(As0 + SA, Bs0 + SB);
      int done = t + 2;
      if (three) {
        ld.store(As0, Bs0, tid, true, vb);
        __syncthreads();
        compute(As0, Bs0);
        done = t + 3;
      }
      // the partial slab at the end of the K range, through the checked path
      const int k0 = kb + done * BK;
      if (k0 < ke) {
        __syncthreads();
        ld.load(A, lda, B, ldb, M, N, m0, n0, k0, ke, tid, va, vb);
        ld.store(As0, Bs0, tid, va, vb);
        __syncthreads();
        compute(As0, Bs0);
      }
      __syncthreads();
      return;
    }
  }
  int nk = ke > kb ? (ke - kb + BK - 1) / BK : 0;
  if (nk > 0) {
    load_slab(kb);
    ld.store(As0, Bs0, tid, va, vb);
  }
  __syncthreads();
  for (int t = 0; t < nk; t++) {
    const int cur = t & 1;
    if (t + 1 < nk) load_slab(kb + (t + 1) * BK);
    compute(As0 + cur * SA, Bs0 + cur * SB);
    if (t + 1 < nk) ld.store(As0 + (cur ^ 1) * SA, Bs0 + (cur ^ 1) * SB, tid, va, vb);
    __syncthreads();
  }
}

template <bool TA, bool TB, int BM, int BN, int WM, int WN, int BK = kBK, bool KCAT = false>
__global__ __launch_bounds__(256) void sgemm_kernel(int M, int N, int K, float alpha, const float *__restrict__ A,
                                                    long lda, long sA, const float *__restrict__ B, long ldb, long sB,
                                                    float beta, float *__restrict__ C, long ldc, long sC, int kchunk,
                                                    int ksplit, float *__restrict__ P, long batches,
                                                    KPair<float> kp, long sA2, long sB2) {
  using L = TileLoader<TA, TB, BM, BN, WM, WN, float, BK>;
  constexpr int MI = BM / WM / 16, NI = BN / WN / 16;
  __shared__ __align__(16) float As[2 * BK * L::LDSA];
  __shared__ __align__(16) float Bs[2 * BK * L::LDSB];
  const long bz = blockIdx.z;
  const long batch = bz / ksplit;
  const int kz = (int)(bz % ksplit);
  A += batch * sA;
  B += batch * sB;
  C += batch * sC;
  if (KCAT) {
    kp.A2 += batch * sA2;
    kp.B2 += batch * sB2;
  }
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int kb = kz * kchunk;
  const int ke = min(K, kb + kchunk);
  f32x4 acc[MI][NI];
#pragma unroll
  for (int mi = 0; mi < MI; mi++)
#pragma unroll
    for (int ni = 0; ni < NI; ni++) acc[mi][ni] = f32x4{0.f, 0.f, 0.f, 0.f};
  // the long-K skinny products take the two-slab pipeline
  tile_mainloop<TA, TB, BM, BN, WM, WN, float, f32x4, BK, BM == 256 && (BN == 32 || BN == 64), KCAT>(
      A, lda, B, ldb, M, N, m0, n0, kb, ke, As, Bs, acc, kp);

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave % WM, wn = wave / WM;
#pragma unroll
  for (int mi = 0; mi < MI; mi++) {
    const int row = m0 + wm * (BM / WM) + mi * 16 + (lane & 15);
    if (row >= M) continue;
#pragma unroll
    for (int ni = 0; ni < NI; ni++) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int col = n0 + wn * (BN / WN) + ni * 16 + 4 * (lane >> 4) + r;
        if (col >= N) continue;
        float v = acc[mi][ni][r];
        if (ksplit > 1) {
          P[((long)kz * batches + batch) * (long)M * N + row + (long)col * M] = v;
        } else {
          long o = row + (long)col * ldc;
          C[o] = (beta == 0.f) ? alpha * v : alpha * v + beta * C[o];
        }
      }
    }
  }
}

// Sum of the K-split partials in split order (eight loads in flight, added
// one after another, so the rounding is that of a plain sequential sum).
__global__ void splitk_reduce_kernel(int M, int N, long batches, int ksplit, const float *__restrict__ P, float alpha,
                                     float beta, float *__restrict__ C, long ldc, long sC) {
  const long mn = (long)M * N, total = mn * batches, stride = batches * mn;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    long b = t / mn;
    long rc = t % mn;
    int row = (int)(rc % M);
    long col = rc / M;
    const float *p = P + b * mn + rc;
    float s = 0.f;
    int kz = 0;
    for (; kz + 8 <= ksplit; kz += 8) {
      float v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = p[(long)(kz + u) * stride];
#pragma unroll
      for (int u = 0; u < 8; u++) s += v[u];
    }
    for (; kz < ksplit; kz++) s += p[(long)kz * stride];
    long o = b * sC + row + col * ldc;
    C[o] = (beta == 0.f) ? alpha * s : alpha * s + beta * C[o];
  }
}

// The second operand pair of a K-concatenated product and its batch strides.
struct KCat {
  KPair<float> kp;
  long sA2 = 0, sB2 = 0;
};

template <bool TA, bool TB, int BM, int BN, int WM, int WN, int BK = kBK>
static int launch_sgemm(evd_ctx *ctx, long m, long n, long k, float alpha, const float *A, long lda, long sA,
                        const float *B, long ldb, long sB, float beta, float *C, long ldc, long sC, long batch,
                        long min_chunk = 256, long target_per_cu = 2, const KCat &kc = KCat()) {
  long tiles = ((m + BM - 1) / BM) * ((n + BN - 1) / BN) * shape_batch(ctx, batch);
  int ksplit = 1;
  const long target = target_per_cu * ctx->num_cus;
  // split K while the tiles alone leave the chip idle, keeping each chunk at
  // least min_chunk deep (a deterministic second pass sums the partials)
  if (tiles < target && k >= 2 * min_chunk) {
    long want = (target + tiles - 1) / tiles;
    long maxs = k / min_chunk;
    ksplit = (int)std::min<long>(std::min<long>(want, maxs), 64);
    if (ksplit < 1) ksplit = 1;
  }
  long kchunk = (k + ksplit - 1) / ksplit;
  kchunk = (kchunk + BK - 1) / BK * BK;
  ksplit = (int)((k + kchunk - 1) / kchunk);
  if (ksplit < 1) ksplit = 1;
  float *P = nullptr;
  ArenaFrame frame(ctx);
  if (ksplit > 1) {
    P = arena_alloc_n<float>(ctx, (size_t)ksplit * batch * m * n);
    if (!P) return EVD_E_NOMEM;
  }
  dim3 grid((unsigned)((m + BM - 1) / BM), (unsigned)((n + BN - 1) / BN), (unsigned)(batch * ksplit));
  if (kc.kp.A2)
    sgemm_kernel<TA, TB, BM, BN, WM, WN, BK, true><<<grid, 256, 0, ctx->stream>>>(
        (int)m, (int)n, (int)k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, (int)kchunk, ksplit, P, batch, kc.kp,
        kc.sA2, kc.sB2);
  else
    sgemm_kernel<TA, TB, BM, BN, WM, WN, BK, false><<<grid, 256, 0, ctx->stream>>>(
        (int)m, (int)n, (int)k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, (int)kchunk, ksplit, P, batch, kc.kp,
        kc.sA2, kc.sB2);
  EVD_LAUNCHED(ctx);
  if (ksplit > 1) {
    long total = m * n * batch;
    unsigned g = (unsigned)std::min<long>((total + 255) / 256, 4096);
    splitk_reduce_kernel<<<g, 256, 0, ctx->stream>>>((int)m, (int)n, batch, ksplit, P, alpha, beta, C, ldc, sC);
    EVD_LAUNCHED(ctx);
  }
  return 0;
}

template <bool TA, bool TB>
static int sgemm_dispatch(evd_ctx *ctx, long m, long n, long k, float alpha, const float *A, long lda, long sA,
                          const float *B, long ldb, long sB, float beta, float *C, long ldc, long sC, long batch,
                          const KCat &kc = KCat()) {
  const long target = 2L * ctx->num_cus;
  if (n <= 32 && m > 64) {
    // tall-skinny: the row tile is chosen for parallelism first (the K
    // dimension of these products is often short), then K is split; the
    // long-K products (Z = A W over the trailing matrix, HBM-bound) split
    // to 4 workgroups per CU, which keeps more of A in flight (2 per CU:
    // SBR 279-283 ms at n = 16384, 4: 272 ms)
    const long hb = shape_batch(ctx, batch);
    const long t256 = (m + 255) / 256 * hb, t128 = (m + 127) / 128 * hb;
    if (t256 >= target / 2 || k >= 4096)
      return launch_sgemm<TA, TB, 256, 32, 4, 1>(ctx, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch,
                                                 256, ctx->tune.skinny_split > 0 ? ctx->tune.skinny_split : 4, kc);
    if (t128 >= target / 2)
      return launch_sgemm<TA, TB, 128, 32, 4, 1>(ctx, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch,
                                                 128, 2, kc);
    return launch_sgemm<TA, TB, 64, 32, 4, 1>(ctx, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, 128,
                                              2, kc);
  }
  if (n <= 64 && m > 64 && k >= 4096) {
    // b = 64 skinny long-K products (Z = A W): 256 x 64 tiles on the
    // two-slab pipeline, split to 4 workgroups per CU like the b = 32 form
    // (64 x 64 tiles moved 2.4 TB/s at n = 32768)
    return launch_sgemm<TA, TB, 256, 64, 4, 1>(ctx, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch,
                                               256, ctx->tune.skinny_split > 0 ? ctx->tune.skinny_split : 4, kc);
  }
  if (m <= 32 && n > 64)
    return launch_sgemm<TA, TB, 32, 256, 1, 4>(ctx, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch,
                                               256, 2, kc);
  long big_tiles = ((m + 127) / 128) * ((n + 127) / 128) * shape_batch(ctx, batch);
  // the gemm_tile knob forces 128 x 128 (1) or 64 x 64 (2) output tiles
  if (ctx->tune.gemm_tile == 1 ||
      (ctx->tune.gemm_tile == 0 && (big_tiles >= target || (m > 64 && n > 64 && k < 1024))))
    return launch_sgemm<TA, TB, 128, 128, 2, 2>(ctx, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch,
                                                256, 2, kc);
  return launch_sgemm<TA, TB, 64, 64, 2, 2>(ctx, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, 128,
                                            2, kc);
}

__global__ void scale_kernel(int M, int N, float beta, float *C, long ldc, long sC, long batch) {
  long total = (long)M * N * batch;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    long b = t / ((long)M * N);
    long rc = t % ((long)M * N);
    long o = b * sC + (rc % M) + (rc / M) * ldc;
    C[o] = beta == 0.f ? 0.f : beta * C[o];
  }
}

static int sgemm_split(evd_ctx *ctx, bool TA, bool TB, long m, long n, long k, float alpha, const float *A, long lda,
                       long sA, const float *B, long ldb, long sB, float beta, float *C, long ldc, long sC, long batch,
                       bool *handled);

// Large plain FP32 products (an output of >= 2^22 elements, i.e. >= 256
// tiles of 128 x 128, and >= 2^31 flops per problem) go to rocBLAS on the
// context's handle: its tuned kernels run these shapes at 126-154 TFLOP/s on
// MI355X against about 100 for the tile kernels here.  Products with a
// small output and a long K stay here: without atomics (needed for
// reproducible bits) the library does not split K, and a 256 x 256 x 16384
// product then ran on a handful of workgroups.  A batch is one call per
// problem, so every problem's arithmetic (the library's kernel choice for
// that shape) is that of a single product.  16-bit split products stay here.
static bool library_gemm(const evd_ctx *ctx, long m, long n, long k) {
  return ctx->blas && !ctx->tune.own_gemm && (double)m * n >= 4194304.0 && 2.0 * m * n * k >= 2147483648.0 &&
         m < (1L << 31) && n < (1L << 31) && k < (1L << 31);
}

static int sgemm_library(evd_ctx *ctx, bool TA, bool TB, long m, long n, long k, float alpha, const float *A,
                         long lda, long sA, const float *B, long ldb, long sB, float beta, float *C, long ldc, long sC,
                         long batch) {
  const rocblas_operation oa = TA ? rocblas_operation_transpose : rocblas_operation_none;
  const rocblas_operation ob = TB ? rocblas_operation_transpose : rocblas_operation_none;
  // the stream of this call (a capture swaps the context's stream)
  if (rocblas_set_stream(ctx->blas, ctx->stream) != rocblas_status_success)
    return set_error(ctx, EVD_E_HIP, "rocblas_set_stream failed");
  for (long q = 0; q < batch; q++) {
    const rocblas_status st = rocblas_sgemm(ctx->blas, oa, ob, (rocblas_int)m, (rocblas_int)n, (rocblas_int)k, &alpha,
                                            A + q * sA, (rocblas_int)lda, B + q * sB, (rocblas_int)ldb, &beta,
                                            C + q * sC, (rocblas_int)ldc);
    if (st != rocblas_status_success)
      return set_error(ctx, EVD_E_HIP, "rocblas_sgemm(%ld x %ld x %ld) failed: status %d", m, n, k, (int)st);
  }
  return 0;
}

int sgemm(evd_ctx *ctx, char ta, char tb, long m, long n, long k, float alpha, const float *A, long lda,
          const float *B, long ldb, float beta, float *C, long ldc, long batch, long sA, long sB, long sC) {
  if (m <= 0 || n <= 0 || batch <= 0) return 0;
  if (k <= 0 || alpha == 0.f) {
    if (beta == 1.f) return 0;
    long total = m * n * batch;
    scale_kernel<<<(unsigned)std::min<long>((total + 255) / 256, 4096), 256, 0, ctx->stream>>>((int)m, (int)n, beta, C,
                                                                                             ldc, sC, batch);
    EVD_LAUNCHED(ctx);
    return 0;
  }
  bool TA = (ta == 'T' || ta == 't'), TB = (tb == 'T' || tb == 't');
  if (ctx->precision == EVD_PREC_FP32 && library_gemm(ctx, m, n, k) && lda < (1L << 31) && ldb < (1L << 31) &&
      ldc < (1L << 31))
    return sgemm_library(ctx, TA, TB, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
  if (ctx->precision != EVD_PREC_FP32) {
    bool handled = false;
    const int r = sgemm_split(ctx, TA, TB, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, &handled);
    if (handled) return r;
  }
  if (!TA && !TB) return sgemm_dispatch<false, false>(ctx, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
  if (TA && !TB) return sgemm_dispatch<true, false>(ctx, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
  if (!TA && TB) return sgemm_dispatch<false, true>(ctx, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
  return sgemm_dispatch<true, true>(ctx, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
}

// C = alpha (op(A1) op(B1) + op(A2) op(B2)) + beta C, the two products (inner
// dimensions k1 and k2, same transposition) summed in one K loop: one launch
// and one K split instead of two.  Falls back to two products when k1 is not
// a multiple of the K slab, for shapes the split-precision modes run on 16-bit
// MFMA (the large square-ish ones) and for products that go to the library.
int sgemm_kcat(evd_ctx *ctx, char ta, char tb, long m, long n, long k1, long k2, float alpha, const float *A1,
               long lda1, const float *B1, long ldb1, const float *A2, long lda2, const float *B2, long ldb2, float beta,
               float *C, long ldc, long batch, long sA1, long sB1, long sA2, long sB2, long sC) {
  if (m <= 0 || n <= 0 || batch <= 0) return 0;
  const long k = k1 + k2;
  const bool TA = (ta == 'T' || ta == 't'), TB = (tb == 'T' || tb == 't');
  const bool split_shape = ctx->precision != EVD_PREC_FP32 && m > 64 && n > 64 && k >= 64;
  if (k1 <= 0 || k2 <= 0 || k1 % kBK != 0 || alpha == 0.f || split_shape ||
      (ctx->precision == EVD_PREC_FP32 && library_gemm(ctx, m, n, k)) || k >= (1L << 31)) {
    EVD_TRY(sgemm(ctx, ta, tb, m, n, k1, alpha, A1, lda1, B1, ldb1, beta, C, ldc, batch, sA1, sB1, sC));
    return sgemm(ctx, ta, tb, m, n, k2, alpha, A2, lda2, B2, ldb2, 1.f, C, ldc, batch, sA2, sB2, sC);
  }
  KCat kc;
  kc.kp.A2 = A2;
  kc.kp.lda2 = lda2;
  kc.kp.B2 = B2;
  kc.kp.ldb2 = ldb2;
  kc.kp.K1 = (int)k1;
  kc.sA2 = sA2;
  kc.sB2 = sB2;
  if (!TA && !TB)
    return sgemm_dispatch<false, false>(ctx, m, n, k, alpha, A1, lda1, sA1, B1, ldb1, sB1, beta, C, ldc, sC, batch, kc);
  if (TA && !TB)
    return sgemm_dispatch<true, false>(ctx, m, n, k, alpha, A1, lda1, sA1, B1, ldb1, sB1, beta, C, ldc, sC, batch, kc);
  if (!TA && TB)
    return sgemm_dispatch<false, true>(ctx, m, n, k, alpha, A1, lda1, sA1, B1, ldb1, sB1, beta, C, ldc, sC, batch, kc);
  return sgemm_dispatch<true, true>(ctx, m, n, k, alpha, A1, lda1, sA1, B1, ldb1, sB1, beta, C, ldc, sC, batch, kc);
}

// ----------------------------------------------------------------- SYR2K
// Lower tiles (ti >= tj) of C; the K loop runs over [A|B] against [B|A].
// blockIdx.y = problem of a same-size batch (operands at the strides).
__global__ __launch_bounds__(256) void syr2k_kernel(int n, int k, float alpha, const float *__restrict__ A, long lda,
                                                    const float *__restrict__ B, long ldb, float beta,
                                                    float *__restrict__ C, long ldc, long sA, long sB, long sC) {
  constexpr int BM = 128, BN = 128, WM = 2, WN = 2, MI = 4, NI = 4, BK = kBK;
  A += blockIdx.y * sA;
  B += blockIdx.y * sB;
  C += blockIdx.y * sC;
  using L = TileLoader<false, true, BM, BN, WM, WN, float, BK>;
  __shared__ __align__(16) float As[2 * BK * L::LDSA];
  __shared__ __align__(16) float Bs[2 * BK * L::LDSB];
  const long t = blockIdx.x;
  long ti = (long)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= t) ti++;
  while (ti * (ti + 1) / 2 > t) ti--;
  const long tj = t - ti * (ti + 1) / 2;
  const int m0 = (int)ti * BM, n0 = (int)tj * BN;
  f32x4 acc[MI][NI];
#pragma unroll
  for (int mi = 0; mi < MI; mi++)
#pragma unroll
    for (int ni = 0; ni < NI; ni++) acc[mi][ni] = f32x4{0.f, 0.f, 0.f, 0.f};
  // A_I * B_J^T  (op(B)(kk, j) = B[j + kk*ldb])
  tile_mainloop<false, true, BM, BN, WM, WN, float, f32x4, BK>(A, lda, B, ldb, n, n, m0, n0, 0, k, As, Bs, acc);
  // B_I * A_J^T
  tile_mainloop<false, true, BM, BN, WM, WN, float, f32x4, BK>(B, ldb, A, lda, n, n, m0, n0, 0, k, As, Bs, acc);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave % WM, wn = wave / WM;
#pragma unroll
  for (int mi = 0; mi < MI; mi++) {
    const int row = m0 + wm * (BM / WM) + mi * 16 + (lane & 15);
    if (row >= n) continue;
#pragma unroll
    for (int ni = 0; ni < NI; ni++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int col = n0 + wn * (BN / WN) + ni * 16 + 4 * (lane >> 4) + r;
        if (col > row || col >= n) continue;
        long o = row + (long)col * ldc;
        float v = (beta == 0.f) ? alpha * acc[mi][ni][r] : alpha * acc[mi][ni][r] + beta * C[o];
        C[o] = v;
        if (col != row) C[col + (long)row * ldc] = v;
      }
  }
}

static int syr2k_split(evd_ctx *ctx, long n, long k, float alpha, const float *A, long lda, const float *B, long ldb,
                       float beta, float *C, long ldc);

int syr2k_lower_mirror(evd_ctx *ctx, long n, long k, float alpha, const float *A, long lda, const float *B, long ldb,
                       float beta, float *C, long ldc) {
  return syr2k_lower_mirror_batched(ctx, 1, n, k, alpha, A, lda, 0, B, ldb, 0, beta, C, ldc, 0);
}

int syr2k_lower_mirror_batched(evd_ctx *ctx, int count, long n, long k, float alpha, const float *A, long lda,
                               long sA, const float *B, long ldb, long sB, float beta, float *C, long ldc, long sC) {
  if (n <= 0 || count <= 0) return 0;
  if (ctx->precision != EVD_PREC_FP32) {
    for (int q = 0; q < count; q++)
      EVD_TRY(syr2k_split(ctx, n, k, alpha, A + q * sA, lda, B + q * sB, ldb, beta, C + q * sC, ldc));
    return 0;
  }
  long nt = (n + 127) / 128;
  long tiles = nt * (nt + 1) / 2;
  syr2k_kernel<<<dim3((unsigned)tiles, (unsigned)count), 256, 0, ctx->stream>>>((int)n, (int)k, alpha, A, lda, B, ldb,
                                                                                 beta, C, ldc, sA, sB, sC);
  EVD_LAUNCHED(ctx);
  return 0;
}

// ------------------------------------------------------- split 16-bit MFMA
// Mixed-precision products with FP32 operands in HBM: while a tile is staged
// into LDS every FP32 value is split into NT 16-bit terms, and the MFMA sums
// the leading cross products in FP32:
//   FP16:  x = hi + lo            (22 significant bits), hh + hl + lh
//   BF16:  x = hi + mid + lo      (24 significant bits), hh + hm + mh + hl + lh + mm
// The neglected products are below 2^-22 (FP16) / 2^-24 (BF16) of |x||y|, so a
// product keeps about FP32 accuracy while running on the 16-bit matrix cores.
// LDS holds each term k-contiguous per row ([row][16 k], padded to 20), so a
// lane reads its 4 consecutive k of one row with a single 8-byte load, the
// operand layout of v_mfma_f32_16x16x16{f16,bf16}.  Two cross products of one
// k-slab share a v_mfma_f32_16x16x32_{f16,bf16}: its 8 K slots per lane take
// the 4 k of one product followed by the 4 k of the other, e.g. A = [lo, hi]
// against B = [hi, hi] sums lo*hi + hi*hi.  On gfx950 the 16x16x32 forms take
// the cycles of one 16x16x16, so FP16 issues 1.5 and BF16 3 instead of 3 and
// 6 instructions per slab.
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short bf16x8s __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
constexpr int kLDK = 20;

struct SplitF16 {
  static constexpr int NT = 2;
  static constexpr int kMinBlocks = 3;  // workgroups per CU the registers are held to
  typedef _Float16 elem;
  typedef f16x4 vec;
  __device__ static inline void split(float x, elem (&t)[NT]) {
    t[0] = (_Float16)x;
    t[1] = (_Float16)(x - (float)t[0]);
  }
  // the products of one slab, smallest first: A = [hi, lo] against
  // B = [lo, hi] sums hi*lo + lo*hi, then hi*hi on the 16x16x16 form (its
  // operands are halves of the same register tuples, so nothing is copied)
  __device__ static inline f32x4 step(const vec (&b)[NT], const vec (&a)[NT], f32x4 c) {
    const f16x8 aa = __builtin_shufflevector(a[0], a[1], 0, 1, 2, 3, 4, 5, 6, 7);
    const f16x8 bb = __builtin_shufflevector(b[1], b[0], 0, 1, 2, 3, 4, 5, 6, 7);
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(bb, aa, c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x16f16(__builtin_shufflevector(bb, bb, 4, 5, 6, 7),
                                                 __builtin_shufflevector(aa, aa, 0, 1, 2, 3), c, 0, 0, 0);
  }
};

__device__ inline short f32_to_bf16_rne(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
__device__ inline float bf16_to_f32(short h) { return __uint_as_float(((unsigned)(unsigned short)h) << 16); }

struct SplitBF16 {
  static constexpr int NT = 3;
  static constexpr int kMinBlocks = 2;
  typedef short elem;
  typedef bf16x4 vec;
  __device__ static inline void split(float x, elem (&t)[NT]) {
    t[0] = f32_to_bf16_rne(x);
    const float r1 = x - bf16_to_f32(t[0]);
    t[1] = f32_to_bf16_rne(r1);
    t[2] = f32_to_bf16_rne(r1 - bf16_to_f32(t[1]));
  }
  // (A term, B term) pairs of the three 16x16x32 products, smallest first:
  // lo*hi + mid*mid, hi*lo + mid*hi, hi*mid + hi*hi
  __device__ static inline f32x4 mfma2(vec b0, vec b1, vec a0, vec a1, f32x4 c) {
    const bf16x8s bb = __builtin_shufflevector(b0, b1, 0, 1, 2, 3, 4, 5, 6, 7);
    const bf16x8s aa = __builtin_shufflevector(a0, a1, 0, 1, 2, 3, 4, 5, 6, 7);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, bb), __builtin_bit_cast(bf16x8, aa), c,
                                                   0, 0, 0);
  }
  __device__ static inline f32x4 step(const vec (&b)[NT], const vec (&a)[NT], f32x4 c) {
    c = mfma2(b[0], b[1], a[2], a[1], c);
    c = mfma2(b[2], b[0], a[0], a[1], c);
    return mfma2(b[1], b[0], a[0], a[0], c);
  }
  // the same six products on the 16x16x16 form (fewer registers: the SYR2K
  // kernel, which holds two accumulator sets, keeps two waves per SIMD)
  __device__ static inline f32x4 step16(const vec (&b)[NT], const vec (&a)[NT], f32x4 c) {
    constexpr int pa[6] = {1, 2, 0, 1, 0, 0}, pb[6] = {1, 0, 2, 0, 1, 0};  // mm, lh, hl, mh, hm, hh
#pragma unroll
    for (int p = 0; p < 6; p++) c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(b[pb[p]], a[pa[p]], c, 0, 0, 0);
    return c;
  }
};

template <class S, bool TA, bool TB, int BM, int BN, int WM, int WN>
__device__ inline void store_split(const TileLoader<TA, TB, BM, BN, WM, WN, float> &ld, typename S::elem *As,
                                   typename S::elem *Bs, int tid) {
  using L = TileLoader<TA, TB, BM, BN, WM, WN, float>;
#pragma unroll
  for (int q = 0; q < L::NA; q++) {
    const int idx = tid + 256 * q;
    int i, kk;
    if (!TA) {
      i = idx % BM;
      kk = idx / BM;
    } else {
      kk = idx % kBK;
      i = idx / kBK;
    }
    typename S::elem t[S::NT];
    S::split(ld.ra[q], t);
#pragma unroll
    for (int u = 0; u < S::NT; u++) As[u * BM * kLDK + i * kLDK + kk] = t[u];
  }
#pragma unroll
  for (int q = 0; q < L::NB; q++) {
    const int idx = tid + 256 * q;
    int j, kk;
    if (TB) {
      j = idx % BN;
      kk = idx / BN;
    } else {
      kk = idx % kBK;
      j = idx / kBK;
    }
    typename S::elem t[S::NT];
    S::split(ld.rb[q], t);
#pragma unroll
    for (int u = 0; u < S::NT; u++) Bs[u * BN * kLDK + j * kLDK + kk] = t[u];
  }
}

// split-precision counterpart of tile_mainloop (kBK = 16 k per slab; PAIR:
// two products per 16x16x32 instruction, else S::step16)
template <class S, bool TA, bool TB, int BM, int BN, int WM, int WN, bool PAIR = true>
__device__ inline void tile_mainloop_split(const float *__restrict__ A, long lda, const float *__restrict__ B, long ldb,
                                           int M, int N, int m0, int n0, int kb, int ke, typename S::elem *As0,
                                           typename S::elem *Bs0, f32x4 (&acc)[BM / WM / 16][BN / WN / 16]) {
  using L = TileLoader<TA, TB, BM, BN, WM, WN, float>;
  constexpr int MI = BM / WM / 16, NI = BN / WN / 16;
  constexpr int SA = S::NT * BM * kLDK, SB = S::NT * BN * kLDK;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  L ld;
  const int nk = ke > kb ? (ke - kb + kBK - 1) / kBK : 0;
  if (nk > 0) {
    ld.load(A, lda, B, ldb, M, N, m0, n0, kb, ke, tid);
    store_split<S>(ld, As0, Bs0, tid);
  }
  __syncthreads();
  for (int t = 0; t < nk; t++) {
    const int cur = t & 1;
    if (t + 1 < nk) ld.load(A, lda, B, ldb, M, N, m0, n0, kb + (t + 1) * kBK, ke, tid);
    const typename S::elem *As = As0 + cur * SA;
    const typename S::elem *Bs = Bs0 + cur * SB;
    typename S::vec a[S::NT][MI], b[S::NT][NI];
    const int ko = 4 * (lane >> 4);
#pragma unroll
    for (int u = 0; u < S::NT; u++) {
#pragma unroll
      for (int mi = 0; mi < MI; mi++)
        a[u][mi] = *reinterpret_cast<const typename S::vec *>(
            &As[u * BM * kLDK + (wm * (BM / WM) + mi * 16 + (lane & 15)) * kLDK + ko]);
#pragma unroll
      for (int ni = 0; ni < NI; ni++)
        b[u][ni] = *reinterpret_cast<const typename S::vec *>(
            &Bs[u * BN * kLDK + (wn * (BN / WN) + ni * 16 + (lane & 15)) * kLDK + ko]);
    }
#pragma unroll
    for (int mi = 0; mi < MI; mi++)
#pragma unroll
      for (int ni = 0; ni < NI; ni++) {
        typename S::vec av[S::NT], bv[S::NT];
#pragma unroll
        for (int u = 0; u < S::NT; u++) {
          av[u] = a[u][mi];
          bv[u] = b[u][ni];
        }
        if constexpr (PAIR)
          acc[mi][ni] = S::step(bv, av, acc[mi][ni]);
        else
          acc[mi][ni] = S::step16(bv, av, acc[mi][ni]);
      }
    if (t + 1 < nk) store_split<S>(ld, As0 + (cur ^ 1) * SA, Bs0 + (cur ^ 1) * SB, tid);
    __syncthreads();
  }
}

template <class S, bool TA, bool TB, int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(256, S::kMinBlocks) void sgemm_split_kernel(int M, int N, int K, float alpha,
                                                          const float *__restrict__ A, long lda, long sA,
                                                          const float *__restrict__ B, long ldb, long sB, float beta,
                                                          float *__restrict__ C, long ldc, long sC, int kchunk,
                                                          int ksplit, float *__restrict__ P, long batches) {
  constexpr int MI = BM / WM / 16, NI = BN / WN / 16;
  __shared__ typename S::elem As[2 * S::NT * BM * kLDK];
  __shared__ typename S::elem Bs[2 * S::NT * BN * kLDK];
  const long bz = blockIdx.z;
  const long batch = bz / ksplit;
  const int kz = (int)(bz % ksplit);
  A += batch * sA;
  B += batch * sB;
  C += batch * sC;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int kb = kz * kchunk;
  const int ke = min(K, kb + kchunk);
  f32x4 acc[MI][NI];
#pragma unroll
  for (int mi = 0; mi < MI; mi++)
#pragma unroll
    for (int ni = 0; ni < NI; ni++) acc[mi][ni] = f32x4{0.f, 0.f, 0.f, 0.f};
  tile_mainloop_split<S, TA, TB, BM, BN, WM, WN>(A, lda, B, ldb, M, N, m0, n0, kb, ke, As, Bs, acc);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave % WM, wn = wave / WM;
#pragma unroll
  for (int mi = 0; mi < MI; mi++) {
    const int row = m0 + wm * (BM / WM) + mi * 16 + (lane & 15);
    if (row >= M) continue;
#pragma unroll
    for (int ni = 0; ni < NI; ni++) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int col = n0 + wn * (BN / WN) + ni * 16 + 4 * (lane >> 4) + r;
        if (col >= N) continue;
        const float v = acc[mi][ni][r];
        if (ksplit > 1) {
          P[((long)kz * batches + batch) * (long)M * N + row + (long)col * M] = v;
        } else {
          const long o = row + (long)col * ldc;
          C[o] = (beta == 0.f) ? alpha * v : alpha * v + beta * C[o];
        }
      }
    }
  }
}

template <class S>
__global__ __launch_bounds__(256) void syr2k_split_kernel(int n, int k, float alpha, const float *__restrict__ A,
                                                          long lda, const float *__restrict__ B, long ldb, float beta,
                                                          float *__restrict__ C, long ldc) {
  constexpr int BM = 128, BN = 128, WM = 2, WN = 2, MI = 4, NI = 4;
  __shared__ typename S::elem As[2 * S::NT * BM * kLDK];
  __shared__ typename S::elem Bs[2 * S::NT * BN * kLDK];
  const long t = blockIdx.x;
  long ti = (long)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= t) ti++;
  while (ti * (ti + 1) / 2 > t) ti--;
  const long tj = t - ti * (ti + 1) / 2;
  const int m0 = (int)ti * BM, n0 = (int)tj * BN;
  f32x4 acc[MI][NI];
#pragma unroll
  for (int mi = 0; mi < MI; mi++)
#pragma unroll
    for (int ni = 0; ni < NI; ni++) acc[mi][ni] = f32x4{0.f, 0.f, 0.f, 0.f};
  constexpr bool pair = S::NT == 2;
  tile_mainloop_split<S, false, true, BM, BN, WM, WN, pair>(A, lda, B, ldb, n, n, m0, n0, 0, k, As, Bs, acc);
  tile_mainloop_split<S, false, true, BM, BN, WM, WN, pair>(B, ldb, A, lda, n, n, m0, n0, 0, k, As, Bs, acc);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave % WM, wn = wave / WM;
#pragma unroll
  for (int mi = 0; mi < MI; mi++) {
    const int row = m0 + wm * (BM / WM) + mi * 16 + (lane & 15);
    if (row >= n) continue;
#pragma unroll
    for (int ni = 0; ni < NI; ni++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int col = n0 + wn * (BN / WN) + ni * 16 + 4 * (lane >> 4) + r;
        if (col > row || col >= n) continue;
        const long o = row + (long)col * ldc;
        const float v = (beta == 0.f) ? alpha * acc[mi][ni][r] : alpha * acc[mi][ni][r] + beta * C[o];
        C[o] = v;
        if (col != row) C[col + (long)row * ldc] = v;
      }
  }
}

template <class S, bool TA, bool TB, int BM, int BN, int WM, int WN>
static int launch_sgemm_split(evd_ctx *ctx, long m, long n, long k, float alpha, const float *A, long lda, long sA,
                              const float *B, long ldb, long sB, float beta, float *C, long ldc, long sC, long batch) {
  const long tiles = ((m + BM - 1) / BM) * ((n + BN - 1) / BN) * shape_batch(ctx, batch);
  int ksplit = 1;
  const long target = 2L * ctx->num_cus;
  if (tiles < target && k >= 512) {
    const long want = (target + tiles - 1) / tiles;
    ksplit = (int)std::min<long>(std::min<long>(want, k / 256), 64);
    if (ksplit < 1) ksplit = 1;
  }
  long kchunk = (k + ksplit - 1) / ksplit;
  kchunk = (kchunk + kBK - 1) / kBK * kBK;
  ksplit = (int)((k + kchunk - 1) / kchunk);
  if (ksplit < 1) ksplit = 1;
  float *P = nullptr;
  ArenaFrame frame(ctx);
  if (ksplit > 1) {
    P = arena_alloc_n<float>(ctx, (size_t)ksplit * batch * m * n);
    if (!P) return EVD_E_NOMEM;
  }
  dim3 grid((unsigned)((m + BM - 1) / BM), (unsigned)((n + BN - 1) / BN), (unsigned)(batch * ksplit));
  sgemm_split_kernel<S, TA, TB, BM, BN, WM, WN><<<grid, 256, 0, ctx->stream>>>(
      (int)m, (int)n, (int)k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, (int)kchunk, ksplit, P, batch);
  EVD_LAUNCHED(ctx);
  if (ksplit > 1) {
    const long total = m * n * batch;
    const unsigned g = (unsigned)std::min<long>((total + 255) / 256, 4096);
    splitk_reduce_kernel<<<g, 256, 0, ctx->stream>>>((int)m, (int)n, batch, ksplit, P, alpha, beta, C, ldc, sC);
    EVD_LAUNCHED(ctx);
  }
  return 0;
}

template <class S>
static int sgemm_split_t(evd_ctx *ctx, bool TA, bool TB, long m, long n, long k, float alpha, const float *A, long lda,
                         long sA, const float *B, long ldb, long sB, float beta, float *C, long ldc, long sC,
                         long batch, bool *handled) {
  *handled = true;
  // The long-K tall-skinny products (A W with K = the trailing size, the
  // long reductions of the ZY updates) stay on the FP32 kernels: they are
  // bound by HBM and latency, and the FP32 kernels' two-slab pipeline runs
  // them faster than the split 16-bit staging (n = 16384 EVD: FP16 form
  // 0.537 -> 0.532 s, BF16 form 0.694 -> 0.638 s).  Large square-ish
  // products (back transformation, WY merges) are the split kernels' work.
  if (m > 64 && n > 64 && k >= 64) {
    if (!TA && !TB) return launch_sgemm_split<S, false, false, 128, 128, 2, 2>(ctx, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
    if (TA && !TB) return launch_sgemm_split<S, true, false, 128, 128, 2, 2>(ctx, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
    if (!TA && TB) return launch_sgemm_split<S, false, true, 128, 128, 2, 2>(ctx, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
  }
  *handled = false;  // short or odd shapes stay on the FP32 kernels
  return 0;
}

static int sgemm_split(evd_ctx *ctx, bool TA, bool TB, long m, long n, long k, float alpha, const float *A, long lda,
                       long sA, const float *B, long ldb, long sB, float beta, float *C, long ldc, long sC, long batch,
                       bool *handled) {
  if (ctx->precision == EVD_PREC_FP16)
    return sgemm_split_t<SplitF16>(ctx, TA, TB, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, handled);
  return sgemm_split_t<SplitBF16>(ctx, TA, TB, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, handled);
}

static int syr2k_split(evd_ctx *ctx, long n, long k, float alpha, const float *A, long lda, const float *B, long ldb,
                       float beta, float *C, long ldc) {
  const long nt = (n + 127) / 128;
  const long tiles = nt * (nt + 1) / 2;
  if (ctx->precision == EVD_PREC_FP16)
    syr2k_split_kernel<SplitF16><<<(unsigned)tiles, 256, 0, ctx->stream>>>((int)n, (int)k, alpha, A, lda, B, ldb, beta,
                                                                          C, ldc);
  else
    syr2k_split_kernel<SplitBF16><<<(unsigned)tiles, 256, 0, ctx->stream>>>((int)n, (int)k, alpha, A, lda, B, ldb,
                                                                           beta, C, ldc);
  EVD_LAUNCHED(ctx);
  return 0;
}

// ----------------------------------------------------------------- FP64
// 64 x 64 output tile on 4 waves (2 x 2, each 32 x 32 of 16 x 16 MFMA
// tiles); a 128 x 128 tile (1 workgroup per CU by registers) measured
// slower (41.9 against 47.6 TFLOP/s at 8192^3).
template <bool TA, bool TB, int BM>
__global__ __launch_bounds__(256) void dgemm_kernel(int M, int N, int K, double alpha, const double *__restrict__ A,
                                                    long lda, const double *__restrict__ B, long ldb, double beta,
                                                    double *__restrict__ C, long ldc) {
  constexpr int BN = BM, WM = 2, WN = 2, MI = BM / WM / 16, NI = BN / WN / 16;
  using L = TileLoader<TA, TB, BM, BN, WM, WN, double>;
  __shared__ __align__(16) double As[2 * kBK * L::LDSA];
  __shared__ __align__(16) double Bs[2 * kBK * L::LDSB];
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  f64x4 acc[MI][NI];
#pragma unroll
  for (int mi = 0; mi < MI; mi++)
#pragma unroll
    for (int ni = 0; ni < NI; ni++) acc[mi][ni] = f64x4{0., 0., 0., 0.};
  tile_mainloop<TA, TB, BM, BN, WM, WN, double, f64x4>(A, lda, B, ldb, M, N, m0, n0, 0, K, As, Bs, acc);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave % WM, wn = wave / WM;
#pragma unroll
  for (int mi = 0; mi < MI; mi++) {
    const int row = m0 + wm * (BM / WM) + mi * 16 + (lane & 15);
    if (row >= M) continue;
#pragma unroll
    for (int ni = 0; ni < NI; ni++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int col = n0 + wn * (BN / WN) + ni * 16 + acc_row(0.0, lane, r);
        if (col >= N) continue;
        long o = row + (long)col * ldc;
        C[o] = (beta == 0.0) ? alpha * acc[mi][ni][r] : alpha * acc[mi][ni][r] + beta * C[o];
      }
  }
}

template <bool TA, bool TB>
static void launch_dgemm(evd_ctx *ctx, long m, long n, long k, double alpha, const double *A, long lda,
                         const double *B, long ldb, double beta, double *C, long ldc) {
  dim3 grid((unsigned)((m + 63) / 64), (unsigned)((n + 63) / 64));
  dgemm_kernel<TA, TB, 64><<<grid, 256, 0, ctx->stream>>>((int)m, (int)n, (int)k, alpha, A, lda, B, ldb, beta, C, ldc);
}

int dgemm(evd_ctx *ctx, char ta, char tb, long m, long n, long k, double alpha, const double *A, long lda,
          const double *B, long ldb, double beta, double *C, long ldc) {
  if (m <= 0 || n <= 0) return 0;
  bool TA = (ta == 'T' || ta == 't'), TB = (tb == 'T' || tb == 't');
  if (k <= 0) {
    // C = beta C through a zero-K product
    k = 0;
  }
  if (!TA && !TB)
    launch_dgemm<false, false>(ctx, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  else if (TA && !TB)
    launch_dgemm<true, false>(ctx, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  else if (!TA && TB)
    launch_dgemm<false, true>(ctx, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  else
    launch_dgemm<true, true>(ctx, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  EVD_LAUNCHED(ctx);
  return 0;
}

// Grouped FP64 GEMM: one workgroup per 64x64 output tile of one problem.
__global__ __launch_bounds__(256) void dgemm_grouped_kernel(const DGemmProblem *__restrict__ probs, int nprob,
                                                            int total_tiles) {
  constexpr int BM = 64, BN = 64, WM = 2, WN = 2, MI = 2, NI = 2;
  using L = TileLoader<false, false, BM, BN, WM, WN, double>;
  __shared__ __align__(16) double As[2 * kBK * L::LDSA];
  __shared__ __align__(16) double Bs[2 * kBK * L::LDSB];
  const int t = blockIdx.x;
  if (t >= total_tiles) return;
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (probs[mid].tile_begin <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  const DGemmProblem p = probs[lo];
  const int local = t - p.tile_begin;
  const int tm = (p.m + BM - 1) / BM;
  if (p.m <= 0 || p.n <= 0) return;
  const int m0 = (local % tm) * BM, n0 = (local / tm) * BN;
  if (n0 >= p.n) return;
  f64x4 acc[MI][NI];
#pragma unroll
  for (int mi = 0; mi < MI; mi++)
#pragma unroll
    for (int ni = 0; ni < NI; ni++) acc[mi][ni] = f64x4{0., 0., 0., 0.};
  tile_mainloop<false, false, BM, BN, WM, WN, double, f64x4>(p.A, p.lda, p.B, p.ldb, p.m, p.n, m0, n0, 0, p.k, As, Bs,
                                                              acc);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave % WM, wn = wave / WM;
#pragma unroll
  for (int mi = 0; mi < MI; mi++) {
    const int row = m0 + wm * (BM / WM) + mi * 16 + (lane & 15);
    if (row >= p.m) continue;
#pragma unroll
    for (int ni = 0; ni < NI; ni++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int col = n0 + wn * (BN / WN) + ni * 16 + acc_row(0.0, lane, r);
        if (col >= p.n) continue;
        const int oc = p.colmap ? p.colmap[col] : col;
        p.C[row + (long)oc * p.ldc] = acc[mi][ni][r];
      }
  }
}

int dgemm_grouped(evd_ctx *ctx, const DGemmProblem *dev_problems, int nproblems, int total_tiles) {
  if (nproblems <= 0 || total_tiles <= 0) return 0;
  dgemm_grouped_kernel<<<(unsigned)total_tiles, 256, 0, ctx->stream>>>(dev_problems, nproblems, total_tiles);
  EVD_LAUNCHED(ctx);
  return 0;
}

}  // namespace evd

using namespace evd;

extern "C" int evd_sgemm(evd_ctx *ctx, char ta, char tb, long m, long n, long k, float alpha, const float *A, long lda,
                         const float *B, long ldb, float beta, float *C, long ldc) {
  if (!ctx) return -1;
  EVD_ARG(ctx, ta == 'N' || ta == 'T' || ta == 'n' || ta == 't', 2);
  EVD_ARG(ctx, tb == 'N' || tb == 'T' || tb == 'n' || tb == 't', 3);
  EVD_ARG(ctx, m >= 0, 4);
  EVD_ARG(ctx, n >= 0, 5);
  EVD_ARG(ctx, k >= 0, 6);
  EVD_ARG(ctx, lda >= 1, 9);
  EVD_ARG(ctx, ldb >= 1, 11);
  EVD_ARG(ctx, ldc >= std::max<long>(1, m), 14);
  // split-K partials live in the arena
  size_t need = 0;
  if (ctx->arena_top == 0) {
    need = (size_t)64 * m * n * sizeof(float);
    if (need > ((size_t)1 << 30)) need = (size_t)1 << 30;
    EVD_TRY(arena_ensure(ctx, need));
  }
  return sgemm(ctx, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

extern "C" int evd_dgemm(evd_ctx *ctx, char ta, char tb, long m, long n, long k, double alpha, const double *A,
                         long lda, const double *B, long ldb, double beta, double *C, long ldc) {
  if (!ctx) return -1;
  EVD_ARG(ctx, ta == 'N' || ta == 'T' || ta == 'n' || ta == 't', 2);
  EVD_ARG(ctx, tb == 'N' || tb == 'T' || tb == 'n' || tb == 't', 3);
  EVD_ARG(ctx, m >= 0, 4);
  EVD_ARG(ctx, n >= 0, 5);
  EVD_ARG(ctx, k >= 0, 6);
  EVD_ARG(ctx, ldc >= std::max<long>(1, m), 14);
  return dgemm(ctx, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
