// Back transformation of the bulge-chasing reflectors (the role of
// my_BC_back_trans_v8_10_noBandU / BC_kernel_computerQ_1Col_V8_10_noBandU,
// api/BC_backTransV8_10_noBandU.cu:60-313).
//
// Reflector (s, j) = I - u u^T acts on rows [s+1+j*b, s+1+(j+1)*b) and is
// stored in U[s+1+j*b.., s].  Two reflectors that share rows must be applied
// in sweep order (lower sweep first for Q <- H_K..H_1 Q, higher sweep first
// for Q <- H_1..H_K Q); reflectors of one sweep are disjoint.
//
// Blocking: sweeps are cut into blocks of BS consecutive sweeps; block (S, j)
// holds the BS reflectors of step j, a staircase covering BS + b - 1 rows.
// For Q <- H_1..H_K Q (reverse) blocks run S descending, j ascending; for
// Q <- H_K..H_1 Q (forward) S ascending, j descending.
//
// MI355X mapping: one workgroup owns 64 columns (lane = column, the column
// window of BS + b rows lives in VGPRs with static indices; reflector entries
// are wave-uniform and come through scalar loads).  NW waves of the workgroup
// form a pipeline over NW consecutive sweep blocks: wave k lags wave k-1 by one
// step and receives the b rows wave k-1 has finished through an LDS double
// buffer, so the column block is streamed from HBM once per NW*BS sweeps and
// each SIMD runs FMA work.  Rows enter and leave HBM through an LDS transpose:
// a wave instruction moves 2 columns x 32 consecutive rows of the
// column-major block (a lane-per-column access would touch 64 cache lines
// with 4 bytes each and leave the L2 as partial-line writes).
#include <type_traits>

#include "common.hpp"

namespace evd {

// Packed reflector blocks.  Block (S, jj) holds the BS reflectors of sweeps
// S*BS .. S*BS+BS-1 at step jj as BS x NB floats, entry [t][i] =
// U[S*BS + 1 + jj*NB + t + i, S*BS + t], and 0 where the sweep or the row does
// not exist (so every block can be applied without checks: a zero reflector
// is the identity).  Sweep block S has J(S) = A + 2 - m S steps, A = (n-2)/NB,
// m = BS/NB, and its blocks start at index prefix(S) = S (A+2) - m S (S-1)/2.
// In this layout reflector t of a block sits at the compile-time offset t*NB
// from the block's base, so its entries reach SGPRs with scalar loads at
// immediate offsets: no per-reflector address registers.
__host__ __device__ inline long pack_steps(long A, long m, long S) { return A + 2 - m * S; }
__host__ __device__ inline long pack_prefix(long A, long m, long S) { return S * (A + 2) - m * S * (S - 1) / 2; }

template <int NB, int BS>
size_t pack_floats(long n) {
  if (n <= 2) return 0;
  const long nsweeps = n - 2, nblk = (nsweeps + BS - 1) / BS, A = (n - 2) / NB;
  return (size_t)pack_prefix(A, BS / NB, nblk) * BS * NB;
}

template <int NB, int BS>
__global__ void bc_pack_kernel(int n, const float *__restrict__ U, long ldu, float *__restrict__ P, int nsweeps) {
  constexpr long m = BS / NB;
  const long A = (n - 2) / NB;
  const long S = blockIdx.y;
  const long base = pack_prefix(A, m, S) * BS * NB;
  const long cnt = pack_steps(A, m, S) * BS * NB;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < cnt; idx += (long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % NB);
    const int t = (int)((idx / NB) % BS);
    const long jj = idx / (NB * BS);
    const long sw = S * BS + t;
    const long row = sw + 1 + jj * NB + i;
    P[base + idx] = (sw < nsweeps && row < n) ? U[row + sw * ldu] : 0.f;
  }
}

// Reflector entries reach the FMAs through DPP row broadcasts.  A reflector
// of NB entries is held in R = NB/16 VGPRs: lane j of every 16-lane row holds
// entries R*j .. R*j+R-1 (one vector load per lane), and entry i is read as
// register i % R broadcast from lane i / R of each row (row_newbcast).  This
// costs no extra instruction per FMA, returns only NB*4 bytes per row per
// reflector from the L1, and lets the loads run PF reflectors ahead.
// v_fmac with a DPP operand is written as inline asm (the compiler does not
// fold a DPP move into a MAC).  The broadcast source is always a register
// written by a vector load, never by a VALU instruction, so the DPP read has
// no VALU-write hazard.
template <int NB>
struct URegs {
  float r[NB / 16];
};

template <int NB>
__device__ __attribute__((always_inline)) inline void load_u(URegs<NB> &u, const float *__restrict__ pt, int j) {
  if constexpr (NB == 32) {
    const float2 v = reinterpret_cast<const float2 *>(pt)[j];
    u.r[0] = v.x;
    u.r[1] = v.y;
  } else {
    const float4 v = reinterpret_cast<const float4 *>(pt)[j];
    u.r[0] = v.x;
    u.r[1] = v.y;
    u.r[2] = v.z;
    u.r[3] = v.w;
  }
}

// d = u[I] * x
template <int NB, int I>
__device__ __attribute__((always_inline)) inline float mul_bc(const URegs<NB> &u, float x) {
  constexpr int R = NB / 16;
  float d;
  asm("v_mul_f32_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
      : "=v"(d)
      : "v"(u.r[I % R]), "v"(x), "n"(I / R));
  return d;
}
// acc += u[I] * x
template <int NB, int I>
__device__ __attribute__((always_inline)) inline void fma_bc(float &acc, const URegs<NB> &u, float x) {
  constexpr int R = NB / 16;
  asm("v_fmac_f32_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
      : "+v"(acc)
      : "v"(u.r[I % R]), "v"(x), "n"(I / R));
}
// acc -= u[I] * x
template <int NB, int I>
__device__ __attribute__((always_inline)) inline void fms_bc(float &acc, const URegs<NB> &u, float x) {
  constexpr int R = NB / 16;
  asm("v_fmac_f32_dpp %0, %1, -%2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
      : "+v"(acc)
      : "v"(u.r[I % R]), "v"(x), "n"(I / R));
}

template <int NB, int W, int T0, int I>
__device__ __attribute__((always_inline)) inline void dot_terms(float (&d)[4], const float (&z)[W], const URegs<NB> &u) {
  if constexpr (I < NB) {
    if constexpr (I < 4)
      d[I] = mul_bc<NB, I>(u, z[T0 + I]);
    else
      fma_bc<NB, I>(d[I % 4], u, z[T0 + I]);
    dot_terms<NB, W, T0, I + 1>(d, z, u);
  }
}

template <int NB, int W, int T0, int I>
__device__ __attribute__((always_inline)) inline void update_terms(float (&z)[W], const URegs<NB> &u, float dot) {
  if constexpr (I < NB) {
    fms_bc<NB, I>(z[T0 + I], u, dot);
    update_terms<NB, W, T0, I + 1>(z, u, dot);
  }
}

constexpr int kPrefetch = 4;  // reflectors in flight ahead of the one applied

// Apply the BS reflectors of one packed block to the window (template
// recursion keeps every window index static).  ring[TT % kPrefetch] holds
// reflector TT; after use it is refilled with reflector TT + kPrefetch.
template <int NB, int BS, int W, bool REV, int TT>
__device__ __attribute__((always_inline)) inline void apply_packed(float (&z)[W], const float *__restrict__ pb,
                                                                   URegs<NB> (&ring)[kPrefetch], int j) {
  if constexpr (TT < BS) {
    constexpr int t = REV ? (BS - 1 - TT) : TT;
    const URegs<NB> u = ring[TT % kPrefetch];
    if constexpr (TT + kPrefetch < BS) {
      constexpr int tn = REV ? (BS - 1 - (TT + kPrefetch)) : (TT + kPrefetch);
      load_u<NB>(ring[TT % kPrefetch], pb + tn * NB, j);
    }
    float d[4];
    dot_terms<NB, W, t, 0>(d, z, u);
    const float dot = (d[0] + d[1]) + (d[2] + d[3]);
    update_terms<NB, W, t, 0>(z, u, dot);
    apply_packed<NB, BS, W, REV, TT + 1>(z, pb, ring, j);
  }
}

template <int NB, int BS, int W, bool REV>
__device__ __attribute__((always_inline)) inline void apply_packed_block(float (&z)[W], const float *__restrict__ pb,
                                                                         int lane) {
  const int j = lane & 15;
  URegs<NB> ring[kPrefetch];
#pragma unroll
  for (int q = 0; q < kPrefetch; q++) load_u<NB>(ring[q], pb + (REV ? (BS - 1 - q) : q) * NB, j);
  apply_packed<NB, BS, W, REV, 0>(z, pb, ring, j);
}

// LDS row stride of a 32-row x 64-column stage (padded against bank conflicts)
constexpr int kStageLd = 65;

// f(std::integral_constant<int, C>) for C = B, B + 32, ... < E: the window
// index stays a compile-time constant however large the body is
template <int B, int E, typename F>
__device__ __attribute__((always_inline)) inline void for_chunks(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    for_chunks<B + 32, E>(f);
  }
}

__device__ __attribute__((always_inline)) inline void lds_order() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// 32 rows [row0, row0 + 32) of the wave's 64 columns from Z into stage[r][c];
// one wave instruction reads 2 columns x 32 consecutive rows.
__device__ __attribute__((always_inline)) inline void stage_load32(float *stage, const float *__restrict__ Z, long ldz,
                                                                   int col0, int ncols, int row0, int n, int lane) {
  const int r = lane & 31, half = lane >> 5;
  const int gr = row0 + r;
#pragma unroll 8
  for (int it = 0; it < 32; it++) {
    const int c = 2 * it + half, gc = col0 + c;
    stage[r * kStageLd + c] = (gc < ncols && gr < n) ? Z[(long)gc * ldz + gr] : 0.f;
  }
  lds_order();
}

// stage[r][c] for rows [row0, row0 + 32) back to Z, rows >= n dropped
__device__ __attribute__((always_inline)) inline void stage_store32(const float *stage, float *__restrict__ Z, long ldz,
                                                                    int col0, int ncols, int row0, int n, int lane) {
  lds_order();
  const int r = lane & 31, half = lane >> 5;
  const int gr = row0 + r;
#pragma unroll 8
  for (int it = 0; it < 32; it++) {
    const int c = 2 * it + half, gc = col0 + c;
    if (gc < ncols && gr < n) Z[(long)gc * ldz + gr] = stage[r * kStageLd + c];
  }
  lds_order();
}

template <int NB, int BS, int NW, bool REV>
__global__ __launch_bounds__(NW * 64) void bcback_kernel(int n, int ncols, float *__restrict__ Z, long ldz,
                                                        const float *__restrict__ P, int nsweeps) {
  static_assert(NB == 32 || NB == 64, "window chunks are 32 rows");
  constexpr int W = BS + NB;
  constexpr int SLOT = NB * kStageLd;
  // slot pair of wave k: the rows it hands to wave k+1 (double-buffered), and
  // its staging area for coalesced global loads and stores
  __shared__ float xch[NW * 2 * SLOT];
  const int lane = threadIdx.x & 63;
  const int k = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int col0 = blockIdx.x * 64;
  const int nblk = (nsweeps + BS - 1) / BS;
  constexpr long m = BS / NB;
  const long A = (n - 2) / NB;
  float z[W];

  const int ngroups = (nblk + NW - 1) / NW;
  for (int gi = 0; gi < ngroups; gi++) {
    int bk, nact, s0min;
    if (REV) {
      const int gtop = nblk - 1 - gi * NW;
      nact = min(NW, gtop + 1);
      bk = gtop - k;
      s0min = (gtop - nact + 1) * BS;
    } else {
      const int gbot = gi * NW;
      nact = min(NW, nblk - gbot);
      bk = gbot + k;
      s0min = gbot * BS;
    }
    const bool active = k < nact;
    const int S0 = bk * BS;
    const int J = (n - 1 - s0min + NB - 1) / NB + 1;
    const int T = J + nact - 1;
    for (int tau = 0; tau < T; tau++) {
      const int e = tau - k;
      if (active && e >= 0 && e < J) {
        const int jj = REV ? e : (J - 1 - e);
        const int wb = S0 + 1 + jj * NB;
        const float *xin = xch + ((k - 1) * 2 + (e & 1)) * SLOT;
        float *xout = xch + (k * 2 + (e & 1)) * SLOT;
        // ---- bring the window up to date (global rows staged through xout,
        // which is only written with this step's hand-off rows at its end)
        auto from_stage = [&](auto cc) __attribute__((always_inline)) {
          constexpr int C = decltype(cc)::value;
#pragma unroll
          for (int i = 0; i < 32; i++) z[C + i] = xout[i * kStageLd + lane];
          lds_order();
        };
        if (REV) {
          if (e == 0) {
            const int ng = (k == 0) ? W : BS;
            for_chunks<0, W>([&](auto cc) __attribute__((always_inline)) {
              constexpr int C = decltype(cc)::value;
              if (C < ng) {
                stage_load32(xout, Z, ldz, col0, ncols, wb + C, n, lane);
                from_stage(cc);
              } else {
#pragma unroll
                for (int i = 0; i < 32; i++) z[C + i] = xin[(C - BS + i) * kStageLd + lane];
              }
            });
          } else {
#pragma unroll
            for (int i = 0; i < BS; i++) z[i] = z[i + NB];
            if (k == 0) {
              for_chunks<BS, W>([&](auto cc) __attribute__((always_inline)) {
                stage_load32(xout, Z, ldz, col0, ncols, wb + decltype(cc)::value, n, lane);
                from_stage(cc);
              });
            } else {
#pragma unroll
              for (int i = 0; i < NB; i++) z[BS + i] = xin[i * kStageLd + lane];
            }
          }
        } else {
          if (e == 0) {
            for_chunks<0, W>([&](auto cc) __attribute__((always_inline)) {
              constexpr int C = decltype(cc)::value;
              if (k > 0 && C < NB) {
#pragma unroll
                for (int i = 0; i < 32; i++) z[C + i] = xin[(C + i) * kStageLd + lane];
              } else {
                stage_load32(xout, Z, ldz, col0, ncols, wb + C, n, lane);
                from_stage(cc);
              }
            });
          } else {
#pragma unroll
            for (int i = BS - 1; i >= 0; i--) z[NB + i] = z[i];
            if (k == 0) {
              for_chunks<0, NB>([&](auto cc) __attribute__((always_inline)) {
                stage_load32(xout, Z, ldz, col0, ncols, wb + decltype(cc)::value, n, lane);
                from_stage(cc);
              });
            } else {
#pragma unroll
              for (int i = 0; i < NB; i++) z[i] = xin[i * kStageLd + lane];
            }
          }
        }
        // ---- apply the BS reflectors of this block at step jj
        if (jj < pack_steps(A, m, bk))
          apply_packed_block<NB, BS, W, REV>(z, P + (pack_prefix(A, m, bk) + jj) * (long)(BS * NB), lane);
        // ---- hand on the finished rows (the last wave writes them to Z)
        const bool last = (k == nact - 1);
        if (REV) {
          for_chunks<0, NB>([&](auto cc) __attribute__((always_inline)) {
            constexpr int C = decltype(cc)::value;
#pragma unroll
            for (int i = 0; i < 32; i++) xout[(C + i) * kStageLd + lane] = z[C + i];
            if (last) stage_store32(xout + C * kStageLd, Z, ldz, col0, ncols, wb + C, n, lane);
          });
        } else {
          if (e == J - 1) {
            // rows above the next wave's window are final
            for_chunks<0, BS>([&](auto cc) __attribute__((always_inline)) {
              constexpr int C = decltype(cc)::value;
#pragma unroll
              for (int i = 0; i < 32; i++) xout[i * kStageLd + lane] = z[C + i];
              stage_store32(xout, Z, ldz, col0, ncols, wb + C, n, lane);
            });
          }
          for_chunks<0, NB>([&](auto cc) __attribute__((always_inline)) {
            constexpr int C = decltype(cc)::value;
#pragma unroll
            for (int i = 0; i < 32; i++) xout[(C + i) * kStageLd + lane] = z[BS + C + i];
            if (last) stage_store32(xout + C * kStageLd, Z, ldz, col0, ncols, wb + BS + C, n, lane);
          });
        }
      }
      __syncthreads();
    }
  }
}

size_t bc_back_pack_bytes(long n, long b) {
  if (b == 64) return align256(pack_floats<64, 64>(n) * sizeof(float));
  return align256(std::max(pack_floats<32, 32>(n), pack_floats<32, 64>(n)) * sizeof(float));
}

template <int NB, int BS, int NW>
static int launch_bcback(evd_ctx *ctx, long n, long ncols, float *Q, long ldq, const float *U, long ldu, bool rev) {
  const int nsweeps = (int)std::max<long>(0, n - 2);
  if (nsweeps == 0 || ncols <= 0) return 0;
  ArenaFrame frame(ctx);
  float *P = arena_alloc_n<float>(ctx, pack_floats<NB, BS>(n));
  if (!P) return EVD_E_NOMEM;
  const unsigned nblk = (unsigned)((nsweeps + BS - 1) / BS);
  bc_pack_kernel<NB, BS><<<dim3(64, nblk), 256, 0, ctx->stream>>>((int)n, U, ldu, P, nsweeps);
  EVD_LAUNCHED(ctx);
  if (ctx->timing) hipEventRecord(ctx->kev[0], ctx->stream);
  dim3 grid((unsigned)((ncols + 63) / 64));
  if (rev)
    bcback_kernel<NB, BS, NW, true><<<grid, NW * 64, 0, ctx->stream>>>((int)n, (int)ncols, Q, ldq, P, nsweeps);
  else
    bcback_kernel<NB, BS, NW, false><<<grid, NW * 64, 0, ctx->stream>>>((int)n, (int)ncols, Q, ldq, P, nsweeps);
  EVD_LAUNCHED(ctx);
  if (ctx->timing) hipEventRecord(ctx->kev[1], ctx->stream);
  return 0;
}

size_t bc_back_wy_bytes(long n);
bool bc_back_wy_fits(long n, long ldq);
int bc_back_wy(evd_ctx *ctx, long n, long ncols, float *Q, long ldq, const float *U, long ldu, bool split16);
size_t bc_back_wy64_bytes(long n);
bool bc_back_wy64_fits(long n, long ldq);
int bc_back_wy64_batched(evd_ctx *ctx, int count, long n, long ncols, float *Q, long ldq, long q_stride,
                         const float *U, long ldu, long u_stride);

// Q <- H_1 .. H_K Q with b = 32 runs as WY blocks on the matrix cores
// (bcback_mfma.hip): split FP16 in the FP16 mixed-precision mode, FP32 MFMA
// otherwise; the other order and b = 64 run on the FP32 VALU kernel below.
static bool use_wy(const evd_ctx *ctx, long n, long b, long ldq, bool rev) {
  if (!rev || ctx->tune.bc_valu) return false;
  return (b == 32 && bc_back_wy_fits(n, ldq)) || (b == 64 && bc_back_wy64_fits(n, ldq));
}

size_t bc_back_bytes(const evd_ctx *ctx, long n, long b, long ldq, bool rev) {
  if (b != 32 && b != 64) return 0;  // the small-order kernel needs no workspace
  if (!use_wy(ctx, n, b, ldq, rev)) return bc_back_pack_bytes(n, b);
  return b == 64 ? bc_back_wy64_bytes(n) : bc_back_wy_bytes(n);
}

// Small orders with any bandwidth b <= 64 (n <= kSmallN, the small buckets
// of a mixed-size batch, where b = min(32, n / 4) can be 16 or 24): one
// workgroup per problem (blockIdx.y) holds all of Q (n x ncols) in LDS, one
// thread per column, and applies the reflectors one by one in the order of
// the product (generation order for order 0, reversed for order 1).
constexpr long kSmallN = 128;

// reflectors of one small problem: (s, p, len) packed as s | p << 8 | len << 16
constexpr int kSmallRefl = 8192;  // n <= 128, b >= 2: at most 126 sweeps x 64 steps

__global__ __launch_bounds__(128) void bcback_small_kernel(int n, int b, int ncols, float *__restrict__ Q_all, long ldq,
                                                           long q_stride, const float *__restrict__ U_all, long ldu,
                                                           long u_stride, int rev) {
  __shared__ float q[kSmallN * (kSmallN + 1)];
  __shared__ float u[64];
  __shared__ int list[kSmallRefl];
  __shared__ int s_total;
  float *Q = Q_all + (long)blockIdx.y * q_stride;
  const float *U = U_all + (long)blockIdx.y * u_stride;
  const int c = threadIdx.x;
  constexpr int LDQ = kSmallN + 1;
  for (int r = 0; r < n; r++)
    if (c < ncols) q[r + c * LDQ] = Q[r + (long)c * ldq];
  if (c == 0) {
    int total = 0;
    for (int sw = 0; sw < n - 2; sw++) {
      int p = sw + 1, len = min(b, n - p);
      while (total < kSmallRefl) {
        list[total++] = sw | (p << 8) | (len << 16);
        const int len2 = max(0, min(b, n - p - len));
        if (len2 <= 1) break;
        p += len;
        len = len2;
      }
    }
    s_total = total;
  }
  __syncthreads();
  const int total = s_total;
  for (int k = 0; k < total; k++) {
    const int e = list[rev ? total - 1 - k : k];
    const int sw = e & 0xFF, p = (e >> 8) & 0xFF, len = e >> 16;
    if (c < len) u[c] = U[(p + c) + (long)sw * ldu];
    __syncthreads();
    if (c < ncols) {
      float d = 0.f;
      for (int i = 0; i < len; i++) d += u[i] * q[(p + i) + c * LDQ];
      for (int i = 0; i < len; i++) q[(p + i) + c * LDQ] -= u[i] * d;
    }
    __syncthreads();
  }
  for (int r = 0; r < n; r++)
    if (c < ncols) Q[r + (long)c * ldq] = q[r + c * LDQ];
}

int bc_back_small_batched(evd_ctx *ctx, int count, long n, long b, long ncols, float *Q, long ldq, long q_stride,
                          const float *U, long ldu, long u_stride, bool rev) {
  if (n <= 2 || ncols <= 0 || count <= 0) return 0;
  if (n > kSmallN || ncols > kSmallN || b < 2 || b > 64)
    return set_error(ctx, EVD_E_UNSUPPORTED, "BC back transformation with b = %ld needs n <= %ld and b >= 2", b,
                     kSmallN);
  bcback_small_kernel<<<dim3(1, (unsigned)count), 128, 0, ctx->stream>>>((int)n, (int)b, (int)ncols, Q, ldq, q_stride,
                                                                         U, ldu, u_stride, rev ? 1 : 0);
  EVD_LAUNCHED(ctx);
  return 0;
}

int bc_back(evd_ctx *ctx, long n, long b, long ncols, float *Q, long ldq, const float *U, long ldu, bool rev) {
  if (b != 32 && b != 64) return bc_back_small_batched(ctx, 1, n, b, ncols, Q, ldq, 0, U, ldu, 0, rev);
  if (use_wy(ctx, n, b, ldq, rev)) {
    if (b == 64) return bc_back_wy64_batched(ctx, 1, n, ncols, Q, ldq, 0, U, ldu, 0);
    return bc_back_wy(ctx, n, ncols, Q, ldq, U, ldu, ctx->precision == EVD_PREC_FP16);
  }
  const int bs = ctx->tune.bc_sweeps;
  const int nw = ctx->tune.bc_waves;
  if (b == 32) {
    if (bs == 32) {
      if (nw == 4) return launch_bcback<32, 32, 4>(ctx, n, ncols, Q, ldq, U, ldu, rev);
      return launch_bcback<32, 32, 8>(ctx, n, ncols, Q, ldq, U, ldu, rev);
    }
    if (nw == 4) return launch_bcback<32, 64, 4>(ctx, n, ncols, Q, ldq, U, ldu, rev);
    return launch_bcback<32, 64, 8>(ctx, n, ncols, Q, ldq, U, ldu, rev);
  }
  if (b == 64) return launch_bcback<64, 64, 4>(ctx, n, ncols, Q, ldq, U, ldu, rev);  // 8 waves exceed LDS
  return set_error(ctx, EVD_E_UNSUPPORTED, "BC back transformation needs b = 32 or 64 (got %ld)", b);
}

}  // namespace evd

using namespace evd;

extern "C" int evd_bc_back_f32(evd_ctx *ctx, long n, long b, long ncols, float *Q, long ldq, const float *U, long ldu,
                               int order) {
  if (!ctx) return -1;
  EVD_ARG(ctx, n >= 0, 2);
  EVD_ARG(ctx, b == 32 || b == 64 || (b >= 2 && b <= 64 && n <= kSmallN && ncols <= kSmallN), 3);
  EVD_ARG(ctx, ncols >= 0, 4);
  EVD_ARG(ctx, ldq >= std::max<long>(1, n), 6);
  EVD_ARG(ctx, ldu >= n, 8);
  EVD_ARG(ctx, order == 0 || order == 1, 9);
  if (n <= 2 || ncols == 0) return 0;
  if (ctx->arena_top == 0) EVD_TRY(arena_ensure(ctx, bc_back_bytes(ctx, n, b, ldq, order == 1) + 4096));
  return bc_back(ctx, n, b, ncols, Q, ldq, U, ldu, order == 1);
}
