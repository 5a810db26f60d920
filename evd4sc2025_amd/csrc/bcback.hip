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
// each SIMD runs one wave of pure FMA work.
#include "common.hpp"

namespace evd {

// Fast path: every reflector of the block exists and lies inside the matrix,
// so the entries are used straight from SGPRs as FMA operands.
template <int NB, int BS, int W, bool REV, int TT>
__device__ __attribute__((always_inline)) inline void apply_block_full(float (&z)[W], const float *__restrict__ U,
                                                                       long ldu, int S0, int wb) {
  if constexpr (TT < BS) {
    constexpr int t = REV ? (BS - 1 - TT) : TT;
    const float *up = U + (wb + t) + (long)(S0 + t) * ldu;
    float d0 = 0.f, d1 = 0.f, d2 = 0.f, d3 = 0.f;
#pragma unroll
    for (int i = 0; i < NB; i += 4) {
      d0 += up[i] * z[t + i];
      d1 += up[i + 1] * z[t + i + 1];
      d2 += up[i + 2] * z[t + i + 2];
      d3 += up[i + 3] * z[t + i + 3];
    }
    const float dot = (d0 + d1) + (d2 + d3);
#pragma unroll
    for (int i = 0; i < NB; i++) z[t + i] -= dot * up[i];
    apply_block_full<NB, BS, W, REV, TT + 1>(z, U, ldu, S0, wb);
  }
}

// General path (blocks at the matrix edge): a reflector is skipped when its
// sweep or its start row does not exist.  Rows of U at and below n must be
// zero (the pipeline's U is zero-filled; the drop-in entry pads a copy), so no
// per-entry masking is needed.
template <int NB, int BS, int W, bool REV, int TT>
__device__ __attribute__((always_inline)) inline void apply_block(float (&z)[W], const float *__restrict__ U,
                                                                  long ldu, int S0, int wb, int n, int nsweeps) {
  if constexpr (TT < BS) {
    constexpr int t = REV ? (BS - 1 - TT) : TT;
    if (S0 + t < nsweeps && wb + t < n - 1) {
      const float *up = U + (wb + t) + (long)(S0 + t) * ldu;
      float d0 = 0.f, d1 = 0.f, d2 = 0.f, d3 = 0.f;
#pragma unroll
      for (int i = 0; i < NB; i += 4) {
        d0 += up[i] * z[t + i];
        d1 += up[i + 1] * z[t + i + 1];
        d2 += up[i + 2] * z[t + i + 2];
        d3 += up[i + 3] * z[t + i + 3];
      }
      const float dot = (d0 + d1) + (d2 + d3);
#pragma unroll
      for (int i = 0; i < NB; i++) z[t + i] -= dot * up[i];
    }
    apply_block<NB, BS, W, REV, TT + 1>(z, U, ldu, S0, wb, n, nsweeps);
  }
}

template <int NB, int BS, int NW, bool REV>
__global__ __launch_bounds__(NW * 64) void bcback_kernel(int n, int ncols, float *__restrict__ Z, long ldz,
                                                        const float *__restrict__ U, long ldu, int nsweeps) {
  constexpr int W = BS + NB;
  __shared__ float xch[(NW > 1 ? NW - 1 : 1) * 2 * NB * 64];
  const int lane = threadIdx.x & 63;
  const int k = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int col = blockIdx.x * 64 + lane;
  const bool colok = col < ncols;
  float *Zc = Z + (long)(colok ? col : 0) * ldz;
  const int nblk = (nsweeps + BS - 1) / BS;
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
        const float *xin = xch + ((k - 1) * 2 + (e & 1)) * NB * 64;
        float *xout = xch + (k * 2 + (e & 1)) * NB * 64;
        // ---- bring the window up to date
        if (REV) {
          if (e == 0) {
            const int ng = (k == 0) ? W : BS;
#pragma unroll
            for (int i = 0; i < W; i++) {
              if (i < ng) {
                const int row = wb + i;
                z[i] = (colok && row < n) ? Zc[row] : 0.f;
              } else {
                z[i] = xin[(i - BS) * 64 + lane];
              }
            }
          } else {
#pragma unroll
            for (int i = 0; i < BS; i++) z[i] = z[i + NB];
#pragma unroll
            for (int i = 0; i < NB; i++) {
              const int row = wb + BS + i;
              z[BS + i] = (k == 0) ? ((colok && row < n) ? Zc[row] : 0.f) : xin[i * 64 + lane];
            }
          }
        } else {
          if (e == 0) {
#pragma unroll
            for (int i = 0; i < W; i++) {
              if (k > 0 && i < NB) {
                z[i] = xin[i * 64 + lane];
              } else {
                const int row = wb + i;
                z[i] = (colok && row < n) ? Zc[row] : 0.f;
              }
            }
          } else {
#pragma unroll
            for (int i = BS - 1; i >= 0; i--) z[NB + i] = z[i];
#pragma unroll
            for (int i = 0; i < NB; i++) {
              const int row = wb + i;
              z[i] = (k == 0) ? ((colok && row < n) ? Zc[row] : 0.f) : xin[i * 64 + lane];
            }
          }
        }
        // ---- apply the BS reflectors of this block at step jj
        if (S0 + BS <= nsweeps && wb + BS + NB <= n)
          apply_block_full<NB, BS, W, REV, 0>(z, U, ldu, S0, wb);
        else
          apply_block<NB, BS, W, REV, 0>(z, U, ldu, S0, wb, n, nsweeps);
        // ---- hand on the finished rows
        const bool last = (k == nact - 1);
        if (REV) {
#pragma unroll
          for (int i = 0; i < NB; i++) {
            const int row = wb + i;
            if (last) {
              if (colok && row < n) Zc[row] = z[i];
            } else {
              xout[i * 64 + lane] = z[i];
            }
          }
        } else {
#pragma unroll
          for (int i = 0; i < NB; i++) {
            const int row = wb + BS + i;
            if (last) {
              if (colok && row < n) Zc[row] = z[BS + i];
            } else {
              xout[i * 64 + lane] = z[BS + i];
            }
          }
          if (e == J - 1) {
            // rows above the next wave's window are final
#pragma unroll
            for (int i = 0; i < BS; i++) {
              const int row = wb + i;
              if (colok && row < n) Zc[row] = z[i];
            }
          }
        }
      }
      __syncthreads();
    }
  }
}

template <int NB, int BS, int NW>
static int launch_bcback(evd_ctx *ctx, long n, long ncols, float *Q, long ldq, const float *U, long ldu, bool rev) {
  const int nsweeps = (int)std::max<long>(0, n - 2);
  if (nsweeps == 0 || ncols <= 0) return 0;
  dim3 grid((unsigned)((ncols + 63) / 64));
  if (rev)
    bcback_kernel<NB, BS, NW, true><<<grid, NW * 64, 0, ctx->stream>>>((int)n, (int)ncols, Q, ldq, U, ldu, nsweeps);
  else
    bcback_kernel<NB, BS, NW, false><<<grid, NW * 64, 0, ctx->stream>>>((int)n, (int)ncols, Q, ldq, U, ldu, nsweeps);
  EVD_LAUNCHED(ctx);
  return 0;
}

int bc_back(evd_ctx *ctx, long n, long b, long ncols, float *Q, long ldq, const float *U, long ldu, bool rev) {
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
  EVD_ARG(ctx, b == 32 || b == 64, 3);
  EVD_ARG(ctx, ncols >= 0, 4);
  EVD_ARG(ctx, ldq >= std::max<long>(1, n), 6);
  EVD_ARG(ctx, ldu >= n, 8);
  EVD_ARG(ctx, order == 0 || order == 1, 9);
  if (n <= 2 || ncols == 0) return 0;
  // the kernel reads up to b rows past the last reflector: run on a copy of
  // the caller's U whose rows n .. n+b-1 are zero
  const long ldp = n + b;
  if (ctx->arena_top == 0) EVD_TRY(arena_ensure(ctx, (size_t)ldp * n * sizeof(float) + 4096));
  ArenaFrame frame(ctx);
  float *Up = arena_alloc_n<float>(ctx, (size_t)ldp * n);
  if (!Up) return EVD_E_NOMEM;
  EVD_TRY(fill_f32(ctx, Up, ldp, n, ldp, 0.f));
  EVD_TRY(copy_f32(ctx, U, ldu, Up, ldp, n, n));
  return bc_back(ctx, n, b, ncols, Q, ldq, Up, ldp, order == 1);
}
