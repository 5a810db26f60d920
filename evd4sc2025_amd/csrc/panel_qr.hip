// Tall-skinny Householder panel QR (the role of panelQR / hou_tsqr_panel /
// my_hou_kernel in the reference: api/panelQR.cu:20-246,
// api/inc/TallShinnyQR.h:9-134, util/inc/kernelQR.h:215-414).
//
// Design (MI355X): one cooperative launch of G workgroups; workgroup g keeps
// rows [g*256*RPT, (g+1)*256*RPT) of the m x nb panel in registers as FP64 for
// the whole factorisation (the panel is read once and written once).  Column
// j needs sigma = ||x(j+1:)||^2 and p_c = x(j+1:)^T a_c(j+1:) for the trailing
// columns; both are sums over rows below the pivot and are reduced in ONE
// grid-wide reduction per column, together with the Gram entries v_i^T v_{j-1}
// of the previous reflector.  After the barrier every workgroup forms the
// same LAPACK dlarfg reflector (beta = -sign(alpha) ||x||, tau = (beta -
// alpha)/beta, v = x / (alpha - beta)) and applies it to its rows.  The Gram
// matrix V^T V gives T by the dlarft forward/columnwise recurrence; W = V T is
// formed row-locally.  Outputs: Y = V (unit lower, zeros above), W = V T, R,
// and optionally [R; 0] into the panel storage.
#include "common.hpp"

namespace evd {

// Reduce NV (power of two, <= 64) per-lane values across the wave.  On exit
// lane l holds the total of value index l / (64 / NV) in v[0].
template <int NV>
__device__ inline double wave_reduce_scatter(double (&v)[NV], int lane) {
#pragma unroll
  for (int L = 0; (NV >> L) > 1; L++) {
    constexpr int dummy = 0;
    (void)dummy;
    const int o = 32 >> L;
    const int half = (NV >> L) >> 1;
    const bool up = (lane & o) != 0;
#pragma unroll
    for (int i = 0; i < NV / 2; i++) {
      if (i < half) {
        double send = up ? v[i] : v[i + half];
        double keep = up ? v[i + half] : v[i];
        v[i] = keep + __shfl_xor(send, o, 64);
      }
    }
  }
  // lanes that share an index: the low log2(64/NV) lane bits
#pragma unroll
  for (int o = (64 / NV) >> 1; o > 0; o >>= 1) v[0] += __shfl_xor(v[0], o, 64);
  return v[0];
}

template <int NB>
__device__ inline double sel(const double (&row)[NB], int j) {
  double x = 0.0;
#pragma unroll
  for (int c = 0; c < NB; c++) x = (c == j) ? row[c] : x;
  return x;
}

template <int NB, int RPT>
__global__ __launch_bounds__(256) void panel_qr_kernel(int m, int nc, const float *__restrict__ Ain, long lda,
                                                       float *__restrict__ Y, long ldy, float *__restrict__ W,
                                                       long ldw, float *__restrict__ R, long ldr,
                                                       float *__restrict__ Aout, long ldo, double *__restrict__ part,
                                                       double *__restrict__ pivrow, double *__restrict__ tau_out,
                                                       unsigned *__restrict__ tags, unsigned epoch,
                                                       unsigned *__restrict__ err) {
  const int G = gridDim.x;
  const int g = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int k = min(m, nc);
  __shared__ double red[4][NB];
  __shared__ double tot[NB];
  __shared__ double gpart[32 * NB];  // partial records of up to 32 groups per pass
  __shared__ int s_ok;
  __shared__ double piv[NB];
  __shared__ double wv[NB];
  __shared__ double sc[4];
  __shared__ double gram[NB][NB + 1];
  __shared__ double taus[NB];
  __shared__ double T[NB][NB + 1];

  double a[RPT][NB];
  long rowv[RPT];
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    long r = (long)g * 256 * RPT + q * 256 + tid;
    rowv[q] = r;
#pragma unroll
    for (int c = 0; c < NB; c++) a[q][c] = (r < m && c < nc) ? (double)Ain[r + (long)c * lda] : 0.0;
  }
  for (int i = tid; i < NB * (NB + 1); i += 256) {
    (&gram[0][0])[i] = 0.0;
    (&T[0][0])[i] = 0.0;
  }
  if (G > 1) {
    // sentinels of steps 0 and 1 in this group's records, then a one-time
    // start barrier so nobody polls a slot before its owner marked it
    const long rec = 2 * NB;
    for (int st = 0; st < 2; st++) {
      const unsigned long long snt = 0x7FF8000000000000ull | ((unsigned long long)(epoch & 0xFFFFFu) << 24) |
                                     (unsigned long long)st;
      double *r0 = part + ((long)st * G + g) * rec;
      for (int c = tid; c < (int)rec; c += 256) st_dev(&r0[c], __longlong_as_double((long long)snt));
    }
    wait_mem();
    __syncthreads();
    if (tid == 0) {
      s_ok = 1;
      st_dev(&tags[g], epoch);
    }
    __syncthreads();
    for (int gg = tid; gg < G; gg += 256) {
      const unsigned long long t0 = clock100();
      while (ld_dev(&tags[gg]) != epoch) {
        __builtin_amdgcn_s_sleep(1);
        if (clock100() - t0 > kWatchdogTicks) {
          atomicOr(err, (unsigned)EVD_E_TIMEOUT);
          s_ok = 0;
          break;
        }
      }
    }
    __syncthreads();
    if (!s_ok) return;
  }

  for (int j = 0; j <= k; j++) {
    // ---- partials: gram(i, j-1) for i < j-1; sigma_j; p_c (c > j), in chunks
    // of 16 columns to bound register pressure
    double vprevq[RPT], xq[RPT];
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      const long r = rowv[q];
      vprevq[q] = (r < m && j >= 1) ? ((r > j - 1) ? sel<NB>(a[q], j - 1) : (r == j - 1 ? 1.0 : 0.0)) : 0.0;
      xq[q] = (r < m && j < k && r > j) ? sel<NB>(a[q], j) : 0.0;
    }
#pragma unroll
    for (int cb = 0; cb < NB; cb += 16) {
      double pv[16];
#pragma unroll
      for (int c = 0; c < 16; c++) pv[c] = 0.0;
#pragma unroll
      for (int q = 0; q < RPT; q++) {
        const long r = rowv[q];
        if (r < m) {
          const double vprev = vprevq[q];
          const double x = xq[q];
#pragma unroll
          for (int cc = 0; cc < 16; cc++) {
            const int c = cb + cc;
            double contrib = 0.0;
            if (c < j - 1)
              contrib = a[q][c] * vprev;  // r >= j-1 > c wherever vprev != 0
            else if (c == j)
              contrib = x * x;
            else if (c > j && c < nc)
              contrib = x * a[q][c];
            pv[cc] += contrib;
          }
        }
      }
      double mine = wave_reduce_scatter<16>(pv, lane);
      if ((lane & 3) == 0) red[wave][cb + (lane >> 2)] = mine;
    }
    __syncthreads();
    // ---- exchange.  One workgroup: straight from LDS.  Several: every group
    // stores its record (partials, and the pivot row from group 0) into ring
    // slot j%4 with device-coherent stores; readers poll the values themselves
    // until none equals the slot's sentinel (a NaN whose payload encodes the
    // launch epoch and the step).  The sentinel of step j+2 is written at step
    // j, into the slot that held step j-2, which every reader has consumed by
    // then; the vmcnt wait of this step's polling orders it before the values
    // of step j+1.
    if (G == 1) {
      if (tid < NB) tot[tid] = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
      if (tid == j && j < k) {
#pragma unroll
        for (int c = 0; c < NB; c++) piv[c] = a[0][c];
      }
      __syncthreads();
    } else {
      const int slot = j & 3;
      const long rec = 2 * NB;
      double *myrec = part + ((long)slot * G + g) * rec;
      if (tid < NB) st_dev(&myrec[tid], red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid]);
      if (g == 0 && tid == j && j < k) {
#pragma unroll
        for (int c = 0; c < NB; c++) st_dev(&myrec[NB + c], a[0][c]);
      }
      // sentinel for step j+2 into its slot (this group's record only)
      {
        const unsigned long long snt = 0x7FF8000000000000ull | ((unsigned long long)(epoch & 0xFFFFFu) << 24) |
                                       (unsigned long long)((j + 2) & 0xFFFFFF);
        double *nrec = part + ((long)((j + 2) & 3) * G + g) * rec;
        for (int c = tid; c < (int)rec; c += 256) st_dev(&nrec[c], __longlong_as_double((long long)snt));
      }
      const unsigned long long want = 0x7FF8000000000000ull | ((unsigned long long)(epoch & 0xFFFFFu) << 24) |
                                      (unsigned long long)(j & 0xFFFFFF);
      if (tid == 0) s_ok = 1;
      if (tid < NB) tot[tid] = 0.0;
      for (int g0 = 0; g0 < G; g0 += 32) {
        const int gn = min(32, G - g0);
        __syncthreads();
        for (int idx = tid; idx < gn * NB; idx += 256) {
          const double *src = part + ((long)slot * G + g0 + idx / NB) * rec + (idx % NB);
          double v = ld_dev(src);
          const unsigned long long t0 = clock100();
          while ((unsigned long long)__double_as_longlong(v) == want) {
            __builtin_amdgcn_s_sleep(1);
            v = ld_dev(src);
            if (clock100() - t0 > kWatchdogTicks) {
              atomicOr(err, (unsigned)EVD_E_TIMEOUT);
              s_ok = 0;
              break;
            }
          }
          gpart[idx] = v;
        }
        __syncthreads();
        if (tid < NB) {
          double sacc = tot[tid];
          for (int gg = 0; gg < gn; gg++) sacc += gpart[gg * NB + tid];
          tot[tid] = sacc;
        }
      }
      if (j < k && tid < NB) {
        const double *src = part + ((long)slot * G + 0) * rec + NB + tid;
        double v = ld_dev(src);
        const unsigned long long t0 = clock100();
        while ((unsigned long long)__double_as_longlong(v) == want) {
          __builtin_amdgcn_s_sleep(1);
          v = ld_dev(src);
          if (clock100() - t0 > kWatchdogTicks) {
            atomicOr(err, (unsigned)EVD_E_TIMEOUT);
            s_ok = 0;
            break;
          }
        }
        piv[tid] = v;
      }
      __syncthreads();
      if (!s_ok) return;
    }
    __syncthreads();
    if (tid < NB && tid < j - 1) gram[tid][j - 1] = tot[tid];
    if (j == k) break;
    if (tid == 0) {
      const double alpha = piv[j];
      const double sigma = tot[j];
      double beta, tau, scale;
      if (sigma == 0.0) {
        beta = alpha;
        tau = 0.0;
        scale = 0.0;
      } else {
        double nrm = sqrt(alpha * alpha + sigma);
        beta = alpha >= 0.0 ? -nrm : nrm;
        tau = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
      }
      sc[0] = beta;
      sc[1] = tau;
      sc[2] = scale;
      taus[j] = tau;
    }
    __syncthreads();
    if (tid < NB) {
      const int c = tid;
      wv[c] = (c > j && c < nc) ? sc[1] * (piv[c] + tot[c] * sc[2]) : 0.0;
    }
    __syncthreads();
    const double beta = sc[0], scale = sc[2];
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      const long r = rowv[q];
      if (r >= m) continue;
      if (r > j) {
        const double v = sel<NB>(a[q], j) * scale;
#pragma unroll
        for (int c = 0; c < NB; c++) {
          if (c == j)
            a[q][c] = v;
          else if (c > j)
            a[q][c] -= v * wv[c];
        }
      } else if (r == j) {
#pragma unroll
        for (int c = 0; c < NB; c++) {
          if (c == j)
            a[q][c] = beta;
          else if (c > j)
            a[q][c] -= wv[c];
        }
      }
    }
  }

  // ---- T by the dlarft forward/columnwise recurrence (all workgroups alike)
  for (int j = 0; j < k; j++) {
    if (tid < j) {
      double s = 0.0;
      for (int l = tid; l < j; l++) s += T[tid][l] * gram[l][j];
      T[tid][j] = -taus[j] * s;
    }
    if (tid == j) T[j][j] = taus[j];
    __syncthreads();
  }
  if (g == 0 && tau_out && tid < k) tau_out[tid] = taus[tid];

  // ---- outputs: R / [R;0] first, then a[] becomes V in place, then W = V T
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    const long r = rowv[q];
    if (r >= m) continue;
#pragma unroll
    for (int c = 0; c < NB; c++) {
      if (c >= nc) continue;
      const double rv = (r <= c) ? a[q][c] : 0.0;
      if (R && r < nc) R[r + (long)c * ldr] = (float)rv;
      if (Aout) Aout[r + (long)c * ldo] = (float)rv;
    }
#pragma unroll
    for (int c = 0; c < NB; c++) a[q][c] = (c < k) ? ((r > c) ? a[q][c] : (r == c ? 1.0 : 0.0)) : 0.0;
#pragma unroll
    for (int c = 0; c < NB; c++) {
      if (c >= nc) continue;
      double w = 0.0;
#pragma unroll
      for (int l = 0; l <= c; l++) w += a[q][l] * T[l][c];
      if (W) W[r + (long)c * ldw] = (float)(c < k ? w : 0.0);
      if (Y) Y[r + (long)c * ldy] = (float)a[q][c];
    }
  }
}

template <int NB, int RPT>
static int launch_panel(evd_ctx *ctx, long m, long nc, const float *A, long lda, float *Y, long ldy, float *W,
                        long ldw, float *R, long ldr, float *Aout, long ldo, double *tau_out) {
  const long rows_per = 256L * RPT;
  int G = (int)((m + rows_per - 1) / rows_per);
  int maxres = 0;
  EVD_TRY(max_coresident_blocks(ctx, (const void *)panel_qr_kernel<NB, RPT>, 256, 0, &maxres));
  if (G > maxres) return set_error(ctx, EVD_E_UNSUPPORTED, "panel of %ld rows needs %d co-resident groups (max %d)", m, G, maxres);
  ArenaFrame frame(ctx);
  double *part = arena_alloc_n<double>(ctx, (size_t)4 * G * 2 * NB);
  double *piv = arena_alloc_n<double>(ctx, (size_t)2 * NB);
  if (!part || !piv) return EVD_E_NOMEM;
  if (G > kTagWords) return set_error(ctx, EVD_E_UNSUPPORTED, "panel needs %d exchange tags", G);
  unsigned *tags = ctx->tags;
  // tags carry a launch epoch, so stale tags of earlier launches never match
  unsigned epoch = (++ctx->epoch) & 0x1FFFFFFu;
  if (epoch == 0) epoch = (++ctx->epoch) & 0x1FFFFFFu;
  unsigned *err = ctx->ctrl + kCtrlError;
  int mi = (int)m, ni = (int)nc;
  void *args[] = {&mi, &ni, (void *)&A, &lda, &Y, &ldy, &W, &ldw, &R, &ldr, &Aout, &ldo, &part, &piv, &tau_out, &tags, &epoch, &err};
  EVD_HIP(ctx, hipLaunchCooperativeKernel((const void *)panel_qr_kernel<NB, RPT>, dim3(G), dim3(256), args, 0,
                                          ctx->stream));
  return 0;
}

// Panel factorisation used by the SBR driver and by evd_panel_qr_f32.
int panel_qr(evd_ctx *ctx, long m, long nc, const float *A, long lda, float *Y, long ldy, float *W, long ldw,
             float *R, long ldr, float *Aout, long ldo, double *tau_out) {
  if (m <= 0 || nc <= 0) return 0;
  if (nc <= 32) return launch_panel<32, 2>(ctx, m, nc, A, lda, Y, ldy, W, ldw, R, ldr, Aout, ldo, tau_out);
  if (nc <= 64) return launch_panel<64, 1>(ctx, m, nc, A, lda, Y, ldy, W, ldw, R, ldr, Aout, ldo, tau_out);
  return set_error(ctx, EVD_E_UNSUPPORTED, "panel width %ld > 64", nc);
}

}  // namespace evd

using namespace evd;

extern "C" int evd_panel_qr_f32(evd_ctx *ctx, long m, long nb, float *A, long lda, float *W, long ldw, float *R,
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
  // reference semantics: A <- Y (unit lower), W, R
  return panel_qr(ctx, m, nb, A, lda, A, lda, W, ldw, R, ldr, nullptr, 0, nullptr);
}
