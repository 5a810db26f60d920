// Tall-skinny Householder panel QR (the role of panelQR / hou_tsqr_panel /
// my_hou_kernel in the reference: api/panelQR.cu:20-246,
// api/inc/TallShinnyQR.h:9-134, util/inc/kernelQR.h:215-414).
//
// Design (MI355X): one cooperative launch of G workgroups; workgroup g keeps
// rows [g*256*RPT, (g+1)*256*RPT) of the m x nb panel in registers as FP64 for
// the whole factorisation (the panel is read once and written once).  Column
// j needs sigma = ||x||^2 and p_c = x^T a_c for the trailing columns, x = the
// pivot column below the diagonal; both are sums over rows below the pivot
// and are reduced in ONE exchange per column, which also yields x^T v_i for
// the finished reflectors and with the pivot row the Gram column v_i^T v_j.
// After the exchange every workgroup forms the same LAPACK dlarfg reflector
// (beta = -sign(alpha) ||x||, tau = (beta - alpha)/beta, v = x / (alpha -
// beta)) and applies it to its rows.  The Gram matrix V^T V gives T by the
// dlarft forward/columnwise recurrence; W = V T is formed row-locally.
// Outputs: Y = V (unit lower, zeros above), W = V T, R, and optionally [R; 0]
// into the panel storage.
#include "common.hpp"

namespace evd {

// One butterfly level of a wave reduce-scatter over N live values: lanes with
// bit o set keep the upper half, the others the lower half, and each adds the
// half its partner kept.  Template recursion keeps every index static.
template <int NV, int N>
__device__ inline void reduce_scatter_level(double (&v)[NV], int lane) {
  if constexpr (N > 1) {
    constexpr int half = N / 2;
    constexpr int o = 32 * N / NV;
    const bool up = (lane & o) != 0;
#pragma unroll
    for (int i = 0; i < half; i++) {
      const double send = up ? v[i] : v[i + half];
      const double keep = up ? v[i + half] : v[i];
      v[i] = keep + __shfl_xor(send, o, 64);
    }
    reduce_scatter_level<NV, half>(v, lane);
  }
}

// Reduce NV (power of two, <= 64) per-lane values across the wave.  On exit
// lane l holds the total of value index l / (64 / NV).
template <int NV>
__device__ inline double wave_reduce_scatter(double (&v)[NV], int lane) {
  reduce_scatter_level<NV, NV>(v, lane);
  double r = v[0];
#pragma unroll
  for (int o = (64 / NV) >> 1; o > 0; o >>= 1) r += __shfl_xor(r, o, 64);
  return r;
}

constexpr unsigned long long panel_sentinel(unsigned epoch, int step) {
  return 0x7FF8000000000000ull | ((unsigned long long)(epoch & 0xFFFFFu) << 24) |
         (unsigned long long)((unsigned)step & 0xFFFFFFu);
}

// Register layout: every thread holds RPT rows of the panel, a[q][s], s a
// rotating slot index.  At the start of step j, column j + s sits in slot s
// for s < NB - j (the active columns; slot 0 is the pivot column) and
// finished column i sits in slot NB - j + i.  Each step writes the finished
// pivot column into the last slot while shifting the others down by one, so
// every register index is a compile-time constant and the step body is a
// short branch-free loop (a runtime column index would force select chains
// or private memory, and a fully unrolled column loop a code footprint far
// beyond the instruction cache).
template <int NB, int RPT, typename T>
__global__ __launch_bounds__(256) void panel_qr_kernel(int m, int nc, const T *__restrict__ Ain, long lda,
                                                       T *__restrict__ Y, long ldy, T *__restrict__ W,
                                                       long ldw, T *__restrict__ R, long ldr,
                                                       T *__restrict__ Aout, long ldo, double *__restrict__ part,
                                                       double *__restrict__ tau_out, unsigned *__restrict__ tags,
                                                       unsigned *__restrict__ epoch_word, unsigned *__restrict__ err,
                                                       long s_in, long s_y, long s_w, long s_out,
                                                       const int *__restrict__ only_bad) {
  constexpr int CH = NB < 32 ? NB : 32;  // slots per in-wave reduction
  constexpr int PASSG = 32;              // group records gathered per pass
  static_assert(NB <= 256, "the pivot rows must live in group 0");
  const int G = gridDim.x;
  const int g = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // same-size panels of several problems (one workgroup each, G = 1):
  // problem blockIdx.y at the given strides
  Ain += blockIdx.y * s_in;
  Y += blockIdx.y * s_y;
  W += blockIdx.y * s_w;
  if (Aout) Aout += blockIdx.y * s_out;
  part += blockIdx.y * (4L * G * 2 * NB);  // each problem's exchange ring and start tags
  tags += blockIdx.y * G;
  const int k = min(m, nc);
  const long rec = 2 * NB;  // exchange record: NB partial sums, NB pivot-row entries
  __shared__ double red[4][NB];
  __shared__ double tot[NB];
  __shared__ double piv[NB];
  __shared__ double wv[NB];
  __shared__ double sc[2];
  __shared__ double taus[NB];
  __shared__ double gram[NB][NB + 1];  // gram[i][j] = v_i^T v_j, i < j
  __shared__ double Tt[NB][NB + 1];    // Tt[c][l] = T[l][c]
  __shared__ double gpart[(PASSG + 1) * NB];
  __shared__ double gsum[256];
  __shared__ int s_ok;

  // fallback of the Cholesky QR panel: problems whose state word is not 1
  // (factored there) load nothing and leave after counting into the launch
  // epoch below
  const bool skip = only_bad && only_bad[blockIdx.y] != 1;
  double a[RPT][NB];
  long rowv[RPT];
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    const long r = (long)g * 256 * RPT + q * 256 + tid;
    rowv[q] = r;
#pragma unroll
    for (int c = 0; c < NB; c++) a[q][c] = (!skip && r < m && c < nc) ? (double)Ain[r + (long)c * lda] : 0.0;
  }
  for (int i = tid; i < NB * (NB + 1); i += 256) {
    (&gram[0][0])[i] = 0.0;
    (&Tt[0][0])[i] = 0.0;
  }
  // Launch epoch: one more than the last launch's, read from a device word
  // (so a launch replayed from a captured graph still gets a fresh epoch and
  // stale tags or sentinels of earlier launches never match).  Thread 0 of
  // every group reads the word, waits for the value, then counts its group
  // in (epoch_word[1]); the last group of the launch (all problems) resets
  // the count and advances the word, after every group has read it.
  __shared__ unsigned s_epoch;
  if (tid == 0) {
    unsigned e = (ld_dev(epoch_word) + 1u) & 0x1FFFFFFu;
    if (e == 0) e = 1;
    s_epoch = e;
    if (G > 1) {
      wait_mem();
      const unsigned total = (unsigned)G * gridDim.y;
      if (__hip_atomic_fetch_add(epoch_word + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == total - 1) {
        st_dev(epoch_word + 1, 0u);
        st_dev(epoch_word, e);
      }
    }
  }
  __syncthreads();
  const unsigned epoch = s_epoch;
  // (every group has counted itself into the epoch above, so the word advances)
  if (skip) return;
  if (G > 1) {
    // sentinels of steps 0 and 1 in this group's records, then a one-time
    // start barrier so nobody polls a slot before its owner marked it
    for (int st = 0; st < 2; st++) {
      double *r0 = part + ((long)st * G + g) * rec;
      for (int c = tid; c < (int)rec; c += 256) st_dev(&r0[c], __longlong_as_double((long long)panel_sentinel(epoch, st)));
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

  for (int j = 0; j < k; j++) {
    const int nact = NB - j;  // active slots [0, nact); finished slots [nact, NB)
    // ---- partial sums over this thread's rows, x = pivot column below the
    // diagonal: active slot s -> x^T a_s (s = 0 gives sigma = ||x||^2),
    // finished slot of column i -> x^T v_i(rows > j), which with the pivot row
    // gives the Gram entry v_i^T v_j = v_i[j] + scale x^T v_i(rows > j)
    double xq[RPT];
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      const long r = rowv[q];
      xq[q] = (r < m && r > j) ? a[q][0] : 0.0;
    }
#pragma unroll
    for (int cb = 0; cb < NB; cb += CH) {
      double e[CH];
#pragma unroll
      for (int cc = 0; cc < CH; cc++) e[cc] = 0.0;
#pragma unroll
      for (int q = 0; q < RPT; q++) {
#pragma unroll
        for (int cc = 0; cc < CH; cc++) e[cc] += xq[q] * a[q][cb + cc];
      }
      const double mine = wave_reduce_scatter<CH>(e, lane);
      if ((lane & (64 / CH - 1)) == 0) red[wave][cb + lane / (64 / CH)] = mine;
    }
    if (g == 0) {
#pragma unroll
      for (int q = 0; q < RPT; q++)
        if (rowv[q] == j) {
#pragma unroll
          for (int s = 0; s < NB; s++) piv[s] = a[q][s];
        }
    }
    if (tid == 0) s_ok = 1;
    __syncthreads();
    // ---- exchange between workgroups.  Every group stores its record (the
    // partial sums; group 0 also the pivot row) into ring slot j % 4 with
    // device-coherent stores; readers poll the values until none equals the
    // slot's sentinel (a NaN whose payload encodes the launch epoch and the
    // step).  The sentinel of step j + 2 is written at step j into the slot
    // that held step j - 2, which every reader consumed before it could write
    // its record of step j - 1; the wait before the next record orders the
    // sentinel ahead of it.
    if (G > 1) {
      const int slot = j & 3;
      wait_mem();
      if (tid < NB) {
        double *myrec = part + ((long)slot * G + g) * rec;
        st_dev(&myrec[tid], red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid]);
        if (g == 0) st_dev(&myrec[NB + tid], piv[tid]);
        const double snt = __longlong_as_double((long long)panel_sentinel(epoch, j + 2));
        double *nrec = part + ((long)((j + 2) & 3) * G + g) * rec;
        st_dev(&nrec[tid], snt);
        st_dev(&nrec[NB + tid], snt);
      }
      const unsigned long long want = panel_sentinel(epoch, j);
      double tacc = 0.0, pacc = 0.0;
      for (int g0 = 0; g0 < G; g0 += PASSG) {
        const int gn = min(PASSG, G - g0);
        const int items = gn * NB + (g0 == 0 ? NB : 0);
        for (int base = 0; base < items; base += 4 * 256) {
          double v[4];
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const int idx = base + u * 256 + tid;
            const int gg = idx / NB, s = idx - gg * NB;
            const double *src = (gg < gn) ? part + ((long)slot * G + g0 + gg) * rec + s
                                          : part + ((long)slot * G) * rec + NB + s;
            v[u] = idx < items ? ld_dev(src) : 0.0;
          }
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const int idx = base + u * 256 + tid;
            if (idx < items) {
              const int gg = idx / NB, s = idx - gg * NB;
              const double *src = (gg < gn) ? part + ((long)slot * G + g0 + gg) * rec + s
                                            : part + ((long)slot * G) * rec + NB + s;
              const unsigned long long t0 = clock100();
              while ((unsigned long long)__double_as_longlong(v[u]) == want) {
                __builtin_amdgcn_s_sleep(1);
                v[u] = ld_dev(src);
                if (clock100() - t0 > kWatchdogTicks) {
                  atomicOr(err, (unsigned)EVD_E_TIMEOUT);
                  s_ok = 0;
                  break;
                }
              }
              gpart[idx] = v[u];
            }
          }
        }
        __syncthreads();
        {
          // the gn records of slot s summed by 256 / NB threads (strided
          // partial sums), then the partials in order: a short dependent
          // chain instead of gn LDS reads in a row
          constexpr int P = 256 / NB;
          const int sl = tid % NB, part = tid / NB;
          double acc = 0.0;
          for (int gg = part; gg < gn; gg += P) acc += gpart[gg * NB + sl];
          gsum[tid] = acc;
        }
        __syncthreads();
        if (tid < NB) {
#pragma unroll
          for (int q = 0; q < 256 / NB; q++) tacc += gsum[q * NB + tid];
          if (g0 == 0) pacc = gpart[gn * NB + tid];
        }
        __syncthreads();
      }
      if (tid < NB) {
        tot[tid] = tacc;
        piv[tid] = pacc;
      }
      __syncthreads();
      if (!s_ok) return;
    }
    // ---- reflector (LAPACK dlarfg), w_s = tau v^T a_s, Gram column j
    if (tid < NB) {
      const int s = tid;
      double ts, t0;
      if (G == 1) {
        ts = red[0][s] + red[1][s] + red[2][s] + red[3][s];
        t0 = red[0][0] + red[1][0] + red[2][0] + red[3][0];
      } else {
        ts = tot[s];
        t0 = tot[0];
      }
      const double alpha = piv[0], sigma = t0;
      double beta, tau, scale;
      if (sigma == 0.0) {
        beta = alpha;
        tau = 0.0;
        scale = 0.0;
      } else {
        const double nrm = sqrt(alpha * alpha + sigma);
        beta = alpha >= 0.0 ? -nrm : nrm;
        tau = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
      }
      wv[s] = (s >= 1 && s < nact) ? tau * (piv[s] + scale * ts) : 0.0;
      if (s >= nact) gram[s - nact][j] = piv[s] + scale * ts;
      if (s == 0) {
        sc[0] = beta;
        sc[1] = scale;
        taus[j] = tau;
      }
    }
    __syncthreads();
    // ---- apply H_j to the trailing columns and rotate the slots
    const double beta = sc[0], scale = sc[1];
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      const long r = rowv[q];
      const double x = a[q][0];
      const double v = (r > j) ? x * scale : (r == j ? 1.0 : 0.0);
      const double fin = (r > j) ? v : (r == j ? beta : x);
#pragma unroll
      for (int s = 0; s < NB - 1; s++) a[q][s] = a[q][s + 1] - v * wv[s + 1];
      a[q][NB - 1] = fin;
    }
  }
  // columns back to natural order (column c in slot c)
  for (int t = k; t < NB; t++) {
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      const double first = a[q][0];
#pragma unroll
      for (int s = 0; s < NB - 1; s++) a[q][s] = a[q][s + 1];
      a[q][NB - 1] = first;
    }
  }
  __syncthreads();

  // ---- T by the dlarft forward/columnwise recurrence, one row per thread:
  // T[i][j] = -tau_j sum_{l=i}^{j-1} T[i][l] (v_l^T v_j)
  if (tid < k) {
    const int i = tid;
    Tt[i][i] = taus[i];
    for (int jj = i + 1; jj < k; jj++) {
      double s = 0.0;
      for (int l = i; l < jj; l++) s += Tt[l][i] * gram[l][jj];
      Tt[jj][i] = -taus[jj] * s;
    }
    if (g == 0 && tau_out) tau_out[i] = taus[i];
  }
  __syncthreads();

  // ---- outputs: R / [R;0], then a[] becomes V in place, Y = V, W = V T
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    const long r = rowv[q];
    if (r < m) {
#pragma unroll
      for (int c = 0; c < NB; c++) {
        if (c >= nc) continue;
        const double rv = (r <= c) ? a[q][c] : 0.0;
        if (R && r < nc) R[r + (long)c * ldr] = (T)rv;
        if (Aout) Aout[r + (long)c * ldo] = (T)rv;
      }
    }
#pragma unroll
    for (int c = 0; c < NB; c++) a[q][c] = (c < k) ? ((r > c) ? a[q][c] : (r == c ? 1.0 : 0.0)) : 0.0;
    if (r < m && Y) {
#pragma unroll
      for (int c = 0; c < NB; c++)
        if (c < nc) Y[r + (long)c * ldy] = (T)a[q][c];
    }
  }
  if (W) {
    for (int c = 0; c < nc; c++) {
#pragma unroll
      for (int q = 0; q < RPT; q++) {
        const long r = rowv[q];
        double w = 0.0;
#pragma unroll
        for (int l = 0; l < NB; l++) w += a[q][l] * Tt[c][l];
        if (r < m) W[r + (long)c * ldw] = (T)w;
      }
    }
  }
}

template <int NB, int RPT, typename T>
static int launch_panel(evd_ctx *ctx, long m, long nc, const T *A, long lda, T *Y, long ldy, T *W, long ldw, T *R,
                        long ldr, T *Aout, long ldo, double *tau_out, int count = 1, long s_in = 0, long s_y = 0,
                        long s_w = 0, long s_out = 0, const int *only_bad = nullptr) {
  const long rows_per = 256L * RPT;
  int G = (int)((m + rows_per - 1) / rows_per);
  if (count > 1 && (R || tau_out)) return set_error(ctx, EVD_E_UNSUPPORTED, "batched panels write no R / tau output");
  int maxres = 0;
  EVD_TRY(max_coresident_blocks(ctx, (const void *)panel_qr_kernel<NB, RPT, T>, 256, 0, &maxres));
  if ((long)G * count > maxres)
    return set_error(ctx, EVD_E_UNSUPPORTED, "%d panels of %ld rows need %ld co-resident groups (max %d)", count, m,
                     (long)G * count, maxres);
  ArenaFrame frame(ctx);
  double *part = arena_alloc_n<double>(ctx, (size_t)4 * G * 2 * NB * count);
  if (!part) return EVD_E_NOMEM;
  if ((long)G * count > kTagWords)
    return set_error(ctx, EVD_E_UNSUPPORTED, "panels need %ld exchange tags", (long)G * count);
  unsigned *tags = ctx->tags;
  unsigned *epoch_word = ctx->ctrl + kCtrlEpoch;
  unsigned *err = ctx->ctrl + kCtrlError;
  int mi = (int)m, ni = (int)nc;
  void *args[] = {&mi, &ni, (void *)&A, &lda, &Y, &ldy, &W, &ldw, &R, &ldr, &Aout, &ldo, &part, &tau_out,
                  &tags, &epoch_word, &err, &s_in, &s_y, &s_w, &s_out, (void *)&only_bad};
  // one workgroup is trivially co-resident: a regular launch, which (unlike a
  // cooperative one) may overlap with the work of other streams
  if (G == 1)
    EVD_HIP(ctx, hipLaunchKernel((const void *)panel_qr_kernel<NB, RPT, T>, dim3(G, count), dim3(256), args, 0,
                                 ctx->stream));
  else
    EVD_HIP(ctx, hipLaunchCooperativeKernel((const void *)panel_qr_kernel<NB, RPT, T>, dim3(G, count), dim3(256), args,
                                            0, ctx->stream));
  return 0;
}

template <typename T>
int panel_cholqr(evd_ctx *ctx, int count, long m, long nc, const T *A, long lda, long s_in, T *Y, long ldy, long s_y,
                 T *W, long ldw, long s_w, T *R, long ldr, T *Aout, long ldo, long s_out, double *tau_out, int *bad);

// Whether a panel goes through the Cholesky QR path (panel_cholqr.hip): up to
// 32 columns and at least twice as many rows, unless the panel_exchange knob
// keeps the column-exchange kernel.
static bool use_cholqr(const evd_ctx *ctx, long m, long nc) {
  return !ctx->tune.panel_exchange && nc <= 32 && m >= 2 * nc;
}

// Panels of 33..64 FP32 columns (b = 64) as two halves on the Cholesky QR
// path: Q = Q1 Q2 with Q1 from the first 32 columns, the rest updated by
// Q1^T and factored from row 32 on.  In WY form
//   V = [V1, [0; V2]],  W = V T = [W1, W2 - W1 (V1^T W2)],
//   R = [R11 R12; 0 R22],  R12 = the top 32 rows of Q1^T A2,
// which is what factoring column by column gives (same reflector signs),
// within rounding.
static bool use_cholqr_halves(const evd_ctx *ctx, long m, long nc) {
  return !ctx->tune.panel_exchange && nc > 32 && nc <= 64 && m >= 2 * nc;
}

template <typename T>
int panel_qr_t(evd_ctx *ctx, long m, long nc, const T *A, long lda, T *Y, long ldy, T *W, long ldw, T *R, long ldr,
               T *Aout, long ldo, double *tau_out);

// Single panel: A (read) may alias Y or Aout, as in the callers.
static int panel_halves(evd_ctx *ctx, long m, long nc, const float *A, long lda, float *Y, long ldy, float *W,
                        long ldw, float *R, long ldr, float *Aout, long ldo, double *tau_out) {
  const long n2 = nc - 32;
  ArenaFrame frame(ctx);
  float *P2 = arena_alloc_n<float>(ctx, (size_t)m * n2);
  float *X = arena_alloc_n<float>(ctx, (size_t)32 * n2);
  if (!P2 || !X) return EVD_E_NOMEM;
  EVD_TRY(copy_f32(ctx, A + 32 * lda, lda, P2, m, m, n2));
  EVD_TRY(panel_qr_t<float>(ctx, m, 32, A, lda, Y, ldy, W, ldw, R, ldr, Aout, ldo, tau_out));
  // A2 <- Q1^T A2 = A2 - V1 (W1^T A2)
  EVD_TRY(sgemm(ctx, 'T', 'N', 32, n2, m, 1.f, W, ldw, P2, m, 0.f, X, 32));
  EVD_TRY(sgemm(ctx, 'N', 'N', m, n2, 32, -1.f, Y, ldy, X, 32, 1.f, P2, m));
  // the second half from row 32 ([R22; 0] back into P2)
  EVD_TRY(panel_qr_t<float>(ctx, m - 32, n2, P2 + 32, m, Y + 32 + 32 * ldy, ldy, W + 32 + 32 * ldw, ldw,
                            R ? R + 32 + 32 * ldr : nullptr, ldr, P2 + 32, m, tau_out ? tau_out + 32 : nullptr));
  if (Aout) EVD_TRY(copy_f32(ctx, P2, m, Aout + 32 * ldo, ldo, m, n2));
  if (R) {
    EVD_TRY(copy_f32(ctx, P2, m, R + 32 * ldr, ldr, 32, n2));
    EVD_TRY(fill_f32(ctx, R + 32, n2, 32, ldr, 0.f));
  }
  EVD_TRY(fill_f32(ctx, Y + 32 * ldy, 32, n2, ldy, 0.f));
  EVD_TRY(fill_f32(ctx, W + 32 * ldw, 32, n2, ldw, 0.f));
  // W2 <- W2 - W1 (V1^T W2) (rows 0..31 of W2 are zero)
  EVD_TRY(sgemm(ctx, 'T', 'N', 32, n2, m - 32, 1.f, Y + 32, ldy, W + 32 + 32 * ldw, ldw, 0.f, X, 32));
  EVD_TRY(sgemm(ctx, 'N', 'N', m, n2, 32, -1.f, W, ldw, X, 32, 1.f, W + 32 * ldw, ldw));
  return 0;
}

// Panel factorisation used by the SBR drivers (FP32 and FP64 storage; the
// arithmetic is FP64 in both) and by evd_panel_qr_f32 / _f64.
template <typename T>
int panel_qr_t(evd_ctx *ctx, long m, long nc, const T *A, long lda, T *Y, long ldy, T *W, long ldw, T *R, long ldr,
               T *Aout, long ldo, double *tau_out) {
  if (m <= 0 || nc <= 0) return 0;
  if constexpr (sizeof(T) == 4) {
    if (use_cholqr_halves(ctx, m, nc)) return panel_halves(ctx, m, nc, A, lda, Y, ldy, W, ldw, R, ldr, Aout, ldo, tau_out);
  }
  if (use_cholqr(ctx, m, nc)) {
    // Cholesky QR + reconstruction; the column-exchange kernel runs after it
    // and factors the panel only if it was flagged
    ArenaFrame frame(ctx);
    int *bad = arena_alloc_n<int>(ctx, 1);
    if (!bad) return EVD_E_NOMEM;
    EVD_TRY(panel_cholqr<T>(ctx, 1, m, nc, A, lda, 0, Y, ldy, 0, W, ldw, 0, R, ldr, Aout, ldo, 0, tau_out, bad));
    if (ctx->tune.panel_rows == 256)
      return launch_panel<32, 1, T>(ctx, m, nc, A, lda, Y, ldy, W, ldw, R, ldr, Aout, ldo, tau_out, 1, 0, 0, 0, 0,
                                    bad);
    return launch_panel<32, 2, T>(ctx, m, nc, A, lda, Y, ldy, W, ldw, R, ldr, Aout, ldo, tau_out, 1, 0, 0, 0, 0, bad);
  }
  // rows per workgroup (the panel_rows knob): 512 (two per thread) by
  // default; 256 doubles the workgroups that share the column exchanges
  if (nc <= 32 && ctx->tune.panel_rows == 256)
    return launch_panel<32, 1, T>(ctx, m, nc, A, lda, Y, ldy, W, ldw, R, ldr, Aout, ldo, tau_out);
  if (nc <= 32) return launch_panel<32, 2, T>(ctx, m, nc, A, lda, Y, ldy, W, ldw, R, ldr, Aout, ldo, tau_out);
  if (nc <= 64) return launch_panel<64, 1, T>(ctx, m, nc, A, lda, Y, ldy, W, ldw, R, ldr, Aout, ldo, tau_out);
  return set_error(ctx, EVD_E_UNSUPPORTED, "panel width %ld > 64", nc);
}
int panel_qr(evd_ctx *ctx, long m, long nc, const float *A, long lda, float *Y, long ldy, float *W, long ldw,
             float *R, long ldr, float *Aout, long ldo, double *tau_out) {
  return panel_qr_t<float>(ctx, m, nc, A, lda, Y, ldy, W, ldw, R, ldr, Aout, ldo, tau_out);
}
int panel_qr_f64(evd_ctx *ctx, long m, long nc, const double *A, long lda, double *Y, long ldy, double *W, long ldw,
                 double *R, long ldr, double *Aout, long ldo, double *tau_out) {
  return panel_qr_t<double>(ctx, m, nc, A, lda, Y, ldy, W, ldw, R, ldr, Aout, ldo, tau_out);
}

// Same-size panels of `count` problems (A, Y, W of problem q at q times the
// strides; the factored block is written back into A) in one launch, grid
// (G, count), when all G * count workgroups can be co-resident; else per
// problem.  Each problem's arithmetic is that of panel_qr.
int panel_qr_batched(evd_ctx *ctx, int count, long m, long nc, float *A, long lda, long sA, float *Y, long ldy,
                     long sY, float *W, long ldw, long sW) {
  if (m <= 0 || nc <= 0 || count <= 0) return 0;
  if (use_cholqr_halves(ctx, m, nc)) {
    // panel_halves for every problem, each step one launch for all of them
    const long n2 = nc - 32;
    ArenaFrame frame(ctx);
    float *P2 = arena_alloc_n<float>(ctx, (size_t)count * m * n2);
    float *X = arena_alloc_n<float>(ctx, (size_t)count * 32 * n2);
    if (!P2 || !X) return EVD_E_NOMEM;
    const long sP = m * n2, sX = 32 * n2;
    for (int q = 0; q < count; q++) EVD_TRY(copy_f32(ctx, A + q * sA + 32 * lda, lda, P2 + q * sP, m, m, n2));
    EVD_TRY(panel_qr_batched(ctx, count, m, 32, A, lda, sA, Y, ldy, sY, W, ldw, sW));
    EVD_TRY(sgemm(ctx, 'T', 'N', 32, n2, m, 1.f, W, ldw, P2, m, 0.f, X, 32, count, sW, sP, sX));
    EVD_TRY(sgemm(ctx, 'N', 'N', m, n2, 32, -1.f, Y, ldy, X, 32, 1.f, P2, m, count, sY, sX, sP));
    EVD_TRY(panel_qr_batched(ctx, count, m - 32, n2, P2 + 32, m, sP, Y + 32 + 32 * ldy, ldy, sY, W + 32 + 32 * ldw,
                             ldw, sW));
    for (int q = 0; q < count; q++) {
      EVD_TRY(copy_f32(ctx, P2 + q * sP, m, A + q * sA + 32 * lda, lda, m, n2));
      EVD_TRY(fill_f32(ctx, Y + q * sY + 32 * ldy, 32, n2, ldy, 0.f));
      EVD_TRY(fill_f32(ctx, W + q * sW + 32 * ldw, 32, n2, ldw, 0.f));
    }
    EVD_TRY(sgemm(ctx, 'T', 'N', 32, n2, m - 32, 1.f, Y + 32, ldy, W + 32 + 32 * ldw, ldw, 0.f, X, 32, count, sY, sW,
                  sX));
    return sgemm(ctx, 'N', 'N', m, n2, 32, -1.f, W, ldw, X, 32, 1.f, W + 32 * ldw, ldw, count, sW, sX, sW);
  }
  if (use_cholqr(ctx, m, nc)) {
    // every problem's panel through the Cholesky QR path in one launch each,
    // then the fallback for the flagged ones (one launch when the groups of
    // all problems can be co-resident, else per problem)
    ArenaFrame frame(ctx);
    int *bad = arena_alloc_n<int>(ctx, (size_t)count);
    if (!bad) return EVD_E_NOMEM;
    EVD_TRY(panel_cholqr<float>(ctx, count, m, nc, A, lda, sA, Y, ldy, sY, W, ldw, sW, nullptr, 0, A, lda, sA,
                                nullptr, bad));
    const long G = (m + 511) / 512;
    int maxres = 0;
    if (count > 1 && G > 1)
      EVD_TRY(max_coresident_blocks(ctx, (const void *)panel_qr_kernel<32, 2, float>, 256, 0, &maxres));
    if (count == 1 || G == 1 || (G * count <= maxres && G * count <= kTagWords))
      return launch_panel<32, 2, float>(ctx, m, nc, A, lda, Y, ldy, W, ldw, nullptr, 0, A, lda, nullptr, count, sA,
                                        sY, sW, sA, bad);
    for (int q = 0; q < count; q++) {
      const int st = launch_panel<32, 2, float>(ctx, m, nc, A + q * sA, lda, Y + q * sY, ldy, W + q * sW, ldw,
                                                nullptr, 0, A + q * sA, lda, nullptr, 1, 0, 0, 0, 0, bad + q);
      if (st) return st;
    }
    return 0;
  }
  const long rows_per = nc <= 32 ? 512 : 256;
  const long G = (m + rows_per - 1) / rows_per;
  int maxres = 0;
  if (count > 1 && G > 1) {
    if (nc <= 32)
      EVD_TRY(max_coresident_blocks(ctx, (const void *)panel_qr_kernel<32, 2, float>, 256, 0, &maxres));
    else
      EVD_TRY(max_coresident_blocks(ctx, (const void *)panel_qr_kernel<64, 1, float>, 256, 0, &maxres));
  }
  if (count > 1 && nc <= 64 && (G == 1 || (G * count <= maxres && G * count <= kTagWords))) {
    if (nc <= 32)
      return launch_panel<32, 2, float>(ctx, m, nc, A, lda, Y, ldy, W, ldw, nullptr, 0, A, lda, nullptr, count, sA,
                                        sY, sW, sA);
    if (nc <= 64)
      return launch_panel<64, 1, float>(ctx, m, nc, A, lda, Y, ldy, W, ldw, nullptr, 0, A, lda, nullptr, count, sA,
                                        sY, sW, sA);
  }
  for (int q = 0; q < count; q++)
    EVD_TRY(panel_qr(ctx, m, nc, A + q * sA, lda, Y + q * sY, ldy, W + q * sW, ldw, nullptr, 0, A + q * sA, lda,
                     nullptr));
  return 0;
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
