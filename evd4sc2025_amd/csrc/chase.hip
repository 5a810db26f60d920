// Band -> tridiagonal bulge chasing (the role of chasing_kernel_one_timeV10
// in util/inc/bugle_chasingV10.h:21-392 and its driver src/EVD/myEVD.cu:365-451).
//
// Algorithm (same arithmetic as the reference): sweep s annihilates column s
// below the subdiagonal with a reflector of length <= b on rows s+1..s+b and
// chases the bulge down in steps of b rows.  One step at row p:
//   x = first column of B1 (the block left of the diagonal block), H = I - u u^T
//   with u = (x/||x|| + sign(x0) e1)/sqrt(|.|) so that ||u||^2 = 2 and
//   Hx = -sign(x0)||x|| e1;  B1 <- H B1;  S <- H S H (diagonal block);
//   N <- N H (block below S, which becomes the next B1 - the bulge).
// u is stored in U[p .. p+len-1, s] ("noBandU" layout, FP32).
//
// MI355X design: persistent cooperative grid; workgroup g runs sweeps
// g, g+G, ...; the three b x b windows (B1, S, N) live in LDS in FP64; the
// block below (N) stays in LDS as the next step's B1, so each step reads S
// and N and writes B1 and S.  Sweeps are ordered through per-sweep progress
// flags in HBM (release/acquire at agent scope, no grid barrier): sweep s may
// run its step at row p once sweep s-1 has published a next-row >= p + 2b.
// Every wait is bounded by a watchdog so the grid always drains.
#include <climits>

#include "common.hpp"

namespace evd {

template <int BW>
__global__ __launch_bounds__(256) void chase_kernel(int n, int b, double *__restrict__ band, long ldb,
                                                    float *__restrict__ U, long ldu, int *__restrict__ com,
                                                    unsigned *__restrict__ err) {
  constexpr int LD = BW + 1;
  __shared__ double sA[BW * LD], sBuf[BW * LD], sS[BW * LD];
  __shared__ double u[BW], tv[BW], yv[BW], qv[BW], wv[BW];
  __shared__ double s_beta;
  __shared__ int s_ok;
  const int G = gridDim.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nsweeps = n - 2;
  double *B1 = sA, *Nb = sBuf;
  double *S = sS;

  for (int s = blockIdx.x; s < nsweeps; s += G) {
    int p = s + 1;
    int len = min(b, n - p);
    int colB1 = 1, cB1 = s;
    bool first = true;
    __syncthreads();
    while (true) {
      // ---- wait until sweep s-1 is 2b rows ahead
      if (tid == 0) {
        s_ok = 1;
        if (s > 0) {
          const int need = p + 2 * b;
          unsigned long long t0 = clock100();
          while (ld_dev(&com[s - 1]) < need) {
            __builtin_amdgcn_s_sleep(1);
            if (clock100() - t0 > kWatchdogTicks) {
              atomicOr(err, (unsigned)EVD_E_TIMEOUT);
              s_ok = 0;
              break;
            }
          }
        }
      }
      __syncthreads();
      if (!s_ok) {
        if (tid == 0) st_dev(&com[s], INT_MAX);
        return;
      }
      if (first) {
        // column s is final once sweep s-1 has passed row s+1+2b
        for (int i = tid; i < len; i += 256) B1[i * LD + 0] = ld_dev(&band[(p + i - s) + (long)s * ldb]);
        first = false;
      }
      const int len2 = max(0, min(b, n - p - len));
      // ---- load S (lower -> full) and N
      for (int idx = tid; idx < len * len; idx += 256) {
        int i = idx % len, kk = idx / len;
        if (i >= kk) {
          double v = ld_dev(&band[(i - kk) + (long)(p + kk) * ldb]);
          S[i * LD + kk] = v;
          S[kk * LD + i] = v;
        }
      }
      for (int idx = tid; idx < len2 * len; idx += 256) {
        int i = idx % len2, kk = idx / len2;
        Nb[i * LD + kk] = ld_dev(&band[(len + i - kk) + (long)(p + kk) * ldb]);
      }
      __syncthreads();
      // ---- reflector from B1[:, 0]
      if (wave == 0) {
        double x = (lane < len) ? B1[lane * LD] : 0.0;
        double nrm = sqrt(wave_sum(x * x));
        double x0 = B1[0];
        double uu = 0.0;
        if (nrm > 0.0) {
          double inv = 1.0 / nrm;
          double u0 = x0 * inv + (x0 >= 0.0 ? 1.0 : -1.0);
          double scl = 1.0 / sqrt(fabs(u0));
          uu = (lane == 0 ? u0 : x * inv) * scl;
        }
        if (lane < BW) u[lane] = (lane < len) ? uu : 0.0;
        if (lane < len) U[(p + lane) + (long)s * ldu] = (float)uu;
        if (lane == 0) s_beta = (nrm > 0.0) ? (x0 >= 0.0 ? -nrm : nrm) : x0;
      }
      __syncthreads();
      // ---- matrix-vector products: t = B1^T u, y = S u, q = N u
      if (wave == 0 && lane < colB1) {
        double acc = 0.0;
        for (int i = 0; i < len; i++) acc += u[i] * B1[i * LD + lane];
        tv[lane] = acc;
      } else if (wave == 1 && lane < len) {
        double acc = 0.0;
        for (int kk = 0; kk < len; kk++) acc += S[lane * LD + kk] * u[kk];
        yv[lane] = acc;
      } else if (wave == 2 && lane < len2) {
        double acc = 0.0;
        for (int kk = 0; kk < len; kk++) acc += Nb[lane * LD + kk] * u[kk];
        qv[lane] = acc;
      }
      __syncthreads();
      if (wave == 0) {
        double pr = (lane < len) ? u[lane] * yv[lane] : 0.0;
        double psi = wave_sum(pr);
        if (lane < len) wv[lane] = yv[lane] - 0.5 * psi * u[lane];
      }
      __syncthreads();
      // ---- rank-1 / rank-2 updates
      for (int idx = tid; idx < len * colB1; idx += 256) {
        int i = idx % len, c = idx / len;
        if (c == 0)
          B1[i * LD] = (i == 0) ? s_beta : 0.0;
        else
          B1[i * LD + c] -= u[i] * tv[c];
      }
      for (int idx = tid; idx < len * len; idx += 256) {
        int i = idx % len, kk = idx / len;
        S[i * LD + kk] -= u[i] * wv[kk] + wv[i] * u[kk];
      }
      for (int idx = tid; idx < len2 * len; idx += 256) {
        int i = idx % len2, kk = idx / len2;
        Nb[i * LD + kk] -= qv[i] * u[kk];
      }
      __syncthreads();
      // ---- write back B1 and S (and N at the end of the sweep)
      for (int idx = tid; idx < len * colB1; idx += 256) {
        int i = idx % len, c = idx / len;
        st_dev(&band[(p + i - (cB1 + c)) + (long)(cB1 + c) * ldb], B1[i * LD + c]);
      }
      for (int idx = tid; idx < len * len; idx += 256) {
        int i = idx % len, kk = idx / len;
        if (i >= kk) st_dev(&band[(i - kk) + (long)(p + kk) * ldb], S[i * LD + kk]);
      }
      const bool last = (len2 <= 1);
      if (last) {
        for (int idx = tid; idx < len2 * len; idx += 256) {
          int i = idx % len2, kk = idx / len2;
          st_dev(&band[(len + i - kk) + (long)(p + kk) * ldb], Nb[i * LD + kk]);
        }
      }
      wait_mem();
      __syncthreads();
      if (tid == 0) st_dev(&com[s], last ? INT_MAX : p + len);
      if (last) break;
      double *t = B1;
      B1 = Nb;
      Nb = t;
      colB1 = len;
      cB1 = p;
      p += len;
      len = len2;
    }
  }
}

__global__ void band_pack_kernel(int n, int b, const float *__restrict__ A, long lda, double *__restrict__ band,
                                 long ldb) {
  long c = blockIdx.y * (long)blockDim.y + threadIdx.y;
  for (; c < n; c += (long)gridDim.y * blockDim.y)
    for (int i = threadIdx.x; i < 2 * b; i += blockDim.x) {
      double v = 0.0;
      if (i <= b && c + i < n) v = (double)A[(c + i) + c * lda];
      band[i + c * ldb] = v;
    }
}

__global__ void band_extract_kernel(int n, const double *__restrict__ band, long ldb, double *__restrict__ D,
                                    double *__restrict__ E) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    D[i] = band[i * ldb];
    if (i < n - 1) E[i] = band[1 + i * ldb];
  }
}

int band_pack(evd_ctx *ctx, long n, long b, const float *A, long lda, double *band, long ldband) {
  if (n <= 0) return 0;
  dim3 blk(64, 4);
  unsigned gy = (unsigned)std::min<long>((n + 3) / 4, 65535);
  band_pack_kernel<<<dim3(1, gy), blk, 0, ctx->stream>>>((int)n, (int)b, A, lda, band, ldband);
  EVD_LAUNCHED(ctx);
  return 0;
}

template <int BW>
static int launch_chase(evd_ctx *ctx, long n, long b, double *band, long ldb, float *U, long ldu, int *com) {
  int maxres = 0;
  EVD_TRY(max_coresident_blocks(ctx, (const void *)chase_kernel<BW>, 256, 0, &maxres));
  long nsweeps = n - 2;
  int G = ctx->tune.chase_groups > 0 ? ctx->tune.chase_groups : ctx->num_cus;
  G = (int)std::min<long>(std::min<long>(G, maxres), std::max<long>(1, nsweeps));
  int ni = (int)n, bi = (int)b;
  unsigned *err = ctx->ctrl + kCtrlError;
  void *args[] = {&ni, &bi, &band, &ldb, &U, &ldu, &com, &err};
  EVD_HIP(ctx, hipLaunchCooperativeKernel((const void *)chase_kernel<BW>, dim3(G), dim3(256), args, 0, ctx->stream));
  return 0;
}

int sb2st(evd_ctx *ctx, long n, long b, double *band, long ldband, float *U, long ldu, double *D, double *E) {
  if (n <= 0) return 0;
  EVD_TRY(fill_f32(ctx, U, ldu, n, ldu, 0.f));
  if (n > 2) {
    ArenaFrame frame(ctx);
    int *com = arena_alloc_n<int>(ctx, (size_t)n);
    if (!com) return EVD_E_NOMEM;
    EVD_TRY(fill_u32(ctx, reinterpret_cast<unsigned *>(com), (size_t)n, 0u));
    if (b <= 32)
      EVD_TRY(launch_chase<32>(ctx, n, b, band, ldband, U, ldu, com));
    else if (b <= 64)
      EVD_TRY(launch_chase<64>(ctx, n, b, band, ldband, U, ldu, com));
    else
      return set_error(ctx, EVD_E_UNSUPPORTED, "bandwidth %ld > 64", b);
  }
  band_extract_kernel<<<(unsigned)std::min<long>((n + 255) / 256, 1024), 256, 0, ctx->stream>>>((int)n, band, ldband,
                                                                                               D, E);
  EVD_LAUNCHED(ctx);
  return 0;
}

}  // namespace evd

using namespace evd;

extern "C" int evd_band_pack_f32(evd_ctx *ctx, long n, long b, const float *A, long lda, double *band, long ldband) {
  if (!ctx) return -1;
  EVD_ARG(ctx, n >= 0, 2);
  EVD_ARG(ctx, b >= 1 && b <= 64, 3);
  EVD_ARG(ctx, lda >= std::max<long>(1, n), 5);
  EVD_ARG(ctx, ldband >= 2 * b, 7);
  return band_pack(ctx, n, b, A, lda, band, ldband);
}

extern "C" int evd_sb2st_f64(evd_ctx *ctx, long n, long b, double *band, long ldband, float *U, long ldu, double *D,
                             double *E) {
  if (!ctx) return -1;
  EVD_ARG(ctx, n >= 0, 2);
  EVD_ARG(ctx, b >= 1 && b <= 64, 3);
  EVD_ARG(ctx, ldband >= 2 * b, 5);
  EVD_ARG(ctx, ldu >= n + b, 7);
  if (ctx->arena_top == 0) EVD_TRY(arena_ensure(ctx, (size_t)n * sizeof(int) + 4096));
  EVD_TRY(sb2st(ctx, n, b, band, ldband, U, ldu, D, E));
  return check_device_error(ctx, "evd_sb2st_f64");
}
