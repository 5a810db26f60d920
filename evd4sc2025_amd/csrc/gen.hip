// Seeded synthetic inputs, generated on the device.
//
// Random stream: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), counter
// = (index lo, index hi, stream, 0), key = (seed lo, seed hi).  Uniform
// doubles take 53 bits from the first two output words; normals use
// Box-Muller on the four words.  The same mapping is restated in NumPy in the
// test oracle so CPU and GPU see identical matrices.
//
//  * evd_gen_symmetric_f32: A = (X + X^T)/2, X i.i.d. uniform[-1,1] or N(0,1).
//  * evd_gen_spectrum_f32:  A = Q diag(lambda) Q^T, Q = H_1 ... H_n with
//    H_k = I - 2 v v^T / v^T v, v ~ N(0, I); built in FP64 in blocks of 32
//    reflectors (compact WY, T^{-1} = striu(V^T V) + diag(V^T V)/2) with the
//    two-sided ZY update, then rounded to FP32.
#include "common.hpp"

namespace evd {

struct Philox {
  __device__ static inline void round(unsigned (&c)[4], unsigned k0, unsigned k1) {
    const unsigned long long p0 = (unsigned long long)0xD2511F53u * c[0];
    const unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c[2];
    const unsigned hi0 = (unsigned)(p0 >> 32), lo0 = (unsigned)p0;
    const unsigned hi1 = (unsigned)(p1 >> 32), lo1 = (unsigned)p1;
    unsigned n0 = hi1 ^ c[1] ^ k0, n1 = lo1, n2 = hi0 ^ c[3] ^ k1, n3 = lo0;
    c[0] = n0;
    c[1] = n1;
    c[2] = n2;
    c[3] = n3;
  }
  __device__ static inline void gen(unsigned long long idx, unsigned stream, unsigned long long seed,
                                    unsigned (&out)[4]) {
    out[0] = (unsigned)idx;
    out[1] = (unsigned)(idx >> 32);
    out[2] = stream;
    out[3] = 0u;
    unsigned k0 = (unsigned)seed, k1 = (unsigned)(seed >> 32);
#pragma unroll
    for (int r = 0; r < 10; r++) {
      round(out, k0, k1);
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
  }
};

__device__ inline double u53(unsigned a, unsigned b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

__device__ inline double rand_entry(unsigned long long idx, unsigned stream, unsigned long long seed, int kind) {
  unsigned w[4];
  Philox::gen(idx, stream, seed, w);
  if (kind == 0) return 2.0 * u53(w[0], w[1]) - 1.0;
  const double u1 = u53(w[0], w[1]) + 0.5 / 9007199254740992.0;
  const double u2 = u53(w[2], w[3]);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
}

__global__ void gen_sym_kernel(long n, unsigned long long seed, int kind, float *__restrict__ A, long lda) {
  for (long c = blockIdx.y; c < n; c += gridDim.y)
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x) {
      const double x = rand_entry((unsigned long long)(r + c * n), 0u, seed, kind);
      const double y = rand_entry((unsigned long long)(c + r * n), 0u, seed, kind);
      A[r + c * lda] = (float)(0.5 * (x + y));
    }
}

__global__ void gen_normal_block_kernel(long n, int nv, long col0, unsigned long long seed, double *__restrict__ V) {
  for (long c = blockIdx.y; c < nv; c += gridDim.y)
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x)
      V[r + c * n] = rand_entry((unsigned long long)(r + (col0 + c) * n), 1u, seed, 1);
}

__global__ void diag_f64_kernel(long n, const double *__restrict__ lam, double *__restrict__ A) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    A[i + i * n] = lam[i];
}

// T = (striu(G) + diag(G)/2)^{-1} for nv <= 32 (single block, thread j solves column j).
__global__ void wy_t_from_gram_kernel(int nv, const double *__restrict__ G, double *__restrict__ T) {
  __shared__ double Ti[32][33];
  const int t = threadIdx.x;
  for (int i = 0; i < nv; i++)
    if (t < nv) Ti[i][t] = (i < t) ? G[i + t * nv] : (i == t ? 0.5 * G[i + i * nv] : 0.0);
  __syncthreads();
  if (t < nv) {
    double x[32];
    for (int i = 0; i < 32; i++) x[i] = 0.0;
    // solve Ti * x = e_t (upper triangular)
    for (int i = t; i >= 0; i--) {
      double s = (i == t) ? 1.0 : 0.0;
      for (int l = i + 1; l <= t; l++) s -= Ti[i][l] * x[l];
      x[i] = s / Ti[i][i];
    }
    for (int i = 0; i < nv; i++) T[i + t * nv] = (i <= t) ? x[i] : 0.0;
  }
}

__global__ void f64_to_f32_sq_kernel(long n, const double *__restrict__ s, float *__restrict__ d, long ldd) {
  for (long c = blockIdx.y; c < n; c += gridDim.y)
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x)
      d[r + c * ldd] = (float)(0.5 * (s[r + c * n] + s[c + r * n]));
}

static dim3 g2(long rows, long cols) {
  return dim3((unsigned)std::min<long>((rows + 255) / 256, 64), (unsigned)std::min<long>(cols, 8192));
}

}  // namespace evd

using namespace evd;

extern "C" int evd_gen_symmetric_f32(evd_ctx *ctx, long n, unsigned long long seed, int kind, float *A, long lda) {
  if (!ctx) return -1;
  EVD_ARG(ctx, n >= 0, 2);
  EVD_ARG(ctx, kind == 0 || kind == 1, 4);
  EVD_ARG(ctx, lda >= std::max<long>(1, n), 6);
  if (n == 0) return 0;
  gen_sym_kernel<<<g2(n, n), 256, 0, ctx->stream>>>(n, seed, kind, A, lda);
  EVD_LAUNCHED(ctx);
  return 0;
}

extern "C" int evd_gen_spectrum_f32(evd_ctx *ctx, long n, unsigned long long seed, const double *lam, float *A,
                                    long lda) {
  if (!ctx) return -1;
  EVD_ARG(ctx, n >= 0, 2);
  EVD_ARG(ctx, lam != nullptr || n == 0, 4);
  EVD_ARG(ctx, lda >= std::max<long>(1, n), 6);
  if (n == 0) return 0;
  if (ctx->arena_top != 0) return set_error(ctx, EVD_E_UNSUPPORTED, "generator cannot nest");
  const int NV = 32;
  EVD_TRY(arena_ensure(ctx, align256((size_t)n * n * 8) + 4 * align256((size_t)n * NV * 8) + 8 * align256(NV * NV * 8)));
  ArenaFrame frame(ctx);
  double *A64 = arena_alloc_n<double>(ctx, (size_t)n * n);
  double *V = arena_alloc_n<double>(ctx, (size_t)n * NV);
  double *X = arena_alloc_n<double>(ctx, (size_t)n * NV);
  double *Zz = arena_alloc_n<double>(ctx, (size_t)n * NV);
  double *G = arena_alloc_n<double>(ctx, NV * NV);
  double *T = arena_alloc_n<double>(ctx, NV * NV);
  double *S = arena_alloc_n<double>(ctx, NV * NV);
  double *M1 = arena_alloc_n<double>(ctx, NV * NV);
  double *M = arena_alloc_n<double>(ctx, NV * NV);
  if (!A64 || !V || !X || !Zz || !G || !T || !S || !M1 || !M) return EVD_E_NOMEM;
  EVD_TRY(fill_f64(ctx, A64, n, n, n, 0.0));
  diag_f64_kernel<<<(unsigned)std::min<long>((n + 255) / 256, 1024), 256, 0, ctx->stream>>>(n, lam, A64);
  EVD_LAUNCHED(ctx);
  for (long c0 = 0; c0 < n; c0 += NV) {
    const int nv = (int)std::min<long>(NV, n - c0);
    gen_normal_block_kernel<<<g2(n, nv), 256, 0, ctx->stream>>>(n, nv, c0, seed, V);
    EVD_LAUNCHED(ctx);
    EVD_TRY(dgemm(ctx, 'T', 'N', nv, nv, n, 1.0, V, n, V, n, 0.0, G, nv));
    wy_t_from_gram_kernel<<<1, 32, 0, ctx->stream>>>(nv, G, T);
    EVD_LAUNCHED(ctx);
    // Q_blk = I - V T V^T;  A <- Q_blk A Q_blk^T = A - V Zz^T - Zz V^T
    EVD_TRY(dgemm(ctx, 'N', 'N', n, nv, n, 1.0, A64, n, V, n, 0.0, X, n));        // X = A V
    EVD_TRY(dgemm(ctx, 'T', 'N', nv, nv, n, 1.0, V, n, X, n, 0.0, S, nv));        // S = V^T A V
    EVD_TRY(dgemm(ctx, 'N', 'N', nv, nv, nv, 1.0, T, nv, S, nv, 0.0, M1, nv));    // T S
    EVD_TRY(dgemm(ctx, 'N', 'T', nv, nv, nv, 1.0, M1, nv, T, nv, 0.0, M, nv));    // M = T S T^T
    EVD_TRY(dgemm(ctx, 'N', 'T', n, nv, nv, 1.0, X, n, T, nv, 0.0, Zz, n));       // X T^T
    EVD_TRY(dgemm(ctx, 'N', 'N', n, nv, nv, -0.5, V, n, M, nv, 1.0, Zz, n));      // - V M / 2
    EVD_TRY(dgemm(ctx, 'N', 'T', n, n, nv, -1.0, V, n, Zz, n, 1.0, A64, n));      // A -= V Zz^T
    EVD_TRY(dgemm(ctx, 'N', 'T', n, n, nv, -1.0, Zz, n, V, n, 1.0, A64, n));      // A -= Zz V^T
  }
  f64_to_f32_sq_kernel<<<g2(n, n), 256, 0, ctx->stream>>>(n, A64, A, lda);
  EVD_LAUNCHED(ctx);
  EVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}
