// Mixed-size batches: problems of different orders solved as a few same-size
// batches.  The reference solves one matrix per process run
// (src/EVD/myEVD.cu:251-706); here each problem of order n is embedded in the
// next bucket order N of a ladder (64, 96, 128, 192, ..., 2^k, 3 * 2^(k-1))
//
//   P = [ A  0 ]      S = diag(s_1 .. s_{N-n}),  s_j > ||A||_inf >= lambda_max(A)
//       [ 0  S ]
//
// so the n smallest eigenvalues of P are those of A and their eigenvectors
// are [v; 0]: w(A) = w(P)[0:n] and Z(A) = Z(P)[0:n, 0:n].  Every bucket (and
// eigenvector flag) is one evd_syevd_batched_f32 solve, i.e. one launch per
// stage for all of its problems.  Orders above the largest bucket are solved
// one at a time, unpadded.  With b = 0 each bucket of order N runs at
// bandwidth min(32, N / 4) (16 and 24 for the two smallest buckets).
#include "common.hpp"

namespace evd {


namespace {

constexpr long kMaxBucket = 4096;

long bucket_order(long n) {
  long p = 64;
  while (true) {
    if (n <= p) return p;
    if (n <= p + p / 2) return p + p / 2;
    p *= 2;
  }
}

// ||A||_inf of the symmetric matrix held in A's lower triangle (column
// major), one workgroup per problem: row i sums |A[i, 0:i]| + |A[i:n, i]|.
__global__ __launch_bounds__(256) void inf_norm_kernel(const long *__restrict__ n_of, float *const *__restrict__ A_of,
                                                       const long *__restrict__ lda_of, double *__restrict__ norm) {
  const int q = blockIdx.x;
  const long n = n_of[q], lda = lda_of[q];
  const float *A = A_of[q];
  __shared__ double red[256];
  double mx = 0.0;
  for (long i = threadIdx.x; i < n; i += 256) {
    double s = 0.0;
    for (long j = 0; j < i; j++) s += fabs((double)A[i + j * lda]);
    for (long r = i; r < n; r++) s += fabs((double)A[r + i * lda]);
    mx = fmax(mx, s);
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) norm[q] = red[0];
}

// P_q (N x N, column major, stride N^2): A_q's lower triangle in the
// top-left block, zeros elsewhere, s_j = (1.5 + j / (N - n)) (||A_q|| + tiny)
// on the padded diagonal.  grid (ceil(N / 256), N, count).
__global__ void pad_kernel(long N, const long *__restrict__ n_of, float *const *__restrict__ A_of,
                           const long *__restrict__ lda_of, const double *__restrict__ norm, float *__restrict__ P) {
  const int q = blockIdx.z;
  const long c = blockIdx.y, n = n_of[q], lda = lda_of[q];
  const float *A = A_of[q];
  float *Pq = P + (long)q * N * N;
  const double base = norm[q] + 1e-30;
  for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < N; r += (long)gridDim.x * blockDim.x) {
    float v = 0.f;
    if (r < n && c < n) {
      if (r >= c) v = A[r + c * lda];
    } else if (r == c) {
      v = (float)(base * (1.5 + (double)(r - n) / (double)(N - n)));
    }
    Pq[r + c * N] = v;
  }
}

// w_q[0:n] = wP_q[0:n]; Z_q = ZP_q[0:n, 0:n] (Z_q may be null).  grid
// (ceil(N / 256), N, count).
__global__ void unpad_kernel(long N, const long *__restrict__ n_of, const double *__restrict__ wP,
                             const float *__restrict__ ZP, double *const *__restrict__ w_of,
                             float *const *__restrict__ Z_of, const long *__restrict__ ldz_of) {
  const int q = blockIdx.z;
  const long c = blockIdx.y, n = n_of[q];
  if (c >= n) return;
  const long r0 = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (c == 0)
    for (long i = r0; i < n; i += (long)gridDim.x * blockDim.x) w_of[q][i] = wP[(long)q * N + i];
  float *Z = Z_of[q];
  if (!Z) return;
  const long ldz = ldz_of[q];
  for (long r = r0; r < n; r += (long)gridDim.x * blockDim.x) Z[r + c * ldz] = ZP[(long)q * N * N + r + c * N];
}

struct Group {
  long N;
  char jobz;
  std::vector<int> members;
};

}  // namespace

}  // namespace evd

using namespace evd;

extern "C" int evd_syevd_mixed_f32(evd_ctx *ctx, int count, const long *n, float *const *A, const long *lda,
                                   double *const *w, float *const *Z, const long *ldz, const char *jobz, long b) {
  if (!ctx) return -1;
  EVD_ARG(ctx, count >= 0, 2);
  EVD_ARG(ctx, count == 0 || (n && A && lda && w && Z && ldz && jobz), 3);
  EVD_ARG(ctx, b == 0 || b == 32 || b == 64, 10);  // 0: b = min(32, N / 4) for each bucket order N
  if (ctx->arena_top != 0) return set_error(ctx, EVD_E_UNSUPPORTED, "evd_syevd_mixed_f32 cannot nest");
  for (int q = 0; q < count; q++) {
    const bool vec = (jobz[q] == 'V' || jobz[q] == 'v');
    if (!(vec || jobz[q] == 'N' || jobz[q] == 'n')) return set_error(ctx, -9, "jobz[%d] invalid", q);
    if (n[q] < 1 || lda[q] < n[q] || !A[q] || !w[q] || (vec && (!Z[q] || ldz[q] < n[q])))
      return set_error(ctx, -3, "problem %d: invalid order, buffer or leading dimension", q);
  }
  // groups: (bucket order, eigenvector flag); large orders alone
  std::vector<Group> groups;
  std::vector<int> alone;
  for (int q = 0; q < count; q++) {
    const char jz = (jobz[q] == 'V' || jobz[q] == 'v') ? 'V' : 'N';
    if (n[q] > kMaxBucket) {
      alone.push_back(q);
      continue;
    }
    const long N = bucket_order(std::max<long>(n[q], b > 0 ? 2 * b : 64));
    auto it = std::find_if(groups.begin(), groups.end(), [&](const Group &g) { return g.N == N && g.jobz == jz; });
    if (it == groups.end()) {
      groups.push_back(Group{N, jz, {}});
      it = groups.end() - 1;
    }
    it->members.push_back(q);
  }
  for (const Group &g : groups) {
    const int c = (int)g.members.size();
    const long N = g.N;
    const bool vec = g.jobz == 'V';
    const long bg = b > 0 ? b : std::min<long>(32, N / 4);  // the bucket's bandwidth
    const long nb = bg * std::max<long>(1, 1024 / bg);
    // staging: per-problem descriptors, norms, padded matrices and results
    const size_t desc = align256((size_t)c * sizeof(long)) * 3 + align256((size_t)c * sizeof(void *)) * 3 +
                        align256((size_t)c * sizeof(double));
    const size_t stage = desc + align256((size_t)c * N * N * sizeof(float)) * (vec ? 2 : 1) +
                         align256((size_t)c * N * sizeof(double));
    EVD_TRY(arena_ensure(ctx, stage + syevd_batched_workspace(c, N, bg, nb, g.jobz, ctx->precision)));
    ArenaFrame frame(ctx);
    long *dn = arena_alloc_n<long>(ctx, c), *dlda = arena_alloc_n<long>(ctx, c), *dldz = arena_alloc_n<long>(ctx, c);
    float **dA = arena_alloc_n<float *>(ctx, c), **dZ = arena_alloc_n<float *>(ctx, c);
    double **dw = arena_alloc_n<double *>(ctx, c);
    double *norm = arena_alloc_n<double>(ctx, c);
    float *P = arena_alloc_n<float>(ctx, (size_t)c * N * N);
    float *ZP = vec ? arena_alloc_n<float>(ctx, (size_t)c * N * N) : nullptr;
    double *wP = arena_alloc_n<double>(ctx, (size_t)c * N);
    if (!dn || !dlda || !dldz || !dA || !dZ || !dw || !norm || !P || (vec && !ZP) || !wP) return EVD_E_NOMEM;
    // descriptors through one pinned-free path: a host staging vector copied
    // with the stream (synchronised below before it is released)
    std::vector<long> hn(c), hlda(c), hldz(c);
    std::vector<float *> hA(c), hZ(c);
    std::vector<double *> hw(c);
    for (int i = 0; i < c; i++) {
      const int q = g.members[i];
      hn[i] = n[q];
      hlda[i] = lda[q];
      hldz[i] = vec ? ldz[q] : 1;
      hA[i] = A[q];
      hZ[i] = vec ? Z[q] : nullptr;
      hw[i] = w[q];
    }
    EVD_HIP(ctx, hipMemcpyAsync(dn, hn.data(), sizeof(long) * c, hipMemcpyHostToDevice, ctx->stream));
    EVD_HIP(ctx, hipMemcpyAsync(dlda, hlda.data(), sizeof(long) * c, hipMemcpyHostToDevice, ctx->stream));
    EVD_HIP(ctx, hipMemcpyAsync(dldz, hldz.data(), sizeof(long) * c, hipMemcpyHostToDevice, ctx->stream));
    EVD_HIP(ctx, hipMemcpyAsync(dA, hA.data(), sizeof(float *) * c, hipMemcpyHostToDevice, ctx->stream));
    EVD_HIP(ctx, hipMemcpyAsync(dZ, hZ.data(), sizeof(float *) * c, hipMemcpyHostToDevice, ctx->stream));
    EVD_HIP(ctx, hipMemcpyAsync(dw, hw.data(), sizeof(double *) * c, hipMemcpyHostToDevice, ctx->stream));
    EVD_HIP(ctx, hipStreamSynchronize(ctx->stream));  // the host vectors go out of scope with this group
    inf_norm_kernel<<<c, 256, 0, ctx->stream>>>(dn, dA, dlda, norm);
    EVD_LAUNCHED(ctx);
    const dim3 grid((unsigned)((N + 255) / 256), (unsigned)N, (unsigned)c);
    pad_kernel<<<grid, 256, 0, ctx->stream>>>(N, dn, dA, dlda, norm, P);
    EVD_LAUNCHED(ctx);
    // the scaling of the input screen follows A's entries, not the pads: the
    // split-FP16 form scales each problem to max |a_ij| in [0.5, 1), and a
    // pad-driven scale would push A's entries toward FP16 subnormals (pads
    // <= 2.5 ||A||_inf <= 2.5 n max |a_ij| stay inside the FP16 range, and
    // the reduction never multiplies them with anything but exact zeros)
    struct AmaxRows {
      evd_ctx *c;
      AmaxRows(evd_ctx *c_, const long *rows) : c(c_) { c->amax_rows = rows; }
      ~AmaxRows() { c->amax_rows = nullptr; }
    } amax_rows(ctx, hn.data());
    EVD_TRY(syevd_batched_call(ctx, g.jobz, c, N, bg, nb, P, N, N * N, wP, N, ZP, N, N * N));
    unpad_kernel<<<grid, 256, 0, ctx->stream>>>(N, dn, wP, ZP, dw, dZ, dldz);
    EVD_LAUNCHED(ctx);
  }
  for (int q : alone) EVD_TRY(evd_syevd_f32(ctx, jobz[q], n[q], b > 0 ? b : 32, 0, A[q], lda[q], w[q], Z[q], ldz[q]));
  return check_device_error(ctx, "evd_syevd_mixed_f32");
}
