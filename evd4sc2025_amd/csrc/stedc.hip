// Symmetric tridiagonal divide and conquer on the GPU, FP64 throughout.
// Replaces the host LAPACKE_dstedc('I') call of the reference driver
// (src/EVD/myEVD.cu:529).  The mathematics is Cuppen's method as published
// (Cuppen 1981; Gu & Eisenstat 1995) in the form LAPACK's dlaed0/1/2/3/4 use:
//
//  * tearing: T = diag(T1 - |e| e_k e_k^T, T2 - |e| e_1 e_1^T) + |e| v v^T,
//    v = [e_k; sign(e) e_1], applied at every split up front;
//  * merge: d = [lambda(T1); lambda(T2)], z = [last row of Q1; sign(e) first
//    row of Q2] / sqrt(2), rho = 2|e|; sort d; deflate small |rho z_j| and
//    close pairs (Givens rotation of the two columns, tol = 8 eps max(|d|,|z|));
//  * secular equation 1 + rho sum z_j^2/(d_j - lambda) = 0 per root, solved
//    relative to the nearer pole with a two-pole rational model step inside a
//    bisection bracket;
//  * Loewner recomputation of z (Gu-Eisenstat) for orthogonal vectors;
//  * Q_new = Q' [U 0; 0 I] by grouped FP64 MFMA GEMMs over the top and bottom
//    row blocks (column types 1/2/3 as in dlaed2/dlaed3).
//
// GPU layout: the binary split tree is built on the host; all merges of one
// tree height run in the same launches (grid.y = merge), every merge working
// in the index range [start, start+N) of n-long scratch arrays and in the
// diagonal block of the n x n Q / scratch matrices.  The deflation scan is
// parallel except for the chain of Givens rotations (processed 64 pairs at a
// time by one wave).
#include <algorithm>
#include <cfloat>

#include <rocblas/rocblas.h>

#include "common.hpp"

namespace evd {

struct MergeDesc {
  int start, n1, n2, pad;
  double rho;  // |e| at the split
  double sgn;  // sign(e)
};
struct MergeOut {
  int K, K1, K2, K3, ndefl, nrot;
  double rho2;  // 2|e|
  double zsum;  // sum of squares of non-deflated z
};

struct DCScratch {
  double *ds, *zs;     // sorted d, z
  int *cs;             // sorted -> local column
  double *nd_d, *nd_z; // non-deflated
  int *nd_col, *qpos;
  double *df_d;
  int *df_col;
  int *rot_p, *rot_n;
  double *rot_c, *rot_s;
  int *types;
  int *orig;
  double *tau, *lam, *zhat;
  int *colmap, *dmap;
  double *Qc, *Uc;  // n x n, ld n
  // Same-size batches run as one concatenated index space: merge starts (and
  // every vector above) are global, problem q = MergeDesc::pad owns indices
  // [q n, q n + n).  Its matrices sit at their own offsets; adding
  // q * zmul (Z) or q * qmul (Qc, Uc) to an index formed from global starts
  // lands in problem q's matrix (0 for a single problem).
  long zmul = 0, qmul = 0;
};

// Tearing of every split at once, one thread per diagonal entry (an entry
// between two splits is adjusted by both): split[k] != 0 marks a split between
// rows k-1 and k.
__global__ void dc_tear_kernel(long n, const int *__restrict__ split, const double *__restrict__ E,
                               double *__restrict__ D) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double d = D[i];
    if (i > 0 && split[i]) d -= fabs(E[i - 1]);
    if (i + 1 < n && split[i + 1]) d -= fabs(E[i]);
    D[i] = d;
  }
}

// Normalisation of each problem's tridiagonal to max(|d_i|, |e_i|) = 1
// before the merges and back afterwards (the scaling step of LAPACK dstedc,
// which the reference's MKL call performs): the deflation tolerance
// 8 eps max(|d|max, |z|max) mixes the matrix scale with the unit-vector
// scale of z, so it is only relative to the matrix on a normalised problem.
// One workgroup per problem (blockIdx.x); nrm[q] keeps the factor.
__global__ __launch_bounds__(256) void dc_normalize_kernel(long n, double *__restrict__ D, double *__restrict__ E,
                                                           double *__restrict__ nrm) {
  __shared__ double red[256];
  const int q = blockIdx.x;
  double *d = D + q * n, *e = E + q * n;
  double mx = 0.0;
  for (long i = threadIdx.x; i < n; i += 256) {
    mx = fmax(mx, fabs(d[i]));
    if (i < n - 1) mx = fmax(mx, fabs(e[i]));
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  const double s = red[0];
  if (threadIdx.x == 0) nrm[q] = s;
  if (s == 0.0 || s == 1.0) return;
  for (long i = threadIdx.x; i < n; i += 256) {
    d[i] /= s;
    if (i < n - 1) e[i] /= s;
  }
}

__global__ void dc_denormalize_kernel(long n, double *__restrict__ D, const double *__restrict__ nrm) {
  const int q = blockIdx.y;
  const double s = nrm[q];
  if (s == 0.0 || s == 1.0) return;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    D[q * n + i] *= s;
}

// |e| and sign(e) at every split, read on the device so no host round trip
// sits between the tridiagonal stage's producer and the merges.
__global__ void dc_rho_kernel(int nmerge, MergeDesc *__restrict__ md, const double *__restrict__ E) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nmerge; i += gridDim.x * blockDim.x) {
    const double e = E[md[i].start + md[i].n1 - 1];
    md[i].rho = fabs(e);
    md[i].sgn = (e >= 0.0) ? 1.0 : -1.0;
  }
}

__global__ void dc_sort_kernel(const MergeDesc *__restrict__ md, const double *__restrict__ D,
                               const double *__restrict__ Z, long ldz, DCScratch sc) {
  const MergeDesc m = md[blockIdx.y];
  const int N = m.n1 + m.n2, st = m.start;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < N; t += gridDim.x * blockDim.x) {
    double d = D[st + t], z;
    int pos;
    if (t < m.n1) {
      z = Z[(long)m.pad * sc.zmul + (st + m.n1 - 1) + (long)(st + t) * ldz];
      int lo = 0, hi = m.n2;
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (D[st + m.n1 + mid] < d)
          lo = mid + 1;
        else
          hi = mid;
      }
      pos = t + lo;
    } else {
      z = m.sgn * Z[(long)m.pad * sc.zmul + (st + m.n1) + (long)(st + t) * ldz];
      int lo = 0, hi = m.n1;
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (D[st + mid] <= d)
          lo = mid + 1;
        else
          hi = mid;
      }
      pos = (t - m.n1) + lo;
    }
    sc.ds[st + pos] = d;
    sc.zs[st + pos] = z * 0.70710678118654752440;
    sc.cs[st + pos] = t;
  }
}

// One thread per merge: deflation scan in sorted order (dlaed2 logic).
// Block-wide helpers for one 256-thread workgroup.
__device__ inline double block_max256(double v, double *red) {
  v = fmax(v, __shfl_xor(v, 32, 64));
  for (int o = 16; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}
// exclusive prefix of per-thread counts (thread order); returns the total
__device__ inline int block_scan256(int v, int *sh, int &excl) {
  __syncthreads();
  sh[threadIdx.x] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int t = 0; t < 256; t++) {
      const int x = sh[t];
      sh[t] = acc;
      acc += x;
    }
    sh[256] = acc;
  }
  __syncthreads();
  excl = sh[threadIdx.x];
  return sh[256];
}

// Deflation of one merge (LAPACK dlaed2), one 256-thread workgroup per merge.
// z-deflation (rho |z_j| <= tol) is decided in parallel and compacted in
// order.  The Givens deflation scan over the remaining entries is sequential
// only through its rotations (pair (p, j) depends on the previous pair only if
// that pair rotated), so wave 0 tests 64 consecutive pairs at once on the
// current values, accepts every pair before the first rotation, applies that
// rotation and continues after it.  The deflated list is the in-order
// z-deflated run merged with the (few) rotated entries; type counts and the
// type-stable positions come from block scans.
__global__ __launch_bounds__(256) void dc_deflate_kernel(const MergeDesc *__restrict__ md, MergeOut *__restrict__ mo,
                                                         DCScratch sc) {
  __shared__ double red[4];
  __shared__ int sh[257];
  __shared__ int s_nc, s_K, s_nrot;
  const MergeDesc m = md[blockIdx.x];
  const int N = m.n1 + m.n2, st = m.start, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double rho = 2.0 * m.rho;
  double *ds = sc.ds + st, *zs = sc.zs + st;
  const int *cs = sc.cs + st;
  int *types = sc.types + st;
  double *nd_d = sc.nd_d + st, *nd_z = sc.nd_z + st;
  int *nd_col = sc.nd_col + st, *qpos = sc.qpos + st;
  double *df_d = sc.df_d + st;
  int *df_col = sc.df_col + st;
  int *rot_p = sc.rot_p + st, *rot_n = sc.rot_n + st;
  double *rot_c = sc.rot_c + st, *rot_s = sc.rot_s + st;
  int *cand = sc.orig + st;     // candidate list (orig is written later by the secular solve)
  double *tmp_d = sc.tau + st;  // merge buffer (tau likewise)
  int *tmp_c = sc.dmap + st;    // (dmap is written after the deflation)
  // chunk of N owned by this thread, in order
  const int per = (N + 255) / 256, i0 = min(N, tid * per), i1 = min(N, i0 + per);

  double dmax = 0.0, zmax = 0.0;
  for (int i = i0; i < i1; i++) {
    types[i] = (i < m.n1) ? 1 : 3;
    dmax = fmax(dmax, fabs(ds[i]));
    zmax = fmax(zmax, fabs(zs[i]));
  }
  dmax = block_max256(dmax, red);
  zmax = block_max256(zmax, red);
  const double eps = DBL_EPSILON * 0.5;
  const double tol = 8.0 * eps * fmax(dmax, zmax);
  const bool all_defl = (rho * zmax <= tol);

  // z-deflated run (in order) and the candidate list, by compaction
  int nz_loc = 0;
  for (int i = i0; i < i1; i++) nz_loc += (all_defl || rho * fabs(zs[i]) <= tol) ? 1 : 0;
  int zoff;
  const int nzd = block_scan256(nz_loc, sh, zoff);
  int coff = (i0 - zoff > 0) ? i0 - zoff : 0;  // candidates before my chunk = i0 - (z-deflated before it)
  coff = i0 - zoff;
  for (int i = i0; i < i1; i++) {
    if (all_defl || rho * fabs(zs[i]) <= tol) {
      df_d[zoff] = ds[i];
      df_col[zoff] = cs[i];
      zoff++;
    } else {
      cand[coff++] = i;
    }
  }
  if (tid == 0) s_nc = N - nzd;
  __syncthreads();
  const int nc = s_nc;

  // Givens scan (wave 0)
  if (wave == 0) {
    int K = 0, nrot = 0;
    int pj = nc > 0 ? cand[0] : -1;
    int k = 1;
    while (k < nc) {
      const int nvalid = min(64, nc - k);
      const bool valid = lane < nvalid;
      const int j = valid ? cand[k + lane] : 0;
      const int prv = (lane == 0) ? pj : (valid ? cand[k + lane - 1] : 0);
      bool rot = false;
      double c = 0.0, sn = 0.0, tau = 0.0;
      if (valid) {
        sn = zs[prv];
        c = zs[j];
        tau = hypot(c, sn);
        const double t = ds[j] - ds[prv];
        c /= tau;
        sn = -sn / tau;
        rot = fabs(t * c * sn) <= tol;
      }
      const unsigned long long mask = __ballot(rot);
      const int first = mask ? __ffsll((long long)mask) - 1 : nvalid;
      if (lane < first) {
        nd_d[K + lane] = ds[prv];
        nd_z[K + lane] = zs[prv];
        nd_col[K + lane] = cs[prv];
      }
      K += first;
      if (first < nvalid) {
        if (lane == first) {
          zs[j] = tau;
          zs[prv] = 0.0;
          const int cp = cs[prv], cn = cs[j];
          if (types[cp] != types[cn]) types[cn] = 2;
          types[cp] = 4;
          rot_p[nrot] = cp;
          rot_n[nrot] = cn;
          rot_c[nrot] = c;
          rot_s[nrot] = sn;
          const double tt = ds[prv] * c * c + ds[j] * sn * sn;
          ds[j] = ds[prv] * sn * sn + ds[j] * c * c;
          ds[prv] = tt;
          df_d[nzd + nrot] = tt;
          df_col[nzd + nrot] = cp;
        }
        nrot++;
        pj = __shfl(j, first, 64);
        k += first + 1;
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
      } else {
        pj = __shfl(j, nvalid - 1, 64);
        k += nvalid;
      }
    }
    if (pj >= 0) {
      if (lane == 0) {
        nd_d[K] = ds[pj];
        nd_z[K] = zs[pj];
        nd_col[K] = cs[pj];
      }
      K++;
    }
    if (lane == 0) {
      s_K = K;
      s_nrot = nrot;
    }
  }
  __syncthreads();
  const int K = s_K, nrot = s_nrot, nd = nzd + nrot;

  // deflated list ascending: the rotated entries (few) sorted in place, then
  // merged with the z-deflated run (already ascending; ties keep the run first)
  if (tid == 0) {
    for (int i = nzd + 1; i < nd; i++) {
      const double v = df_d[i];
      const int cc = df_col[i];
      int j = i - 1;
      while (j >= nzd && df_d[j] > v) {
        df_d[j + 1] = df_d[j];
        df_col[j + 1] = df_col[j];
        j--;
      }
      df_d[j + 1] = v;
      df_col[j + 1] = cc;
    }
  }
  __syncthreads();
  if (nrot > 0) {
    for (int e = tid; e < nd; e += 256) {
      const double v = df_d[e];
      int pos;
      if (e < nzd) {  // rank among rotated entries with value < v
        int lo = nzd, hi = nd;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (df_d[mid] < v) lo = mid + 1; else hi = mid;
        }
        pos = e + (lo - nzd);
      } else {  // rank among z-deflated entries with value <= v
        int lo = 0, hi = nzd;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (df_d[mid] <= v) lo = mid + 1; else hi = mid;
        }
        pos = (e - nzd) + lo;
      }
      tmp_d[pos] = v;
      tmp_c[pos] = df_col[e];
    }
    __syncthreads();
    for (int e = tid; e < nd; e += 256) {
      df_d[e] = tmp_d[e];
      df_col[e] = tmp_c[e];
    }
  }

  // type counts, z sum and type-stable positions of the non-deflated entries
  const int kp = (K + 255) / 256, k0 = min(K, tid * kp), k1 = min(K, k0 + kp);
  int c1 = 0, c2 = 0, c3 = 0;
  double zs2 = 0.0;
  for (int kk = k0; kk < k1; kk++) {
    const int ty = types[nd_col[kk]];
    c1 += (ty == 1);
    c2 += (ty == 2);
    c3 += (ty == 3);
    zs2 += nd_z[kk] * nd_z[kk];
  }
  int e1, e2, e3;
  const int K1 = block_scan256(c1, sh, e1);
  const int K2 = block_scan256(c2, sh, e2);
  const int K3 = block_scan256(c3, sh, e3);
  for (int kk = k0; kk < k1; kk++) {
    const int ty = types[nd_col[kk]];
    qpos[kk] = (ty == 1) ? e1++ : (ty == 2) ? K1 + e2++ : K1 + K2 + e3++;
  }
  // z sum in a fixed order (deterministic): per-thread chunks, then thread order
  __syncthreads();
  __shared__ double zpart[256];
  zpart[tid] = zs2;
  __syncthreads();
  if (tid == 0) {
    double zsum = 0.0;
    for (int t = 0; t < 256; t++) zsum += zpart[t];
    MergeOut o;
    o.K = K;
    o.K1 = K1;
    o.K2 = K2;
    o.K3 = K3;
    o.ndefl = nd;
    o.nrot = nrot;
    o.rho2 = rho;
    o.zsum = zsum;
    mo[blockIdx.x] = o;
  }
}

// Row-parallel: apply the deflation rotations to Q's columns.
__global__ void dc_rotate_kernel(const MergeDesc *__restrict__ md, const MergeOut *__restrict__ mo,
                                 double *__restrict__ Z, long ldz, DCScratch sc) {
  const MergeDesc m = md[blockIdx.y];
  const MergeOut o = mo[blockIdx.y];
  if (o.nrot == 0) return;
  const int N = m.n1 + m.n2, st = m.start;
  double *Zb = Z + ((long)m.pad * sc.zmul + st + (long)st * ldz);
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < N; r += gridDim.x * blockDim.x) {
    for (int t = 0; t < o.nrot; t++) {
      const int p = sc.rot_p[st + t], q = sc.rot_n[st + t];
      const double c = sc.rot_c[st + t], s = sc.rot_s[st + t];
      const double x = Zb[r + (long)p * ldz], y = Zb[r + (long)q * ldz];
      Zb[r + (long)p * ldz] = c * x + s * y;
      Zb[r + (long)q * ldz] = c * y - s * x;
    }
  }
}

// Gather the permuted copy Qc = [non-deflated by type | deflated]: blocks of
// 256 rows x 64 destination columns (grid z over merges).
__global__ void dc_gather_kernel(const MergeDesc *__restrict__ md, const MergeOut *__restrict__ mo,
                                 const double *__restrict__ Z, long ldz, DCScratch sc, long ldq) {
  const MergeDesc m = md[blockIdx.z];
  const MergeOut o = mo[blockIdx.z];
  const int N = m.n1 + m.n2, st = m.start;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c0 = blockIdx.y * 64, c1 = min(N, c0 + 64);
  if (r >= N || c0 >= N) return;
  const double *Zb = Z + ((long)m.pad * sc.zmul + st + (long)st * ldz);
  double *Qc = sc.Qc + ((long)m.pad * sc.qmul + st + (long)st * ldq);
  for (int k = c0; k < c1; k++) {
    if (k < o.K)
      Qc[r + (long)sc.qpos[st + k] * ldq] = Zb[r + (long)sc.nd_col[st + k] * ldz];
    else
      Qc[r + (long)k * ldq] = Zb[r + (long)sc.df_col[st + (k - o.K)] * ldz];
  }
}

// f(tau) pieces relative to pole `org`: psi over poles <= lower, phi over the rest.
// The secular function and its split derivatives at d[org] + tau, summed by
// one wave: lane l takes poles l, l+64, ...; a butterfly allreduce leaves the
// same sums in every lane, so the root iteration stays wave-uniform.
__device__ inline void secular_eval(const double *__restrict__ d, const double *__restrict__ z, int K, int org,
                                    int lower, double tau, double rho, double &f, double &psi, double &dpsi,
                                    double &phi, double &dphi) {
  const double dorg = d[org];
  psi = dpsi = phi = dphi = 0.0;
  const int lane = threadIdx.x & 63;
  for (int k = lane; k < K; k += 64) {
    const double del = (d[k] - dorg) - tau;
    const double zk = z[k];
    const double t = zk / del;
    const double term = zk * t;
    const double dterm = t * t;
    if (k <= lower) {
      psi += term;
      dpsi += dterm;
    } else {
      phi += term;
      dphi += dterm;
    }
  }
  psi = wave_sum(psi);
  dpsi = wave_sum(dpsi);
  phi = wave_sum(phi);
  dphi = wave_sum(dphi);
  psi *= rho;
  dpsi *= rho;
  phi *= rho;
  dphi *= rho;
  f = 1.0 + psi + phi;
}

__global__ void dc_secular_kernel(const MergeDesc *__restrict__ md, const MergeOut *__restrict__ mo, DCScratch sc,
                                  unsigned *err) {
  const MergeDesc m = md[blockIdx.y];
  const MergeOut o = mo[blockIdx.y];
  const int st = m.start, K = o.K;
  const double *d = sc.nd_d + st, *z = sc.nd_z + st;
  const double rho = o.rho2;
  const double eps = DBL_EPSILON * 0.5;
  // one wave per root
  const int waves = blockDim.x >> 6, lane = threadIdx.x & 63;
  for (int i = blockIdx.x * waves + (threadIdx.x >> 6); i < K; i += gridDim.x * waves) {
    int org;
    double lo, hi;
    if (i < K - 1) {
      const double gap = d[i + 1] - d[i];
      const double mid = 0.5 * gap;
      double f, psi, dpsi, phi, dphi;
      secular_eval(d, z, K, i, i, mid, rho, f, psi, dpsi, phi, dphi);
      if (f >= 0.0) {
        org = i;
        lo = 0.0;
        hi = mid;
      } else {
        org = i + 1;
        lo = mid - gap;  // relative to d[i+1]
        hi = 0.0;
      }
    } else {
      org = K - 1;
      lo = 0.0;
      hi = rho * o.zsum;
      if (hi <= 0.0) hi = rho;
    }
    const int lower = i;  // poles d[0..i] lie below the root
    double tau = 0.5 * (lo + hi);
    bool conv = false;
    for (int it = 0; it < 200; it++) {
      double f, psi, dpsi, phi, dphi;
      secular_eval(d, z, K, org, lower, tau, rho, f, psi, dpsi, phi, dphi);
      if (f < 0.0)
        lo = tau;
      else
        hi = tau;
      const double erretm = 10.0 * (fabs(psi) + fabs(phi)) + 4.0;
      if (fabs(f) <= eps * erretm || (hi - lo) <= 4.0 * eps * fmax(fabs(lo), fabs(hi))) {
        conv = true;
        break;
      }
      // two-pole rational model around the current point
      const double dl = (d[i] - d[org]) - tau;  // < 0
      double tn;
      bool ok = false;
      if (i < K - 1) {
        const double du = (d[i + 1] - d[org]) - tau;  // > 0
        const double q = dl * dl * dpsi, s = du * du * dphi;
        const double C = f - dl * dpsi - du * dphi;
        const double a2 = C, a1 = -(C * (dl + du) + q + s), a0 = C * dl * du + q * du + s * dl;
        double eta = 0.0;
        if (fabs(a2) < 1e-300) {
          if (a1 != 0.0) {
            eta = -a0 / a1;
            ok = true;
          }
        } else {
          const double disc = a1 * a1 - 4.0 * a2 * a0;
          if (disc >= 0.0) {
            const double sq = sqrt(disc);
            const double qq = -0.5 * (a1 + (a1 >= 0.0 ? sq : -sq));
            const double r1 = qq / a2;
            const double r2 = (qq != 0.0) ? a0 / qq : r1;
            const bool in1 = (r1 > dl && r1 < du), in2 = (r2 > dl && r2 < du);
            if (in1 && (!in2 || fabs(r1) <= fabs(r2))) {
              eta = r1;
              ok = true;
            } else if (in2) {
              eta = r2;
              ok = true;
            }
          }
        }
        tn = tau + eta;
      } else {
        // single pole below: C + q/(dl - eta) = 0
        const double q = dl * dl * dpsi;
        const double C = f - dl * dpsi;
        if (C > 0.0) {
          tn = tau + dl + q / C;
          ok = true;
        } else {
          tn = tau;
        }
      }
      if (!ok || !(tn > lo && tn < hi)) tn = 0.5 * (lo + hi);
      if (tn == tau || tn <= lo || tn >= hi) {
        // the bracket cannot be split further in FP64: tau is as good as it gets
        conv = true;
        break;
      }
      tau = tn;
    }
    if (lane == 0) {
      if (!conv) atomicOr(err, (unsigned)EVD_E_NOCONVERGE);
      sc.orig[st + i] = org;
      sc.tau[st + i] = tau;
      sc.lam[st + i] = d[org] + tau;
    }
  }
}

// zhat_j^2 = prod_i (lambda_i - d_j) / (rho prod_{k != j} (d_k - d_j))
__global__ void dc_loewner_kernel(const MergeDesc *__restrict__ md, const MergeOut *__restrict__ mo, DCScratch sc) {
  const MergeDesc m = md[blockIdx.y];
  const MergeOut o = mo[blockIdx.y];
  const int st = m.start, K = o.K;
  const double *d = sc.nd_d + st, *z = sc.nd_z + st;
  const int *orig = sc.orig + st;
  const double *tau = sc.tau + st;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < K; j += gridDim.x * blockDim.x) {
    const double dj = d[j];
    double prod = ((d[orig[j]] - dj) + tau[j]) / o.rho2;  // (lambda_j - d_j)/rho
    for (int k = 0; k < K; k++) {
      if (k == j) continue;
      const double lk = (d[orig[k]] - dj) + tau[k];  // lambda_k - d_j
      prod *= lk / (d[k] - dj);
    }
    const double zh = sqrt(fabs(prod));
    sc.zhat[st + j] = (z[j] >= 0.0) ? zh : -zh;
  }
}

// One workgroup per eigenvector column i of one merge: Uc[qpos[j], i].
__global__ __launch_bounds__(256) void dc_vectors_kernel(const MergeDesc *__restrict__ md,
                                                         const MergeOut *__restrict__ mo, DCScratch sc, long ldu) {
  const MergeDesc m = md[blockIdx.y];
  const MergeOut o = mo[blockIdx.y];
  const int st = m.start, K = o.K;
  const int i = blockIdx.x;
  if (i >= K) return;
  const double *d = sc.nd_d + st;
  const double dorg = d[sc.orig[st + i]];
  const double ti = sc.tau[st + i];
  double *Uc = sc.Uc + ((long)m.pad * sc.qmul + st + (long)st * ldu);
  __shared__ double red[4];
  double ss = 0.0;
  for (int j = threadIdx.x; j < K; j += 256) {
    const double v = sc.zhat[st + j] / ((d[j] - dorg) - ti);
    ss += v * v;
  }
  ss = wave_sum(ss);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ss;
  __syncthreads();
  const double inv = 1.0 / sqrt(red[0] + red[1] + red[2] + red[3]);
  for (int j = threadIdx.x; j < K; j += 256) {
    const double v = sc.zhat[st + j] / ((d[j] - dorg) - ti);
    Uc[sc.qpos[st + j] + (long)i * ldu] = v * inv;
  }
}

// Final ascending order of [lambda (K) ; deflated (ndefl)], new D and column maps.
__global__ void dc_order_kernel(const MergeDesc *__restrict__ md, const MergeOut *__restrict__ mo, DCScratch sc,
                                double *__restrict__ D) {
  const MergeDesc m = md[blockIdx.y];
  const MergeOut o = mo[blockIdx.y];
  const int st = m.start, K = o.K, nd = o.ndefl;
  const double *lam = sc.lam + st, *df = sc.df_d + st;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < K + nd; t += gridDim.x * blockDim.x) {
    if (t < K) {
      const double v = lam[t];
      int lo = 0, hi = nd;
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (df[mid] < v)
          lo = mid + 1;
        else
          hi = mid;
      }
      const int pos = t + lo;
      D[st + pos] = v;
      sc.colmap[st + t] = st + pos;
    } else {
      const int u = t - K;
      const double v = df[u];
      int lo = 0, hi = K;
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (lam[mid] <= v)
          lo = mid + 1;
        else
          hi = mid;
      }
      const int pos = u + lo;
      D[st + pos] = v;
      sc.dmap[st + u] = st + pos;
    }
  }
}

__global__ void dc_copy_deflated_kernel(const MergeDesc *__restrict__ md, const MergeOut *__restrict__ mo,
                                        DCScratch sc, long ldq, double *__restrict__ Z, long ldz) {
  const MergeDesc m = md[blockIdx.y];
  const MergeOut o = mo[blockIdx.y];
  const int st = m.start, N = m.n1 + m.n2;
  const long total = (long)N * o.ndefl;
  const double *Qc = sc.Qc + ((long)m.pad * sc.qmul + st + (long)st * ldq);
  const long zsh = (long)m.pad * sc.zmul;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const int r = (int)(e % N);
    const int u = (int)(e / N);
    Z[zsh + (st + r) + (long)sc.dmap[st + u] * ldz] = Qc[r + (long)(o.K + u) * ldq];
  }
}

// Grouped-GEMM problems of one tree level, built on the device from the
// deflation results: merge i gives problem 2i (top rows: Qc[:, 0:K1+K2]) and
// 2i+1 (bottom rows: Qc[:, K1:K]).  Tile offsets are fixed per level from the
// bound K <= N; tiles past a problem's actual K columns exit at once, and a
// merge with K = 0 gets m = 0.
__global__ void dc_problems_kernel(int nm, const MergeDesc *__restrict__ md, const MergeOut *__restrict__ mo,
                                   const int *__restrict__ tile_begin, DCScratch sc, long ldq, double *Z, long ldz,
                                   DGemmProblem *__restrict__ probs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nm) return;
  const MergeDesc m = md[i];
  const MergeOut o = mo[i];
  const int st = m.start;
  const long zsh = (long)m.pad * sc.zmul, qsh = (long)m.pad * sc.qmul;
  DGemmProblem top;
  top.m = o.K > 0 ? m.n1 : 0;
  top.n = o.K;
  top.k = o.K1 + o.K2;
  top.tile_begin = tile_begin[2 * i];
  top.A = sc.Qc + (qsh + st + (long)st * ldq);
  top.lda = ldq;
  top.B = sc.Uc + (qsh + st + (long)st * ldq);
  top.ldb = ldq;
  top.C = Z + (zsh + st);
  top.ldc = ldz;
  top.colmap = sc.colmap + st;
  probs[2 * i] = top;
  DGemmProblem bot;
  bot.m = o.K > 0 ? m.n2 : 0;
  bot.n = o.K;
  bot.k = o.K2 + o.K3;
  bot.tile_begin = tile_begin[2 * i + 1];
  bot.A = sc.Qc + (qsh + (st + m.n1) + (long)(st + o.K1) * ldq);
  bot.lda = ldq;
  bot.B = sc.Uc + (qsh + (st + o.K1) + (long)st * ldq);
  bot.ldb = ldq;
  bot.C = Z + (zsh + st + m.n1);
  bot.ldc = ldz;
  bot.colmap = sc.colmap + st;
  probs[2 * i + 1] = bot;
}

struct TreeNode {
  int start, size, left, right, height;
};

static int build_tree(std::vector<TreeNode> &nodes, int start, int size) {
  int id = (int)nodes.size();
  nodes.push_back({start, size, -1, -1, 0});
  if (size > 1) {
    int n1 = size / 2;
    int l = build_tree(nodes, start, n1);
    int r = build_tree(nodes, start + n1, size - n1);
    nodes[id].left = l;
    nodes[id].right = r;
    nodes[id].height = 1 + std::max(nodes[l].height, nodes[r].height);
  }
  return id;
}

// Merges of order >= kDcLibMinN form their FP64 products with rocBLAS (the
// library's dgemm runs these shapes at 71-75 TFLOP/s on MI355X, the grouped
// tile kernel here at about 48) when the host may read the level's
// deflation counts, i.e. outside graph capture; the product goes through a
// temporary of (n/2 + 1) x n doubles and is scattered to its columns.
constexpr long kDcLibMinN = 4096;
static size_t dc_lib_temp_words(long n) { return n >= kDcLibMinN ? (size_t)(n / 2 + 1) * n : 0; }

static size_t dc_per_problem(long n) {
  size_t b = 0;
  b += align256((size_t)n * n * sizeof(double)) * 2;  // Qc, Uc
  b += align256((size_t)n * sizeof(double)) * 12;
  b += align256((size_t)n * sizeof(int)) * 12;
  b += align256((size_t)n * sizeof(MergeOut));
  b += align256((size_t)2 * n * sizeof(DGemmProblem));
  return b;
}

size_t stedc_workspace(long n) { return dc_per_problem(n) + align256(dc_lib_temp_words(n) * sizeof(double)) + (1 << 16); }

// Merge plan of order n, built once per context: the tree (halving splits
// down to single rows), merges grouped by height, the split flags and the
// per-level tile offsets of the grouped GEMMs, copied to device memory that
// lives as long as the context.  Only the split values (rho, sign) change per
// solve; dc_rho_kernel writes them into the device descriptors.
int dc_plan(evd_ctx *ctx, long n, int count, const DcPlan **out) {
  const long key = n | ((long)count << 32);
  auto it = ctx->dc_plans.find(key);
  if (it != ctx->dc_plans.end()) {
    *out = &it->second;
    return 0;
  }
  if (ctx->capturing)
    return set_error(ctx, EVD_E_UNSUPPORTED, "D&C plan for %d x n = %ld missing during capture", count, n);
  std::vector<TreeNode> nodes;
  nodes.reserve(2 * n);
  build_tree(nodes, 0, (int)n);
  DcPlan plan;
  plan.H = nodes[0].height;
  const int H = plan.H;
  std::vector<std::vector<int>> by_height(H + 1);
  for (int i = 0; i < (int)nodes.size(); i++)
    if (nodes[i].size > 1) by_height[nodes[i].height].push_back(i);
  std::vector<MergeDesc> all;
  std::vector<int> tiles;
  plan.level_off.assign(H + 2, 0);
  plan.level_maxN.assign(H + 1, 0);
  plan.level_tiles.assign(H + 1, 0);
  auto ceil_div = [](int a, int b) { return (a + b - 1) / b; };
  for (int h = 1; h <= H; h++) {
    plan.level_off[h] = (int)all.size();
    int t = 0;
    for (int q = 0; q < count; q++)
    for (int id : by_height[h]) {
      const TreeNode &nd = nodes[id];
      MergeDesc m;
      m.start = (int)(q * n) + nd.start;
      m.n1 = nodes[nd.left].size;
      m.n2 = nodes[nd.right].size;
      m.pad = q;
      m.rho = 0.0;
      m.sgn = 1.0;
      all.push_back(m);
      const int N = m.n1 + m.n2, tn = ceil_div(N, kDGroupTile);
      plan.level_maxN[h] = std::max(plan.level_maxN[h], N);
      tiles.push_back(t);
      t += ceil_div(m.n1, kDGroupTile) * tn;
      tiles.push_back(t);
      t += ceil_div(m.n2, kDGroupTile) * tn;
    }
    plan.level_tiles[h] = t;
  }
  plan.level_off[H + 1] = (int)all.size();
  plan.nmerge = (int)all.size();
  const long nt = n * count;
  std::vector<int> split(nt, 0);
  for (const MergeDesc &m : all) split[m.start + m.n1] = 1;
  const size_t b_md = align256(sizeof(MergeDesc) * all.size()), b_split = align256(sizeof(int) * nt),
               b_tiles = align256(sizeof(int) * tiles.size());
  std::vector<char> &host = plan.host;
  host.assign(b_md + b_split + b_tiles, 0);
  memcpy(host.data(), all.data(), sizeof(MergeDesc) * all.size());
  memcpy(host.data() + b_md, split.data(), sizeof(int) * nt);
  memcpy(host.data() + b_md + b_split, tiles.data(), sizeof(int) * tiles.size());
  // device copy: carved from the context's plan pool (one allocation for all
  // plans), else an allocation of its own; the upload is stream-ordered and
  // its source stays alive with the plan
  EVD_TRY(plan_pool_reserve(ctx, kPlanPoolBytes));
  const size_t need = align256(host.size());
  if (ctx->plan_pool_used + need <= ctx->plan_pool_bytes) {
    plan.dev = ctx->plan_pool + ctx->plan_pool_used;
    ctx->plan_pool_used += need;
  } else {
    EVD_HIP(ctx, hipMalloc(&plan.dev, host.size()));
    plan.own = true;
  }
  const hipError_t e = hipMemcpyAsync(plan.dev, host.data(), host.size(), hipMemcpyHostToDevice, ctx->stream);
  if (e != hipSuccess) {
    if (plan.own) hipFree(plan.dev);
    return set_error(ctx, EVD_E_HIP, "D&C plan upload failed: %s", hipGetErrorString(e));
  }
  auto ins = ctx->dc_plans.emplace(key, std::move(plan));
  *out = &ins.first->second;
  return 0;
}

int stedc_batched(evd_ctx *ctx, int count, long n, double *D, double *E, double *Z, long ldz, long zs);

int stedc(evd_ctx *ctx, long n, double *D, double *E, double *Z, long ldz) {
  return stedc_batched(ctx, 1, n, D, E, Z, ldz, 0);
}

size_t stedc_batched_workspace(int count, long n) {
  const int per = (int)std::max<long>(1, 65535 / ((n + 1) / 2));
  return dc_per_problem(n) * (size_t)std::min(count, per) + align256(dc_lib_temp_words(n) * sizeof(double)) +
         (1 << 16);
}

// C[:, colmap[j]] = T[:, j] (j < n, rows < m) of one merge product
__global__ void dc_scatter_cols_kernel(int m, int n, const double *__restrict__ T, double *__restrict__ C, long ldc,
                                       const int *__restrict__ colmap) {
  const long total = (long)m * n;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const int r = (int)(e % m), j = (int)(e / m);
    const int oc = colmap ? colmap[j] : j;
    C[r + (long)oc * ldc] = T[e];
  }
}

// One merge level's products on rocBLAS: the problem list the device built
// (sizes from the deflation results) is read back, each product goes to the
// temporary T and is scattered to its output columns.  Per problem, so the
// problems of a batch get the arithmetic of their single solves.
static int dc_level_library(evd_ctx *ctx, const DGemmProblem *dprob, int nprob, double *T, size_t t_words) {
  std::vector<DGemmProblem> hp((size_t)nprob);
  EVD_HIP(ctx, hipMemcpyAsync(hp.data(), dprob, sizeof(DGemmProblem) * nprob, hipMemcpyDeviceToHost, ctx->stream));
  EVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const double one = 1.0, zero = 0.0;
  if (rocblas_set_stream(ctx->blas, ctx->stream) != rocblas_status_success)
    return set_error(ctx, EVD_E_HIP, "rocblas_set_stream failed");
  for (const DGemmProblem &p : hp) {
    if (p.m <= 0 || p.n <= 0) continue;
    if ((size_t)p.m * p.n > t_words)
      return set_error(ctx, EVD_E_UNSUPPORTED, "D&C product %d x %d exceeds its temporary", p.m, p.n);
    if (p.k > 0) {
      const rocblas_status st = rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_none, p.m, p.n,
                                              p.k, &one, p.A, (rocblas_int)p.lda, p.B, (rocblas_int)p.ldb, &zero, T,
                                              p.m);
      if (st != rocblas_status_success)
        return set_error(ctx, EVD_E_HIP, "rocblas_dgemm(%d x %d x %d) failed: status %d", p.m, p.n, p.k, (int)st);
    } else {
      EVD_HIP(ctx, hipMemsetAsync(T, 0, sizeof(double) * p.m * p.n, ctx->stream));
    }
    const long total = (long)p.m * p.n;
    dc_scatter_cols_kernel<<<(unsigned)std::min<long>((total + 255) / 256, 8192), 256, 0, ctx->stream>>>(
        p.m, p.n, T, p.C, p.ldc, p.colmap);
    EVD_LAUNCHED(ctx);
  }
  return 0;
}

// `count` same-size tridiagonals (D, E of problem q at q n, Z at q zs): every
// merge level of all problems is one set of launches.  Each problem's merges
// and their arithmetic are those of its single solve.
int stedc_batched(evd_ctx *ctx, int count, long n, double *D, double *E, double *Z, long ldz, long zs) {
  if (n <= 0 || count <= 0) return 0;
  if (n == 1) {
    for (int q = 0; q < count; q++) EVD_TRY(set_identity_f64(ctx, Z + q * zs, 1, ldz));
    return 0;
  }
  // the merges of one level are a launch dimension (at most 65535)
  const int per = (int)std::max<long>(1, 65535 / ((n + 1) / 2));
  if (count > per) {
    for (int q0 = 0; q0 < count; q0 += per) {
      const int c = std::min(per, count - q0);
      EVD_TRY(stedc_batched(ctx, c, n, D + q0 * n, E + q0 * n, Z + q0 * zs, ldz, zs));
    }
    return 0;
  }
  ArenaFrame frame(ctx);
  DCScratch sc;
  const long ldq = n, nt = n * count;
  sc.Qc = arena_alloc_n<double>(ctx, (size_t)n * n * count);
  sc.Uc = arena_alloc_n<double>(ctx, (size_t)n * n * count);
  double **dbufs[] = {&sc.ds, &sc.zs, &sc.nd_d, &sc.nd_z, &sc.df_d, &sc.rot_c, &sc.rot_s, &sc.tau, &sc.lam, &sc.zhat};
  for (auto p : dbufs) *p = arena_alloc_n<double>(ctx, (size_t)nt);
  int **ibufs[] = {&sc.cs, &sc.nd_col, &sc.qpos, &sc.df_col, &sc.rot_p, &sc.rot_n, &sc.types, &sc.orig, &sc.colmap, &sc.dmap};
  for (auto p : ibufs) *p = arena_alloc_n<int>(ctx, (size_t)nt);
  if (!sc.Qc || !sc.Uc || !sc.zhat || !sc.dmap) return EVD_E_NOMEM;
  sc.zmul = zs - n * (1 + ldz);
  sc.qmul = -n;

  const DcPlan *plan = nullptr;
  EVD_TRY(dc_plan(ctx, n, count, &plan));
  const int nmerge = plan->nmerge, H = plan->H;
  char *pdev = static_cast<char *>(plan->dev);
  MergeDesc *dmd = reinterpret_cast<MergeDesc *>(pdev);
  const int *dsplit = reinterpret_cast<const int *>(pdev + align256(sizeof(MergeDesc) * nmerge));
  const int *dtiles = reinterpret_cast<const int *>(pdev + align256(sizeof(MergeDesc) * nmerge) +
                                                    align256(sizeof(int) * nt));
  MergeOut *dmo = static_cast<MergeOut *>(arena_alloc(ctx, sizeof(MergeOut) * nmerge));
  // temporary of the library products (large merges, eager solves)
  double *libT = nullptr;
  if (ctx->blas && !ctx->tune.own_gemm && !ctx->capturing && dc_lib_temp_words(n) > 0) {
    libT = arena_alloc_n<double>(ctx, dc_lib_temp_words(n));
    if (!libT) return EVD_E_NOMEM;
  }
  DGemmProblem *dprob = static_cast<DGemmProblem *>(arena_alloc(ctx, sizeof(DGemmProblem) * 2 * nmerge));
  double *nrm = arena_alloc_n<double>(ctx, (size_t)count);
  if (!dmo || !dprob || !nrm) return EVD_E_NOMEM;
  unsigned *err = ctx->ctrl + kCtrlError;

  // split values and signs, tearing of every split, then Q = I
  dc_normalize_kernel<<<(unsigned)count, 256, 0, ctx->stream>>>(n, D, E, nrm);
  EVD_LAUNCHED(ctx);
  dc_rho_kernel<<<(unsigned)std::min((nmerge + 255) / 256, 1024), 256, 0, ctx->stream>>>(nmerge, dmd, E);
  EVD_LAUNCHED(ctx);
  dc_tear_kernel<<<(unsigned)std::min<long>((nt + 255) / 256, 1024), 256, 0, ctx->stream>>>(nt, dsplit, E, D);
  EVD_LAUNCHED(ctx);
  for (int q = 0; q < count; q++) EVD_TRY(set_identity_f64(ctx, Z + q * zs, n, ldz));

  // Merges level by level (by height).  Launch shapes come from the largest
  // merge of the level; every kernel bounds its work by the deflation
  // results it reads on the device, so no level waits for the host.
  for (int h = 1; h <= H; h++) {
    const int off = plan->level_off[h], nm = plan->level_off[h + 1] - off;
    if (nm == 0) continue;
    const MergeDesc *lmd = dmd + off;
    MergeOut *lmo = dmo + off;
    const int maxN = plan->level_maxN[h];
    const unsigned gx = (unsigned)std::min(64, (maxN + 255) / 256);
    dc_sort_kernel<<<dim3(gx, nm), 256, 0, ctx->stream>>>(lmd, D, Z, ldz, sc);
    EVD_LAUNCHED(ctx);
    dc_deflate_kernel<<<nm, 256, 0, ctx->stream>>>(lmd, lmo, sc);
    EVD_LAUNCHED(ctx);
    dc_rotate_kernel<<<dim3(gx, nm), 256, 0, ctx->stream>>>(lmd, lmo, Z, ldz, sc);
    EVD_LAUNCHED(ctx);
    dc_gather_kernel<<<dim3((unsigned)((maxN + 255) / 256), (unsigned)((maxN + 63) / 64), nm), 256, 0, ctx->stream>>>(
        lmd, lmo, Z, ldz, sc, ldq);
    EVD_LAUNCHED(ctx);
    const unsigned gk = (unsigned)std::min(256, (maxN + 255) / 256);
    const unsigned gw = (unsigned)std::min(16384, (maxN + 3) / 4);  // 4 roots (waves) per workgroup
    dc_secular_kernel<<<dim3(gw, nm), 256, 0, ctx->stream>>>(lmd, lmo, sc, err);
    EVD_LAUNCHED(ctx);
    dc_loewner_kernel<<<dim3(gk, nm), 256, 0, ctx->stream>>>(lmd, lmo, sc);
    EVD_LAUNCHED(ctx);
    dc_vectors_kernel<<<dim3(maxN, nm), 256, 0, ctx->stream>>>(lmd, lmo, sc, ldq);
    EVD_LAUNCHED(ctx);
    dc_order_kernel<<<dim3(gx, nm), 256, 0, ctx->stream>>>(lmd, lmo, sc, D);
    EVD_LAUNCHED(ctx);
    dc_problems_kernel<<<(unsigned)((nm + 255) / 256), 256, 0, ctx->stream>>>(nm, lmd, lmo, dtiles + 2 * off, sc, ldq,
                                                                             Z, ldz, dprob);
    EVD_LAUNCHED(ctx);
    if (libT && maxN >= kDcLibMinN && !ctx->capturing)
      EVD_TRY(dc_level_library(ctx, dprob, 2 * nm, libT, dc_lib_temp_words(n)));
    else
      EVD_TRY(dgemm_grouped(ctx, dprob, 2 * nm, plan->level_tiles[h]));
    const unsigned gd = (unsigned)std::min<long>(((long)maxN * maxN + 255) / 256, 1024);
    dc_copy_deflated_kernel<<<dim3(gd, nm), 256, 0, ctx->stream>>>(lmd, lmo, sc, ldq, Z, ldz);
    EVD_LAUNCHED(ctx);
  }
  dc_denormalize_kernel<<<dim3((unsigned)std::min<long>((n + 255) / 256, 64), (unsigned)count), 256, 0,
                          ctx->stream>>>(n, D, nrm);
  EVD_LAUNCHED(ctx);
  return 0;
}

}  // namespace evd

using namespace evd;

extern "C" int evd_stedc_f64(evd_ctx *ctx, long n, double *D, double *E, double *Z, long ldz) {
  if (!ctx) return -1;
  EVD_ARG(ctx, n >= 0, 2);
  EVD_ARG(ctx, D != nullptr || n == 0, 3);
  EVD_ARG(ctx, E != nullptr || n <= 1, 4);
  EVD_ARG(ctx, Z != nullptr || n == 0, 5);
  EVD_ARG(ctx, ldz >= std::max<long>(1, n), 6);
  if (ctx->arena_top == 0) EVD_TRY(arena_ensure(ctx, stedc_workspace(n)));
  EVD_TRY(stedc(ctx, n, D, E, Z, ldz));
  return check_device_error(ctx, "evd_stedc_f64");
}
