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
// MI355X design: sweeps are handed out by an atomic ticket in increasing
// order, so a sweep's predecessor is always owned by a running workgroup and
// no co-residency (cooperative launch) is needed.  The band lives in HBM and
// every access of the chase is device-coherent (sc1 loads and stores, which
// bypass the non-coherent per-XCD caches; on this chip the L1-bypassing sc0
// loads do not observe other CUs' stores, so an XCD-local exchange is not
// available), and a vmcnt(0) wait orders a step's stores before its flag.
// Per step the S and N windows are held in registers (a row per 256/BW
// lanes), B1 and the next B1 in LDS; H B1 is applied by one wave while the
// others wait for their window loads.  Step (s, p) needs step (s-1, p-1+2b)
// only for one element, B1[0][0] of that step (row p+2b-1, column p+b-1),
// which becomes beta; the owner publishes that beta at the end of its
// previous step, so sweep s may start row p as soon as sweep s-1 has
// published a next-row >= p + 2b - 1 (a lag of two steps, not three).
// Every wait is bounded by a watchdog so the grid always drains.
#include <climits>
#include <map>
#include <mutex>

#include "common.hpp"

namespace evd {

// one progress word per 128-byte line: the word a sweep publishes is polled by
// its successor only, and words sharing a line would serialise unrelated
// sweeps' polls and stores on that line
constexpr long kComStride = 32;
constexpr unsigned kBufferWord3 = 0x00020000u;  // raw buffer descriptor: 32-bit data format

constexpr int kSc1 = 16;  // buffer-op cache policy: sc1, device-coherent

__device__ inline double ld_band(__amdgpu_buffer_rsrc_t r, unsigned off_elems) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off_elems * 8u, 0, kSc1));
}
__device__ inline void st_band(__amdgpu_buffer_rsrc_t r, unsigned off_elems, double v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((__vector_size__(2 * sizeof(unsigned)))) unsigned, v),
                                        r, off_elems * 8u, 0, kSc1);
}

// LAPACK-style reflector of the chase (reference sign convention): from the
// column x = Bc[0..lenx) (LDS, stride LD), u = (x/||x|| + sign(x0) e1) /
// sqrt(|x0/||x|| + sign(x0)|), ||u||^2 = 2, H x = beta e1.  Wave 0 only.
template <int BW>
__device__ inline void chase_reflector(const double *Bc, int lenx, int lane, double *u, double *s_beta, float *U,
                                       long urow, long ucol) {
  constexpr int LD = BW + 1;
  const double x = (lane < lenx) ? Bc[lane * LD] : 0.0;
  const double nrm = sqrt(wave_sum(x * x));
  const double x0 = Bc[0];
  double uu = 0.0;
  if (nrm > 0.0) {
    const double inv = 1.0 / nrm;
    const double u0 = x0 * inv + (x0 >= 0.0 ? 1.0 : -1.0);
    const double scl = 1.0 / sqrt(fabs(u0));
    uu = (lane == 0 ? u0 : x * inv) * scl;
  }
  if (lane < BW) u[lane] = (lane < lenx) ? uu : 0.0;
  if (lane < lenx) U[(urow + lane) + ucol] = (float)uu;
  if (lane == 0) *s_beta = (nrm > 0.0) ? (x0 >= 0.0 ? -nrm : nrm) : x0;
}

// A batch of `count` same-size problems shares one launch: ticket t is sweep
// t / count of problem t % count, so each problem's sweeps are still handed
// out in increasing order.  Problem q's band, reflectors and progress words
// start at band + q * band_stride, U + q * u_stride, com + q * com_stride.
template <int BW>
__global__ __launch_bounds__(256, BW == 32 ? 4 : 1) void chase_kernel(int n, int b, double *__restrict__ band_all,
                                                    long ldb, unsigned band_bytes, float *__restrict__ U_all,
                                                    long ldu, int *__restrict__ com_all,
                                                    unsigned *__restrict__ ticket, unsigned *__restrict__ err,
                                                    int count, long band_stride, long u_stride, long com_stride) {
  constexpr int LD = BW + 1;
  constexpr int LPR = 256 / BW;  // lanes per window row
  constexpr int EPT = BW / LPR;  // window columns per lane
  constexpr int HALVES = 64 / BW;
  __shared__ double sB[2][BW * LD];
  __shared__ double u[BW], yv[BW], qv[BW];
  __shared__ double s_beta;
  __shared__ int s_ticket, s_ok;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ri = tid / LPR, kc = tid % LPR;
  const int nsweeps = n - 2;
  const unsigned ld = (unsigned)ldb;  // band offsets fit 32 bits (checked by the launcher)

  // tid 0 polls the predecessor's progress word; false = watchdog expired
  auto wait_prev = [&](const int *com, int s, int need) -> bool {
    if (tid == 0) {
      s_ok = 1;
      if (s > 0) {
        const unsigned long long t0 = clock100();
        while (ld_dev(&com[(long)(s - 1) * kComStride]) < need) {
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
    return s_ok != 0;
  };

  while (true) {
    if (tid == 0)
      s_ticket = (int)__hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int tk = s_ticket;
    __syncthreads();
    if (tk >= nsweeps * count) return;
    const int q = tk % count, s = tk / count;
    const __amdgpu_buffer_rsrc_t rb =
        __builtin_amdgcn_make_buffer_rsrc(band_all + (long)q * band_stride, 0, (int)band_bytes, kBufferWord3);
    float *U = U_all + (long)q * u_stride;
    int *com = com_all + (long)q * com_stride;
    int p = s + 1;
    int len = min(b, n - p);
    int colB1 = 1, cB1 = s;
    bool first = true;
    double *B1 = sB[0], *Bn = sB[1];
    bool ok = wait_prev(com, s, p + 2 * b - 1);
    if (ok) {
      // column s below the diagonal is final: it becomes B1[:, 0]
      for (int i = tid; i < BW; i += 256) B1[i * LD] = (i < len) ? ld_band(rb, (unsigned)(1 + i) + (unsigned)s * ld) : 0.0;
      __syncthreads();
      if (wave == 0) chase_reflector<BW>(B1, len, lane, u, &s_beta, U, p, (long)s * ldu);
      __syncthreads();
    }
    while (ok) {
      const int len2 = max(0, min(b, n - p - len));
      const bool last = (len2 <= 1);
      // ---- S (symmetric, from the lower band) and N windows into registers
      double Sr[EPT], Nr[EPT];
#pragma unroll
      for (int r = 0; r < EPT; r++) {
        const int kk = kc + LPR * r;
        const int lo = min(ri, kk), dd = ri > kk ? ri - kk : kk - ri;
        Sr[r] = (ri < len && kk < len) ? ld_band(rb, (unsigned)dd + (unsigned)(p + lo) * ld) : 0.0;
        Nr[r] = (ri < len2 && kk < len) ? ld_band(rb, (unsigned)(len + ri - kk) + (unsigned)(p + kk) * ld) : 0.0;
      }
      // ---- B1 <- H B1 and write-back (wave 3, while the loads are in flight):
      // t_c = u^T B1[:, c] by lane c (rows split over HALVES lane sets), then
      // the write-back with lanes along rows, so each store instruction
      // covers whole runs of one or two band columns instead of one element
      // of 32 or 64 columns
      if (wave == 3) {
        const int c = lane % BW, h = lane / BW;
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < BW; i += HALVES) t += u[i + h] * B1[(i + h) * LD + c];
        if (HALVES == 2) t += __shfl_xor(t, 32, 64);
        const double beta = s_beta;
        const int ii = lane % BW;
        const double uii = u[ii];
#pragma unroll
        for (int c0 = 0; c0 < BW; c0 += HALVES) {
          const int cc = c0 + h;
          const double tc = __shfl(t, cc, 64);
          if (cc < colB1 && ii < len && !(cc == 0 && ii == 0 && !first)) {
            const int col = cB1 + cc;
            const double v = (cc == 0) ? (ii == 0 ? beta : 0.0) : B1[ii * LD + cc] - uii * tc;
            st_band(rb, (unsigned)(p + ii - col) + (unsigned)col * ld, v);
          }
        }
      }
      // ---- y = S u and q = N u (row sums over the LPR lanes of a row)
      double yp = 0.0, qp = 0.0;
#pragma unroll
      for (int r = 0; r < EPT; r++) {
        const double uk = u[kc + LPR * r];
        yp += Sr[r] * uk;
        qp += Nr[r] * uk;
      }
#pragma unroll
      for (int o = LPR / 2; o > 0; o >>= 1) {
        yp += __shfl_xor(yp, o, 64);
        qp += __shfl_xor(qp, o, 64);
      }
      if (kc == 0) {
        yv[ri] = yp;
        qv[ri] = qp;
      }
      __syncthreads();
      // ---- S <- H S H (rank-2 with w = y - (u^T y / 2) u), N <- N H
      const double psi = wave_sum((lane < BW) ? u[lane] * yv[lane] : 0.0);
      const double ui = u[ri], wi = yv[ri] - 0.5 * psi * ui, qi = qv[ri];
#pragma unroll
      for (int r = 0; r < EPT; r++) {
        const int kk = kc + LPR * r;
        const double uk = u[kk], wk = yv[kk] - 0.5 * psi * uk;
        Sr[r] -= ui * wk + wi * uk;
        Nr[r] -= qi * uk;
        if (ri >= kk && ri < len && kk < len) st_band(rb, (unsigned)(ri - kk) + (unsigned)(p + kk) * ld, Sr[r]);
        Bn[ri * LD + kk] = Nr[r];
        if (last && ri < len2 && kk < len) st_band(rb, (unsigned)(len + ri - kk) + (unsigned)(p + kk) * ld, Nr[r]);
      }
      __syncthreads();
      // ---- next reflector from N[:, 0]; its beta (row p+len, column p) is
      // published now, before the progress word, for the successor sweep
      if (!last && wave == 0) {
        chase_reflector<BW>(Bn, len2, lane, u, &s_beta, U, p + len, (long)s * ldu);
        if (lane == 0) st_band(rb, (unsigned)len + (unsigned)p * ld, s_beta);
      }
      wait_mem();
      __syncthreads();
      if (tid == 0) st_dev(&com[(long)s * kComStride], last ? INT_MAX : p + len);
      if (last) break;
      double *t = B1;
      B1 = Bn;
      Bn = t;
      colB1 = len;
      cB1 = p;
      p += len;
      len = len2;
      first = false;
      ok = wait_prev(com, s, p + 2 * b - 1);
    }
    if (!ok) {
      if (tid == 0) st_dev(&com[(long)s * kComStride], INT_MAX);
      return;
    }
  }
}

// ---------------------------------------------------------------------------
// Grouped chase (b <= 32): a workgroup owns a group of D consecutive sweeps,
// s0 .. s0 + D - 1, one wave per sweep.  The waves step in lockstep: at time
// step t, sweep s0 + i does its step t - 2i (the two-step lag of the sweep
// order), all on a window of band columns held in LDS (column c in slot
// c mod kGW, offsets 0 .. 2b-1), so the hand-off between the group's sweeps
// never leaves the CU.  One more wave moves the window (I/O wave):
//  * it brings in the columns the front sweep s0 first needs at
//    step j (columns s0 .. p_0 + len_0 at j = 0, p_j .. p_j + len_j later, up
//    to n at the front's last step) from HBM: the loads are issued at time
//    j - 3, once the previous group's progress word is >= p_j + 2b - 1, into
//    one of two register buffers, and written to LDS at time j - 1.  The
//    last column of a chunk, p_j + b - 1, is a step row of the previous
//    group's last sweep, which writes it back down to its early-published
//    beta only (offset b): the fill below is zeroed by that sweep's next
//    step, so the I/O wave writes those zeros;
//  * it writes back, at time t, the columns the group's last sweep finished
//    at t - 1 (all columns up to its step row P; column P down to offset b)
//    and publishes that sweep's next row P + len at t + 1, once those stores
//    have completed; the last write-back covers every remaining column and
//    publishes INT_MAX.
// Every step issues the same number of memory operations in the same order
// (kGChunk stores, one poll of the predecessor's progress word, kGChunk
// loads; unused lanes address past the end of the band, which the buffer
// range check drops), so the wait before the publish is a fixed count - the
// newest kGChunk loads stay in flight - and the loads have two steps to land.
// Within a time step the waves' LDS accesses are disjoint and every slot is
// rewritten only after its previous column was written back, provided
// kGW >= (2D + 2) b - D + 1; a CPU model of this schedule (random
// interleavings of groups, access-set and slot checks) reproduces the
// sequential chase.  Groups are handed out by ticket in increasing order, so
// a group's predecessor is always running; every wait is bounded.
constexpr int kGW = 256;  // window columns
// doubles per window column (2b <= 64 offsets).  Element (row R, column C)
// sits at C * kGLD + R - C, so a band row runs through the window with stride
// kGLD - 1 = 63 doubles (odd: the 32 lanes of a row access hit 32 distinct
// banks) and a band column with stride 1.
constexpr int kGLD = 64;
// most columns the front brings in at one step: b + 2 at b = 32 (the first
// chunk s0 .. n of a sweep whose first step is also its last)
constexpr int kGChunk = 34;
constexpr unsigned kOutOfRange = 0x80000000u;  // byte offset past every band (< 2 GiB)

__device__ __attribute__((always_inline)) inline void grp_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}
__device__ __attribute__((always_inline)) inline void wave_lds_order() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// s_waitcnt with only the vector-memory count: vmcnt(N), N <= 63
template <int N>
__device__ __attribute__((always_inline)) inline void wait_vm() {
  static_assert(N >= 0 && N <= 63, "vmcnt is 6 bits");
  __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

__device__ inline int sweep_nsteps(int n, int b, int s) {
  int p = s + 1, len = min(b, n - p), k = 1;
  while (true) {
    const int len2 = max(0, min(b, n - p - len));
    if (len2 <= 1) return k;
    p += len;
    len = len2;
    k++;
  }
}

// 64-bit DPP move of v (rows outside RM read 0)
template <int CTRL, int RM>
__device__ __attribute__((always_inline)) inline double dpp_f64(double v) {
  const unsigned long long x = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)x, CTRL, RM, 0xf, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(x >> 32), CTRL, RM, 0xf, false);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// sum over the 64 lanes, returned in every lane: quad, half-row and row
// reductions by DPP, then row 15 -> rows 1 and 3 and lane 31 -> rows 2 and 3
// broadcasts; lane 63 holds the total
__device__ __attribute__((always_inline)) inline double wave_total(double v) {
  v += dpp_f64<0xb1, 0xf>(v);   // quad_perm [1, 0, 3, 2]
  v += dpp_f64<0x4e, 0xf>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp_f64<0x141, 0xf>(v);  // row_half_mirror
  v += dpp_f64<0x140, 0xf>(v);  // row_mirror
  v += dpp_f64<0x142, 0xa>(v);  // row_bcast:15
  v += dpp_f64<0x143, 0xc>(v);  // row_bcast:31
  const unsigned long long x = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)x, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(x >> 32), 63);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// v of lane ^ 32.  v_permlane32_swap exchanges one half of its destination
// with the other half of its source; with both operands v, one result holds
// this lane's own value and the other the partner's, so xor-ing out the own
// value leaves the partner's whichever half was exchanged.
__device__ __attribute__((always_inline)) inline double half_swap(double v) {
  const unsigned long long x = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)x, hi = (unsigned)(x >> 32);
  const auto rl = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
  const auto rh = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
  const unsigned plo = rl[0] ^ rl[1] ^ lo, phi = rh[0] ^ rh[1] ^ hi;
  return __builtin_bit_cast(double, ((unsigned long long)phi << 32) | plo);
}

// reflector of the chase's convention from x (lane < lenx, zero elsewhere),
// wave-wide; uu is zero for lanes >= lenx
__device__ __attribute__((always_inline)) inline void grp_reflector(double x, int lane, int lenx, double &uu,
                                                                   double &beta) {
  const double nrm = sqrt(wave_total(x * x));
  const unsigned long long xb = __builtin_bit_cast(unsigned long long, x);
  const double x0 = __builtin_bit_cast(double, ((unsigned long long)(unsigned)__builtin_amdgcn_readlane(
                                                    (int)(unsigned)(xb >> 32), 0)
                                                << 32) |
                                                   (unsigned)__builtin_amdgcn_readlane((int)(unsigned)xb, 0));
  uu = 0.0;
  if (nrm > 0.0) {
    const double inv = 1.0 / nrm;
    const double u0 = x0 * inv + (x0 >= 0.0 ? 1.0 : -1.0);
    const double scl = 1.0 / sqrt(fabs(u0));
    uu = (lane == 0 ? u0 : x * inv) * scl;
  }
  if (lane >= lenx) uu = 0.0;
  beta = (nrm > 0.0) ? (x0 >= 0.0 ? -nrm : nrm) : x0;
}

// one chunk of the front's columns in flight: registers + where they go
// (columns lo .. hi-1; hi == lo when the step loads nothing)
struct GrpChunk {
  double v[kGChunk];
  int lo, hi, syn;
};

template <int D>
__global__ __launch_bounds__((D + 1) * 64, 1) void chase_group_kernel(
    int n, int b, double *__restrict__ band_all, long ldb, unsigned band_bytes, float *__restrict__ U_all, long ldu,
    int *__restrict__ com_all, unsigned *__restrict__ ticket, unsigned *__restrict__ err, int count, long band_stride,
    long u_stride, long com_stride) {
  // the window, then one column-sized slot that takes the chunk writes of
  // unused registers (every loaded register is written somewhere, so no load
  // is dropped on some paths and the memory-operation count per step stays
  // fixed)
  __shared__ double win[kGW * kGLD + 64];
  __shared__ double uS[D][32], yS[D][32];
  __shared__ int s_ticket, s_abort;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nsweeps = n - 2;
  const int ngroups = (nsweeps + D - 1) / D;
  const unsigned ld = (unsigned)ldb;
  const int b2 = 2 * b;

  while (true) {
    if (tid == 0) {
      s_ticket = (int)__hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s_abort = 0;
    }
    grp_barrier();
    const int tk = s_ticket;
    grp_barrier();
    if (tk >= ngroups * count) return;
    const int q = tk % count, g = tk / count;
    const int s0 = g * D, nsw = min(D, nsweeps - s0);
    const __amdgpu_buffer_rsrc_t rb =
        __builtin_amdgcn_make_buffer_rsrc(band_all + (long)q * band_stride, 0, (int)band_bytes, kBufferWord3);
    float *U = U_all + (long)q * u_stride;
    int *com = com_all + (long)q * com_stride;
    const int tail_steps = sweep_nsteps(n, b, s0 + nsw - 1);
    const int front_steps = sweep_nsteps(n, b, s0);
    const int te = tail_steps - 1 + 2 * (nsw - 1);  // time step of the last sweep's last step
    // the time loop runs t = -3 .. t_last, an even number of steps (the I/O
    // wave's loop body is two steps, one per register buffer, with no exit
    // between them: an exit there costs the loads their overlap)
    const int t_last = te + 2 + ((te + 6) & 1);

    // Each role runs its own copy of the time loop (one barrier per time
    // step in every copy), so the roles' registers do not add up.
    if (wave < D) {
      // ---- worker: sweep ws of this group.  Lane (c, h) = (lane & 31,
      // lane >> 5) holds rows 2r + h of B1 column c; lane (ri, kc) = the same
      // pair holds row ri, columns kc + 2r of the S and N blocks.  LDS reads
      // are gathered into registers before any write of the same phase.
      const int ws = s0 + wave;
      const int w_steps = (wave < nsw) ? sweep_nsteps(n, b, ws) : 0;
      const int c = lane & 31, h = lane >> 5;
      int wp = ws + 1, wlen = min(b, n - wp), cB1 = ws, colB1 = 1;
      bool wfirst = true;
      double wbeta = 0.0;
      double *u = uS[wave], *yv = yS[wave];
      for (int t = -3; t <= t_last; ++t) {
        const int js = t - 2 * wave;
        if (js >= 0 && js < w_steps) {
          const int p = wp, len = wlen;
          const int len2 = max(0, min(b, n - p - len));
          const bool last = (len2 <= 1);
          if (wfirst) {
            // the sweep's first reflector, from column ws below the diagonal
            const double x = (lane < len) ? win[(ws & (kGW - 1)) * kGLD + 1 + lane] : 0.0;
            double uu, beta;
            grp_reflector(x, lane, len, uu, beta);
            if (lane < 32) u[lane] = uu;
            if (lane < len) U[(p + lane) + (long)ws * ldu] = (float)uu;
            wbeta = beta;
          }
          double u16[16];
#pragma unroll
          for (int r = 0; r < 16; r++) u16[r] = u[2 * r + h];
          const double uc = u[c];
          // ---- B1 <- H B1 (columns 1 .. colB1-1); column 0 becomes beta e1
          {
            const bool act = (c >= 1 && c < colB1);
            // row p of column cB1 + c (lanes past the block read slot 0, unused)
            const int base = (c < colB1) ? ((cB1 + c) & (kGW - 1)) * kGLD + (p - cB1 - c) : 0;
            double bv[16];
#pragma unroll
            for (int r = 0; r < 16; r++) bv[r] = win[base + 2 * r + h];
            double tsum = 0.0;
#pragma unroll
            for (int r = 0; r < 16; r++) {
              const int i = 2 * r + h;
              if (!(act && i < len)) bv[r] = 0.0;
              tsum += u16[r] * bv[r];
            }
            tsum += half_swap(tsum);
#pragma unroll
            for (int r = 0; r < 16; r++) {
              const int i = 2 * r + h;
              if (act && i < len) win[base + i] = bv[r] - u16[r] * tsum;
              if (c == 0 && i < len && !(i == 0 && !wfirst)) win[base + i] = (i == 0) ? wbeta : 0.0;
            }
          }
          // ---- S <- H S H, N <- N H (S from the lower band)
          const int ri = c, kc = h;
          double Sr[16], Nr[16];
#pragma unroll
          for (int r = 0; r < 16; r++) {
            const int kk = kc + 2 * r;
            const int lo = min(ri, kk), hi = max(ri, kk);
            Sr[r] = win[((p + lo) & (kGW - 1)) * kGLD + (hi - lo)];
            Nr[r] = win[((p + kk) & (kGW - 1)) * kGLD + ((len + ri - kk) & 63)];
          }
          double yp = 0.0, qp = 0.0;
#pragma unroll
          for (int r = 0; r < 16; r++) {
            const int kk = kc + 2 * r;
            if (!(ri < len && kk < len)) Sr[r] = 0.0;
            if (!(ri < len2 && kk < len)) Nr[r] = 0.0;
            yp += Sr[r] * u16[r];
            qp += Nr[r] * u16[r];
          }
          yp += half_swap(yp);  // y = S u, row ri (both halves)
          qp += half_swap(qp);  // q = N u, row ri
          if (kc == 0) yv[ri] = yp;
          double yk[16];
#pragma unroll
          for (int r = 0; r < 16; r++) yk[r] = yv[kc + 2 * r];
          // u^T y over the 64 lanes counts every row twice
          const double hpsi = 0.5 * (0.5 * wave_total(uc * yp));
          const double wi = yp - hpsi * uc;
#pragma unroll
          for (int r = 0; r < 16; r++) {
            const double wk = yk[r] - hpsi * u16[r];
            Sr[r] -= uc * wk + wi * u16[r];
            Nr[r] -= qp * u16[r];
          }
#pragma unroll
          for (int r = 0; r < 16; r++) {
            const int kk = kc + 2 * r;
            if (ri >= kk && ri < len && kk < len) win[((p + kk) & (kGW - 1)) * kGLD + (ri - kk)] = Sr[r];
            if (ri < len2 && kk < len) win[((p + kk) & (kGW - 1)) * kGLD + (len + ri - kk)] = Nr[r];
          }
          // ---- next reflector from N[:, 0] (column p, held by lanes kc = 0);
          // its beta goes in now
          if (!last) {
            const double x = (kc == 0 && ri < len2) ? Nr[0] : 0.0;
            double uu, beta;
            grp_reflector(x, lane, len2, uu, beta);
            if (lane < 32) u[lane] = uu;
            if (lane < len2) U[(p + len + lane) + (long)ws * ldu] = (float)uu;
            if (lane == 0) win[(p & (kGW - 1)) * kGLD + len] = beta;
            wbeta = beta;
            wfirst = false;
            cB1 = p;
            colB1 = len;
            wp = p + len;
            wlen = len2;
          }
        }
        grp_barrier();
        if (((t + 3) & 1) && s_abort) break;
      }
    } else {
      // ---- I/O wave (see the schedule above); chunk j lives in buffer j & 1
      int fp = s0 + 1, flen = min(b, n - s0 - 1), lj = 0;  // next front step to load
      int known = (g == 0) ? INT_MAX : -1, pv = -1;
      const int *pcom = com + (long)max(g - 1, 0) * kComStride;
      int tp = s0 + nsw, tlen = min(b, n - s0 - nsw), wb = s0, pval = 0;
      bool pend = false;
      GrpChunk ca, cb;
#pragma unroll
      for (int k = 0; k < kGChunk; k++) ca.v[k] = cb.v[k] = 0.0;
      ca.lo = ca.hi = cb.lo = cb.hi = 0;
      ca.syn = cb.syn = -1;
      auto io_step = [&](int t, GrpChunk &ch) {
        // all but the newest kGChunk loads (the other buffer's chunk) are done:
        // this buffer's chunk, last step's stores and poll
        wait_vm<kGChunk>();
        if (pend) {
          if (pval == INT_MAX) wait_mem();
          if (lane == 0) st_dev(&com[(long)g * kComStride], pval);
          pend = false;
        }
        known = max(known, pv);  // group 0: known is INT_MAX already
        // chunk t + 1 -> LDS
#pragma unroll
        for (int k = 0; k < kGChunk; k++) {
          const int cc = ch.lo + k;
          const int at = (cc < ch.hi && lane < b2) ? (cc & (kGW - 1)) * kGLD + lane : kGW * kGLD + lane;
          win[at] = (cc == ch.syn && lane > b) ? 0.0 : ch.v[k];
        }
        // write-back of the columns the last sweep finished at t - 1
        const int k = t - 1 - 2 * (nsw - 1);
        const bool wbk = (k >= 0 && k < tail_steps);
        const bool wlast = wbk && (k == tail_steps - 1);
        const int P = tp;
        const int whi = wbk ? (wlast ? n : P + 1) : wb;
        int c0 = wb;
        do {
          // kGChunk stores (two halves of reads then stores); lanes past the
          // columns or past a column's valid offsets store out of range
#pragma unroll
          for (int half = 0; half < 2; half++) {
            double sv[kGChunk / 2];
#pragma unroll
            for (int k2 = 0; k2 < kGChunk / 2; k2++) {
              const int cc = min(c0 + half * (kGChunk / 2) + k2, n - 1);
              sv[k2] = win[(cc & (kGW - 1)) * kGLD + lane];
            }
#pragma unroll
            for (int k2 = 0; k2 < kGChunk / 2; k2++) {
              const int cc = c0 + half * (kGChunk / 2) + k2;
              const int dmax = (wlast || cc < P) ? b2 : b + 1;
              const unsigned off = (cc < whi && lane < dmax) ? ((unsigned)lane + (unsigned)cc * ld) * 8u : kOutOfRange;
              __builtin_amdgcn_raw_buffer_store_b64(
                  __builtin_bit_cast(__attribute__((__vector_size__(2 * sizeof(unsigned)))) unsigned, sv[k2]), rb, off,
                  0, kSc1);
            }
          }
          c0 += kGChunk;
        } while (c0 < whi);
        if (wbk) {
          wb = whi;
          pend = true;
          pval = wlast ? INT_MAX : P + tlen;
          const int nl = max(0, min(b, n - P - tlen));
          tp = P + tlen;
          tlen = nl;
        }
        // poll of the predecessor's progress word (its own word for group 0)
        pv = ld_dev(pcom);
        // loads of front step t + 3
        const bool ld_now = (lj < front_steps);
        if (ld_now && known < fp + 2 * b - 1 && !s_abort) {
          const int need = fp + 2 * b - 1;
          const unsigned long long t0 = clock100();
          while (true) {
            known = max(known, ld_dev(pcom));
            if (known >= need) break;
            if (clock100() - t0 > kWatchdogTicks) {
              if (lane == 0) {
                atomicOr(err, (unsigned)EVD_E_TIMEOUT);
                s_abort = 1;
              }
              break;
            }
            __builtin_amdgcn_s_sleep(1);
          }
        }
        const int lo = (lj == 0) ? s0 : fp;
        const int hi = ld_now ? ((lj == front_steps - 1) ? n : fp + flen) : lo;
#pragma unroll
        for (int k2 = 0; k2 < kGChunk; k2++) {
          const int cc = lo + k2;
          const unsigned off = (cc < hi && lane < b2) ? ((unsigned)lane + (unsigned)cc * ld) * 8u : kOutOfRange;
          ch.v[k2] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rb, off, 0, kSc1));
        }
        ch.lo = lo;
        ch.hi = hi;
        ch.syn = fp + b - 1;
        if (ld_now) {
          fp += flen;
          flen = min(b, n - fp);
          lj++;
        }
        grp_barrier();
      };
      for (int t = -3; t <= t_last; t += 2) {
        io_step(t, ca);
        io_step(t + 1, cb);
        if (s_abort) break;
      }
      wait_mem();
    }
    if (s_abort) {
      if (wave == D && lane == 0) st_dev(&com[(long)g * kComStride], INT_MAX);
      return;
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

int band_pack(evd_ctx *ctx, long n, long b, const float *A, long lda, double *band, long ldband) {
  if (n <= 0) return 0;
  dim3 blk(64, 4);
  unsigned gy = (unsigned)std::min<long>((n + 3) / 4, 65535);
  band_pack_kernel<<<dim3(1, gy), blk, 0, ctx->stream>>>((int)n, (int)b, A, lda, band, ldband);
  EVD_LAUNCHED(ctx);
  return 0;
}

// largest sweeps-per-group whose window fits kGW columns: (2D + 2) b - D + 1 <= kGW
static int group_depth_max(long b) { return (int)std::max<long>(0, (kGW - 1 - 2 * b) / (2 * b - 1)); }

template <int D>
static int launch_chase_group(evd_ctx *ctx, int count, long n, long b, double *band, long ldb, long band_stride,
                              float *U, long ldu, long u_stride, int *com, long com_stride, unsigned *ticket) {
  const long bytes = ldb * n * (long)sizeof(double);
  if (bytes >= (1L << 31))
    return set_error(ctx, EVD_E_UNSUPPORTED, "band of %ld bytes exceeds the 2 GiB buffer range of the chase", bytes);
  // groups in flight: a group lags its predecessor by about 2D + 4 steps and
  // lives about n / b steps
  int resident = 0;
  EVD_TRY(max_coresident_blocks(ctx, (const void *)chase_group_kernel<D>, (D + 1) * 64, 0, &resident));
  const long ngroups = (n - 2 + D - 1) / D;
  const long inflight = (((n + b - 1) / b + 2 * D + 3) / (2 * D + 4) + 2) * count;
  long G = ctx->tune.chase_groups > 0 ? ctx->tune.chase_groups : inflight;
  G = std::max<long>(1, std::min<long>(std::min<long>(G, resident), ngroups * count));
  unsigned *err = ctx->ctrl + kCtrlError;
  if (ctx->timing && !ctx->capturing) hipEventRecord(ctx->cev[0], ctx->stream);
  chase_group_kernel<D><<<(unsigned)G, (D + 1) * 64, 0, ctx->stream>>>((int)n, (int)b, band, ldb, (unsigned)bytes, U,
                                                                        ldu, com, ticket, err, count, band_stride,
                                                                        u_stride, com_stride);
  EVD_LAUNCHED(ctx);
  if (ctx->timing && !ctx->capturing) hipEventRecord(ctx->cev[1], ctx->stream);
  return 0;
}

template <int BW>
static int launch_chase(evd_ctx *ctx, int count, long n, long b, double *band, long ldb, long band_stride, float *U,
                        long ldu, long u_stride, int *com, long com_stride, unsigned *ticket) {
  const long bytes = ldb * n * (long)sizeof(double);
  if (bytes >= (1L << 31))
    return set_error(ctx, EVD_E_UNSUPPORTED, "band of %ld bytes exceeds the 2 GiB buffer range of the chase", bytes);
  // workgroups: sweep s+1 trails sweep s by two b-row steps, so at most
  // about n / (2b) sweeps of a problem are in flight; by default that many
  // plus a few per problem (a workgroup takes its next sweep by ticket as
  // soon as it finishes one), at most the resident count.  Workgroups beyond
  // that only spin on progress words, slow the chase through the polling
  // traffic (n = 16384: 148 ms with 260 workgroups, 163 ms with 1.25x,
  // 191 ms with all resident ones) and crowd out concurrent solves.
  int resident = 0;
  EVD_TRY(max_coresident_blocks(ctx, (const void *)chase_kernel<BW>, 256, 0, &resident));
  const long inflight = ((n + 2 * b - 1) / (2 * b) + 4) * count;
  long G = ctx->tune.chase_groups > 0 ? ctx->tune.chase_groups
                                      : std::min<long>(inflight, (long)std::max(ctx->num_cus, resident));
  G = std::min<long>(G, std::max<long>(1, (n - 2) * count));
  int ni = (int)n, bi = (int)b;
  unsigned nbytes = (unsigned)bytes;
  unsigned *err = ctx->ctrl + kCtrlError;
  if (ctx->timing && !ctx->capturing) hipEventRecord(ctx->cev[0], ctx->stream);
  chase_kernel<BW><<<(unsigned)G, 256, 0, ctx->stream>>>(ni, bi, band, ldb, nbytes, U, ldu, com, ticket, err, count,
                                                          band_stride, u_stride, com_stride);
  EVD_LAUNCHED(ctx);
  if (ctx->timing && !ctx->capturing) hipEventRecord(ctx->cev[1], ctx->stream);
  return 0;
}

__global__ void band_extract_batched_kernel(int n, const double *__restrict__ band, long ldb, long band_stride,
                                            double *__restrict__ D, long d_stride, double *__restrict__ E,
                                            long e_stride) {
  const long q = blockIdx.y;
  band += q * band_stride;
  D += q * d_stride;
  E += q * e_stride;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    D[i] = band[i * ldb];
    if (i < n - 1) E[i] = band[1 + i * ldb];
  }
}

// Bulge chasing of `count` same-size bands in one persistent launch (problem q
// at band + q band_stride, U + q u_stride, D + q d_stride, E + q e_stride).
// The progress words and the ticket come from the arena, or from ext_com
// (at least sb2st_batched_workspace(count, n) bytes) when it is given.
int sb2st_batched(evd_ctx *ctx, int count, long n, long b, double *band, long ldband, long band_stride, float *U,
                  long ldu, long u_stride, double *D, long d_stride, double *E, long e_stride, int *ext_com) {
  if (n <= 0 || count <= 0) return 0;
  for (int q = 0; q < count; q++) EVD_TRY(fill_f32(ctx, U + (long)q * u_stride, ldu, n, ldu, 0.f));
  if (n > 2) {
    ArenaFrame frame(ctx);
    // progress words com[q * n * kComStride + s * kComStride], then the ticket
    const long com_stride = n * kComStride;
    const size_t words = (size_t)count * com_stride + 1;
    int *com = ext_com ? ext_com : arena_alloc_n<int>(ctx, words);
    if (!com) return EVD_E_NOMEM;
    unsigned *ticket = reinterpret_cast<unsigned *>(com + (size_t)count * com_stride);
    EVD_TRY(fill_u32(ctx, reinterpret_cast<unsigned *>(com), words, 0u));
    // sweeps per workgroup: the knob (0 = 1, one sweep per workgroup), at most
    // what the LDS window holds and at most 3 (with more than four waves a
    // workgroup's waves share SIMDs and the grouped kernel's registers spill)
    const int depth = std::min({ctx->tune.chase_depth > 0 ? ctx->tune.chase_depth : 1, group_depth_max(b), 3});
    if (b <= 32 && depth == 3)
      EVD_TRY(launch_chase_group<3>(ctx, count, n, b, band, ldband, band_stride, U, ldu, u_stride, com, com_stride,
                                    ticket));
    else if (b <= 32 && depth == 2)
      EVD_TRY(launch_chase_group<2>(ctx, count, n, b, band, ldband, band_stride, U, ldu, u_stride, com, com_stride,
                                    ticket));
    else if (b <= 32)
      EVD_TRY(launch_chase<32>(ctx, count, n, b, band, ldband, band_stride, U, ldu, u_stride, com, com_stride, ticket));
    else if (b <= 64)
      EVD_TRY(launch_chase<64>(ctx, count, n, b, band, ldband, band_stride, U, ldu, u_stride, com, com_stride, ticket));
    else
      return set_error(ctx, EVD_E_UNSUPPORTED, "bandwidth %ld > 64", b);
  }
  band_extract_batched_kernel<<<dim3((unsigned)std::min<long>((n + 255) / 256, 1024), (unsigned)count), 256, 0,
                                ctx->stream>>>((int)n, band, ldband, band_stride, D, d_stride, E, e_stride);
  EVD_LAUNCHED(ctx);
  return 0;
}

size_t sb2st_batched_workspace(int count, long n) { return align256(((size_t)count * n * kComStride + 1) * sizeof(int)); }

int sb2st(evd_ctx *ctx, long n, long b, double *band, long ldband, float *U, long ldu, double *D, double *E) {
  return sb2st_batched(ctx, 1, n, b, band, ldband, 0, U, ldu, 0, D, 0, E, 0, nullptr);
}

// ---- the chase without allocation: scratch in a file-scope device array.
// Progress words (one per sweep, kComStride apart) and the sweep ticket of
// one problem of order <= kNoAllocMaxN: 4 MiB per device, held until the
// process ends.  Host code never passes the array's name to a kernel (it is a
// host-side stand-in); it uses the device address that hipGetSymbolAddress
// returns, looked up once per device with that device current.
constexpr long kNoAllocMaxN = 32768;
constexpr size_t kNoAllocWords = (size_t)kNoAllocMaxN * kComStride + 64;
static __device__ __align__(256) int scratch[kNoAllocWords];

static int noalloc_scratch(evd_ctx *ctx, int **out) {
  static std::mutex mu;
  static std::map<int, int *> ptrs;
  std::lock_guard<std::mutex> lk(mu);
  auto it = ptrs.find(ctx->device);
  if (it != ptrs.end()) {
    *out = it->second;
    return 0;
  }
  int cur = -1;
  EVD_HIP(ctx, hipGetDevice(&cur));
  if (cur != ctx->device) EVD_HIP(ctx, hipSetDevice(ctx->device));
  int *p = nullptr;
  const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
  if (cur != ctx->device) EVD_HIP(ctx, hipSetDevice(cur));
  if (e != hipSuccess || !p)
    return set_error(ctx, EVD_E_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
  ptrs[ctx->device] = p;
  *out = p;
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
  if (ctx->arena_top == 0) EVD_TRY(arena_ensure(ctx, sb2st_batched_workspace(1, n) + 4096));
  EVD_TRY(sb2st(ctx, n, b, band, ldband, U, ldu, D, E));
  return check_device_error(ctx, "evd_sb2st_f64");
}

extern "C" int evd_sb2st_f64_noalloc(evd_ctx *ctx, long n, long b, double *band, long ldband, float *U, long ldu,
                                     double *D, double *E) {
  if (!ctx) return -1;
  EVD_ARG(ctx, n >= 0, 2);
  EVD_ARG(ctx, b >= 1 && b <= 64, 3);
  EVD_ARG(ctx, ldband >= 2 * b, 5);
  EVD_ARG(ctx, ldu >= n + b, 7);
  if (n > kNoAllocMaxN)
    return set_error(ctx, EVD_E_UNSUPPORTED, "evd_sb2st_f64_noalloc: n = %ld exceeds the scratch limit n <= %ld", n,
                     kNoAllocMaxN);
  static_assert(kNoAllocWords * sizeof(int) >= ((size_t)kNoAllocMaxN * kComStride + 1) * sizeof(int),
                "scratch holds the progress words and the ticket of the largest order");
  int *com = nullptr;
  EVD_TRY(noalloc_scratch(ctx, &com));
  EVD_TRY(sb2st_batched(ctx, 1, n, b, band, ldband, 0, U, ldu, 0, D, 0, E, 0, com));
  return check_device_error(ctx, "evd_sb2st_f64_noalloc");
}

extern "C" int evd_sb2st_batched_f64(evd_ctx *ctx, int count, long n, long b, double *band, long ldband,
                                     long band_stride, float *U, long ldu, long u_stride, double *D, long d_stride,
                                     double *E, long e_stride) {
  if (!ctx) return -1;
  EVD_ARG(ctx, count >= 0 && count <= 65535, 2);  // the band extraction carries the problem in grid y
  EVD_ARG(ctx, n >= 0, 3);
  EVD_ARG(ctx, b >= 1 && b <= 64, 4);
  EVD_ARG(ctx, ldband >= 2 * b, 6);
  EVD_ARG(ctx, count <= 1 || band_stride >= ldband * n, 7);
  EVD_ARG(ctx, ldu >= n + b, 9);
  EVD_ARG(ctx, count <= 1 || u_stride >= ldu * n, 10);
  EVD_ARG(ctx, count <= 1 || d_stride >= n, 12);
  EVD_ARG(ctx, count <= 1 || e_stride >= std::max<long>(1, n - 1), 14);
  EVD_ARG(ctx, (long)count * std::max<long>(n, 1) < (1L << 30), 2);
  if (ctx->arena_top == 0) EVD_TRY(arena_ensure(ctx, sb2st_batched_workspace(count, n) + 4096));
  EVD_TRY(sb2st_batched(ctx, count, n, b, band, ldband, band_stride, U, ldu, u_stride, D, d_stride, E, e_stride,
                        nullptr));
  return check_device_error(ctx, "evd_sb2st_batched_f64");
}
