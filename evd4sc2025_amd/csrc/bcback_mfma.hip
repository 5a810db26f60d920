// Mixed-precision back transformation of the bulge-chasing reflectors on the
// 16-bit matrix cores: Q <- H_1 H_2 .. H_K Q (the product the eigenvector
// path needs, the role of my_BC_back_trans_v8_10_noBandU,
// api/BC_backTransV8_10_noBandU.cu:60-313), with the reflectors aggregated
// into compact WY blocks and applied as split-FP16 MFMA products.
//
// Blocks.  Sweeps are cut into blocks of 32; block (S, j) holds the 32
// reflectors of sweeps 32S-1 .. 32S+30 at step j (sweep -1 does not exist and
// counts as the identity).  Reflector t of the block (I - u u^T, |u|^2 = 2)
// covers rows r0 + t .. r0 + t + 31 with r0 = 32S + 32j, so the block is a
// 64 x 32 staircase V (column t nonzero in window rows t .. t+31; row 63 is
// always zero) on a window that starts on a 128-byte line of Q's columns.  The block's product in
// sweep order is H_0 H_1 .. H_31 = I - V T V^T with T upper triangular (the
// forward columnwise recurrence T[0:t, t] = -T[0:t, 0:t] V[:, 0:t]^T v_t,
// T[t][t] = 1), and applying it is
//     X = V^T Zw   (32 x c),     Zw <- Zw - (V T) X   (64 x c)
// on the 64-row window Zw of the block.  Blocks run S descending, j ascending
// (reflectors that share rows are applied higher sweep first).
//
// Precision.  V T and V are formed in FP64 (bcw_prep_kernel) and stored as
// FP16 hi + lo pairs; the window is kept in FP32 accumulators scaled by 2^10
// and split into hi + lo right before it feeds an MFMA, X likewise.  Each
// product sums lo*hi + hi*lo + hi*hi in FP32 (22 significant bits per
// operand).  The scale keeps the lo terms of eigenvector entries out of the
// FP16 subnormal range; it requires columns of Q with 2-norm <= 8 (the
// eigenvector matrix of the tridiagonal stage has unit columns), so that
// |Zw| and |X| stay below the FP16 range.
//
// MI355X mapping.  One workgroup owns 64 columns of Q; its NW waves form a
// pipeline over NW consecutive sweep blocks (wave k works on block S_top - k
// one step behind wave k-1) and hand each other the 32 rows they have
// finished through an LDS double buffer, so Q is streamed from HBM once per
// 32 NW sweeps.  A wave holds its 64 x 64 window as 4 x 4 accumulator tiles of
// v_mfma_f32_16x16x32_f16; the K index of the 32-deep operands is permuted so
// that the two 16-row accumulator tiles a lane holds are exactly the 8 K
// values its operand slot needs (kmap below): accumulators become operands
// without any data movement between lanes.  The two window halves alternate
// roles (top/bottom) from step to step instead of being copied.
#include <type_traits>

#include "common.hpp"

namespace evd {

namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef float f2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kB = 32;                 // reflector length = sweeps per block
constexpr int kFrag = 64;              // one operand fragment: 64 lanes x 8 halves
constexpr int kBlockFrags = 16;        // V^T (4 fragments) and -V T (4), hi and lo each
constexpr float kScale = 1024.f;       // window scale (2^10)
constexpr float kHalfMax = 60000.f;    // range guard for the stored operands

// K slot e (0..7) of lane group g of a 32-deep operand <-> row of the two
// 16-row accumulator tiles the lane holds (rows 4g..4g+3 and 16+4g..16+4g+3).
__host__ __device__ constexpr int kmap(int g, int e) { return e < 4 ? 4 * g + e : 16 + 4 * g + (e - 4); }

// sweep block S has Jn - S steps (Jn = ceil(n / 32): windows start below n),
// and its blocks start at index prefix(S)
__host__ __device__ inline long wy_steps(long Jn, long S) { return Jn - S; }
__host__ __device__ inline long wy_prefix(long Jn, long S) { return S * Jn - S * (S - 1) / 2; }
__host__ __device__ inline long wy_sweep_blocks(long nsweeps) { return (nsweeps + kB) / kB; }

// ------------------------------------------------------------------ prep
// One workgroup per block (S, j): V from U, T (forward recurrence, solved
// column-parallel as T = (I + strict upper part of V^T V)^-1), V T,
// and the MFMA operand fragments of V^T and -(V T), hi and lo halves.
// A block is 16 fragments of 64 lanes x 16 bytes.
// Split FP16 (h8 per lane, K = 32 per MFMA):
//   [0..7]  V^T: fragment f = 2 ks + mt (ks: 32-row K step, mt: 16-sweep tile), term u at 2 f + u
//   [8..15] -V T: fragment 4 + mt (mt: 16-row tile), term u at 8 + 2 mt + u
// FP32 (f4 per lane = the 4 K=4 MFMAs over one 16-row tile, component i
// taking rows 4g + i of the tile, the rows an accumulator lane holds):
//   [0..7]  V^T: fragment 2 rt + mt (rt: 16-row window tile, mt: 16-sweep tile)
//   [8..15] -V T: fragment 8 + 2 mt + st (mt: 16-row window tile, st: 16-sweep tile)
template <bool F32>
__global__ __launch_bounds__(256) void bcw_prep_kernel(int n, const float *__restrict__ U, long ldu, int nsweeps,
                                                       h8 *__restrict__ F, unsigned *__restrict__ err, long u_stride,
                                                       long f_stride) {
  U += blockIdx.z * u_stride;  // problem blockIdx.z of a batch
  F += blockIdx.z * f_stride;
  const long Jn = (n + kB - 1) / kB;
  const int S = blockIdx.y, j = blockIdx.x;
  if (j >= wy_steps(Jn, S)) return;
  __shared__ double V[64][33];
  __shared__ double G[32][33];
  __shared__ double T[32][33];
  __shared__ double VT[64][33];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < 64 * 32; idx += 256) {
    const int r = idx & 63, t = idx >> 6, i = r - t;
    const long sw = (long)S * kB - 1 + t;
    const long row = (long)S * kB + (long)j * kB + r;
    double v = 0.0;
    if (i >= 0 && i < kB && sw >= 0 && sw < nsweeps && row < n) v = (double)U[row + sw * ldu];
    V[r][t] = v;
  }
  __syncthreads();
  // G[a][c] = v_a . v_c for a < c: the columns overlap in rows c .. a+31
  for (int idx = tid; idx < 32 * 32; idx += 256) {
    const int a = idx & 31, c = idx >> 5;
    double s = 0.0;
    if (a < c)
      for (int r = c; r < a + kB; r++) s += V[r][a] * V[r][c];
    G[a][c] = s;
  }
  __syncthreads();
  // With tau = 1 the recurrence gives T = (I + strict upper part of G)^-1:
  // column c by back substitution, one thread per column, no barriers.
  if (tid < kB) {
    const int c = tid;
    for (int i = 0; i < kB; i++) T[i][c] = (i == c) ? 1.0 : 0.0;
    for (int i = c - 1; i >= 0; i--) {
      double s = 0.0;
      for (int j = i + 1; j <= c; j++) s += G[i][j] * T[j][c];
      T[i][c] = -s;
    }
  }
  __syncthreads();
  for (int idx = tid; idx < 64 * 32; idx += 256) {
    const int r = idx & 63, t = idx >> 6;
    double s = 0.0;
    for (int q = 0; q <= t; q++) s += V[r][q] * T[q][t];
    VT[r][t] = s;
  }
  __syncthreads();
  const long blk = wy_prefix(Jn, S) + j;
  h8 *out = F + blk * (kBlockFrags * kFrag);
  if constexpr (F32) {
    f4 *o4 = reinterpret_cast<f4 *>(out);
    for (int idx = tid; idx < kBlockFrags * kFrag; idx += 256) {
      const int q = idx >> 6, lane = idx & 63, g = lane >> 4, m = lane & 15;
      f4 v;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        if (q < 8)
          v[i] = (float)V[16 * (q >> 1) + 4 * g + i][16 * (q & 1) + m];
        else
          v[i] = (float)-VT[16 * ((q - 8) >> 1) + m][16 * ((q - 8) & 1) + 4 * g + i];
      }
      o4[idx] = v;
    }
    return;
  }
  bool bad = false;
  for (int idx = tid; idx < 8 * kFrag; idx += 256) {
    const int f = idx >> 6, lane = idx & 63, g = lane >> 4, m = lane & 15;
    h8 hi, lo;
#pragma unroll
    for (int e = 0; e < 8; e++) {
      double x;
      if (f < 4) {
        const int ks = f >> 1, mt = f & 1;
        x = V[32 * ks + kmap(g, e)][16 * mt + m];
      } else {
        x = -VT[16 * (f - 4) + m][kmap(g, e)];
      }
      const float xf = (float)x;
      bad |= !(fabsf(xf) < kHalfMax);
      const _Float16 h = (_Float16)xf;
      hi[e] = h;
      lo[e] = (_Float16)(float)(x - (double)(float)h);
    }
    out[(2 * f) * kFrag + lane] = hi;
    out[(2 * f + 1) * kFrag + lane] = lo;
  }
  if (bad) atomicOr(err, (unsigned)EVD_E_RANGE);
}

// ------------------------------------------------------------------ main
// x - float(h) for the low / high FP16 half of h, one v_fma_mix each
__device__ __attribute__((always_inline)) inline float resid_lo(h2 h, float x) {
  float r;
  asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h), "v"(x));
  return r;
}
__device__ __attribute__((always_inline)) inline float resid_hi(h2 h, float x) {
  float r;
  asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h), "v"(x));
  return r;
}

// split 8 FP32 values (4 from each of two accumulator tiles) into FP16 hi, lo
__device__ __attribute__((always_inline)) inline void split8(const f4 &a, const f4 &b, h8 &hi, h8 &lo) {
#pragma unroll
  for (int e = 0; e < 8; e += 2) {
    const float x0 = e < 4 ? a[e] : b[e - 4], x1 = e < 4 ? a[e + 1] : b[e - 3];
    const h2 h = __builtin_convertvector(f2{x0, x1}, h2);
    const h2 l = __builtin_convertvector(f2{resid_lo(h, x0), resid_hi(h, x1)}, h2);
    hi[e] = h[0];
    hi[e + 1] = h[1];
    lo[e] = l[0];
    lo[e + 1] = l[1];
  }
}

__device__ __attribute__((always_inline)) inline f4 mma(const h8 &a, const h8 &b, const f4 &c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

__device__ __attribute__((always_inline)) inline f4 mma(float a, float b, const f4 &c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Element (row r, column c) of a 32 x 64 half window in accumulator order
// (tile 4 mt + nt, lane 16 g + c16, component i), as a float index.
__device__ __attribute__((always_inline)) inline int half_index(int r, int c) {
  return (((r >> 4) * 4 + (c >> 4)) * 64 + ((r >> 2) & 3) * 16 + (c & 15)) * 4 + (r & 3);
}

// LDS slot: 32 rows x 64 columns in accumulator order
__device__ __attribute__((always_inline)) inline void put_half(f4 *slot, const f4 (&h)[2][4], int lane) {
#pragma unroll
  for (int mt = 0; mt < 2; mt++)
#pragma unroll
    for (int nt = 0; nt < 4; nt++) slot[(mt * 4 + nt) * 64 + lane] = h[mt][nt];
}
__device__ __attribute__((always_inline)) inline void get_half(f4 (&h)[2][4], const f4 *slot, int lane) {
#pragma unroll
  for (int mt = 0; mt < 2; mt++)
#pragma unroll
    for (int nt = 0; nt < 4; nt++) h[mt][nt] = slot[(mt * 4 + nt) * 64 + lane];
}

// fragments of a block: V^T of K step ks (4 x h8: mt 0 hi, lo, mt 1 hi, lo), -V T (8 x h8)
template <typename FR, int N>
__device__ __attribute__((always_inline)) inline void load_frags(FR (&fr)[N], const FR *__restrict__ src, int lane) {
#pragma unroll
  for (int q = 0; q < N; q++) fr[q] = src[q * kFrag + lane];
}

// Zw <- (I - V T V^T) Zw for one block; top = window rows 0..31, bot = 32..63.
// v0 / v1 hold V^T for window rows 0..31 / 32..63, wt holds -V T.  With nxt
// given, the next block's v0, v1 and wt are loaded as soon as the current ones
// are consumed, so their latency hides behind the rest of the step and the
// next one's start.
__device__ __attribute__((always_inline)) inline void apply_block(f4 (&top)[2][4], f4 (&bot)[2][4], h8 (&v0)[4],
                                                                  h8 (&v1)[4], h8 (&wt)[8],
                                                                  const h8 *__restrict__ nxt, int lane) {
  f4 x[2][4];
#pragma unroll
  for (int mt = 0; mt < 2; mt++)
#pragma unroll
    for (int nt = 0; nt < 4; nt++) x[mt][nt] = f4{0.f, 0.f, 0.f, 0.f};
  // X = V^T Zw: K = 64 window rows in two steps, one column tile at a time
  // (products lo*hi, hi*lo, hi*hi)
#pragma unroll
  for (int nt = 0; nt < 4; nt++) {
    h8 zh, zl;
    split8(top[0][nt], top[1][nt], zh, zl);
#pragma unroll
    for (int mt = 0; mt < 2; mt++) x[mt][nt] = mma(v0[2 * mt + 1], zh, x[mt][nt]);
#pragma unroll
    for (int mt = 0; mt < 2; mt++) x[mt][nt] = mma(v0[2 * mt], zl, x[mt][nt]);
#pragma unroll
    for (int mt = 0; mt < 2; mt++) x[mt][nt] = mma(v0[2 * mt], zh, x[mt][nt]);
  }
  __builtin_amdgcn_sched_barrier(0);
  if (nxt) load_frags(v0, nxt, lane);
#pragma unroll
  for (int nt = 0; nt < 4; nt++) {
    h8 zh, zl;
    split8(bot[0][nt], bot[1][nt], zh, zl);
#pragma unroll
    for (int mt = 0; mt < 2; mt++) x[mt][nt] = mma(v1[2 * mt + 1], zh, x[mt][nt]);
#pragma unroll
    for (int mt = 0; mt < 2; mt++) x[mt][nt] = mma(v1[2 * mt], zl, x[mt][nt]);
#pragma unroll
    for (int mt = 0; mt < 2; mt++) x[mt][nt] = mma(v1[2 * mt], zh, x[mt][nt]);
  }
  __builtin_amdgcn_sched_barrier(0);
  if (nxt) load_frags(v1, nxt + 4 * kFrag, lane);
  // Zw += (-V T) X, K = 32 sweeps
#pragma unroll
  for (int nt = 0; nt < 4; nt++) {
    h8 xh, xl;
    split8(x[0][nt], x[1][nt], xh, xl);
#pragma unroll
    for (int u = 0; u < 3; u++) {
      const int ua = u == 0 ? 1 : 0;  // lo*hi, hi*lo, hi*hi
      const h8 &b = (u == 1) ? xl : xh;
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        if (mt < 2)
          top[mt][nt] = mma(wt[2 * mt + ua], b, top[mt][nt]);
        else
          bot[mt - 2][nt] = mma(wt[2 * mt + ua], b, bot[mt - 2][nt]);
      }
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  if (nxt) load_frags(wt, nxt + 8 * kFrag, lane);
}

// The FP32 form of apply_block: V^T and -V T as f4 fragments, K = 4 per
// v_mfma_f32_16x16x4f32; accumulator tiles are the B operands as they are.
__device__ __attribute__((always_inline)) inline void apply_block(f4 (&top)[2][4], f4 (&bot)[2][4], f4 (&v0)[4],
                                                                  f4 (&v1)[4], f4 (&wt)[8],
                                                                  const f4 *__restrict__ nxt, int lane) {
  f4 x[2][4];
#pragma unroll
  for (int mt = 0; mt < 2; mt++)
#pragma unroll
    for (int nt = 0; nt < 4; nt++) x[mt][nt] = f4{0.f, 0.f, 0.f, 0.f};
  // X = V^T Zw.  Sweep tile mt of the staircase only reaches window tiles
  // mt .. mt+2, and column t of V T only rows 0 .. t+31: the zero tiles
  // (window tile 0 x sweep tile 1, window tile 3 x sweep tile 0) are skipped.
#pragma unroll
  for (int rt = 0; rt < 2; rt++)
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int nt = 0; nt < 4; nt++)
          if (!(rt == 0 && mt == 1)) x[mt][nt] = mma(v0[2 * rt + mt][i], top[rt][nt][i], x[mt][nt]);
  __builtin_amdgcn_sched_barrier(0);
  if (nxt) load_frags(v0, nxt, lane);
#pragma unroll
  for (int rt = 0; rt < 2; rt++)
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int nt = 0; nt < 4; nt++)
          if (!(rt == 1 && mt == 0)) x[mt][nt] = mma(v1[2 * rt + mt][i], bot[rt][nt][i], x[mt][nt]);
  __builtin_amdgcn_sched_barrier(0);
  if (nxt) load_frags(v1, nxt + 4 * kFrag, lane);
  // Zw += (-V T) X
#pragma unroll
  for (int st = 0; st < 2; st++)
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
          if (mt == 3 && st == 0) continue;
          if (mt < 2)
            top[mt][nt] = mma(wt[2 * mt + st][i], x[st][nt][i], top[mt][nt]);
          else
            bot[mt - 2][nt] = mma(wt[2 * mt + st][i], x[st][nt][i], bot[mt - 2][nt]);
        }
  __builtin_amdgcn_sched_barrier(0);
  if (nxt) load_frags(wt, nxt + 8 * kFrag, lane);
}

__device__ __attribute__((always_inline)) inline void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// HBM <-> LDS traffic of a half window, spread over the 8 waves: wave k moves
// columns 8k .. 8k+7, one instruction = 2 columns x 32 consecutive rows.
// The accesses are raw buffer operations over the workgroup's 64 columns that
// are always issued: a row or column that must be skipped gets an offset past
// the buffer, which the hardware drops (stores) or reads as 0 (loads).  With a
// fixed number of memory operations per step the compiler can wait for the
// block fragments without also waiting for this traffic.
constexpr unsigned kBufWord3 = 0x00020000u;  // raw buffer descriptor, 32-bit data
constexpr unsigned kOutside = 0x80000000u;   // offset beyond every buffer here
struct HalfIO {
  __amdgpu_buffer_rsrc_t rs;
  int r, c;         // row of the lane, first column of the lane (within the 64)
  unsigned cbytes;  // byte offset of the lane's first column
  int ncl;          // columns of this workgroup
  float scale;      // window scale (loads multiply, stores divide)
  __device__ HalfIO(float *Zc, long ldz, int n, int ncols_wg, int lane, int k, float sc)
      : rs(__builtin_amdgcn_make_buffer_rsrc(Zc, 0, (int)(((long)(ncols_wg - 1) * ldz + n) * 4), kBufWord3)),
        r(lane & 31),
        c(8 * k + (lane >> 5)),
        cbytes((unsigned)(c * ldz * 4)),
        ncl(ncols_wg),
        scale(sc) {}
  __device__ __attribute__((always_inline)) inline unsigned off(bool on, long row0, int n, long ldz, int q) const {
    const long gr = row0 + r;
    return (on && gr < n && c + 2 * q < ncl) ? cbytes + (unsigned)((2 * q * ldz + gr) * 4) : kOutside;
  }
  __device__ __attribute__((always_inline)) inline void load(float (&v)[4], bool on, long row0, int n, long ldz) const {
#pragma unroll
    for (int q = 0; q < 4; q++)
      v[q] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, off(on, row0, n, ldz, q), 0, 0)) * scale;
  }
  __device__ __attribute__((always_inline)) inline void park(f4 *slot, const float (&v)[4]) const {
    float *s = reinterpret_cast<float *>(slot);
#pragma unroll
    for (int q = 0; q < 4; q++) s[half_index(r, c + 2 * q)] = v[q];
  }
  __device__ __attribute__((always_inline)) inline void store(const f4 *slot, bool on, long row0, int n,
                                                              long ldz) const {
    const float *s = reinterpret_cast<const float *>(slot);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const float v = s[half_index(r, c + 2 * q)] / scale;
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rs, off(on, row0, n, ldz, q), 0, 0);
    }
  }
};

template <int NW, bool F32>
__global__ __launch_bounds__(NW * 64) void bcw_kernel(int n, int ncols, float *__restrict__ Z, long ldz,
                                                      const h8 *__restrict__ F16, int nsweeps, long z_stride,
                                                      long f_stride) {
  typedef typename std::conditional<F32, f4, h8>::type Frag;
  Z += blockIdx.y * z_stride;  // problem blockIdx.y of a batch
  const Frag *F = reinterpret_cast<const Frag *>(F16 + blockIdx.y * f_stride);
  static_assert(NW == 8, "each wave moves 8 of the 64 columns of a half window");
  // hand-off slots (2 per wave), the rows wave 0 takes in at each step, and
  // the top rows of each wave's first step (both double-buffered): 160 KiB
  __shared__ f4 xch[NW * 2 * 512];
  __shared__ f4 stage[2 * 512];
  __shared__ f4 stage_top[2 * 512];
  const int lane = threadIdx.x & 63;
  const int k = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int col0 = blockIdx.x * 64;
  const int nblk = (int)wy_sweep_blocks(nsweeps);
  const long Jn = (n + kB - 1) / kB;
  const HalfIO io(Z + (long)col0 * ldz, ldz, n, min(64, ncols - col0), lane, k, F32 ? 1.f : kScale);
  f4 zt[2][4], zb[2][4];
  Frag v0[4], v1[4], wt[8];

  const int ngroups = (nblk + NW - 1) / NW;
  for (int gi = 0; gi < ngroups; gi++) {
    const int gtop = nblk - 1 - gi * NW;
    const int nact = min(NW, gtop + 1);
    const int bk = gtop - k;
    // steps of the lowest block, plus one that hands its last rows on
    const int J = (int)wy_steps(Jn, gtop - nact + 1) + 1;
    const int T = J + nact - 1;
    const int kl = nact - 1;  // last active wave
    const long nst = wy_steps(Jn, bk);
    const Frag *fb = F + wy_prefix(Jn, bk) * (long)(kBlockFrags * kFrag);
    // One pipeline step.  Rows are loaded two steps before their use into
    // (lb, lt) and parked in LDS one step later from (pb, pt); the loop is
    // unrolled by two so the two register sets alternate without copies
    // (a copy would wait for the load).
    auto body = [&](int tau, float(&lb)[4], float(&lt)[4], const float(&pb)[4], const float(&pt)[4])
                    __attribute__((always_inline)) {
      const int e = tau - k;
      const bool act = tau >= 0 && tau < T && k < nact && e >= 0 && e < J;
      const bool app = act && e < nst;
      const Frag *cur = fb + (long)e * (kBlockFrags * kFrag);
      // (0) a wave's first block: its fragments before the HBM traffic
      // issued below, so waiting for them never waits for HBM (later blocks'
      // fragments are loaded one block ahead inside apply_block)
      if (app && e == 0) {
        load_frags(v0, cur, lane);
        load_frags(v1, cur + 4 * kFrag, lane);
        load_frags(wt, cur + 8 * kFrag, lane);
      }
      // (1) rows the last wave handed on at step tau-1 -> Q
      const int el = tau - 1 - kl;
      io.store(xch + (kl * 2 + (el & 1)) * 512, el >= 0 && el < J, (long)(gtop - kl + el) * kB, n, ldz);
      // (2) rows needed at step tau+2: wave 0's new bottom half, and the top
      // half of wave tau+2's first step
      io.load(lb, tau + 2 < J, (long)(gtop + tau + 3) * kB, n, ldz);
      io.load(lt, tau + 2 < nact, (long)(gtop - tau - 2) * kB, n, ldz);
      // (3) this wave's step
      if (act) {
        if (e == 0) {
          get_half(zt, stage_top + (k & 1) * 512, lane);
        } else {
#pragma unroll
          for (int mt = 0; mt < 2; mt++)
#pragma unroll
            for (int nt = 0; nt < 4; nt++) zt[mt][nt] = zb[mt][nt];
        }
        get_half(zb, k == 0 ? stage + (e & 1) * 512 : xch + ((k - 1) * 2 + (e & 1)) * 512, lane);
        if (app) apply_block(zt, zb, v0, v1, wt, (e + 1 < nst && e + 1 < J) ? cur + kBlockFrags * kFrag : nullptr, lane);
        put_half(xch + (k * 2 + (e & 1)) * 512, zt, lane);
      }
      // (4) park the rows loaded at step tau-1 (for step tau+1)
      if (tau >= -1 && tau + 1 < J) io.park(stage + ((tau + 1) & 1) * 512, pb);
      if (tau >= -1 && tau + 1 < nact) io.park(stage_top + ((tau + 1) & 1) * 512, pt);
      lds_barrier();
    };
    float la[4], ta[4], lb2[4], tb2[4];
    for (int tau = -2; tau <= T; tau += 2) {
      body(tau, la, ta, lb2, tb2);
      if (tau + 1 <= T) body(tau + 1, lb2, tb2, la, ta);
    }
    // the next group reads rows stored above
    wait_mem();
    __syncthreads();
  }
}

// ------------------------------------------------------------------ b = 64
// The same pipeline for reflectors of length 64 (FP32 MFMA).  Block (S, j)
// holds the 64 reflectors of sweeps 64S-1 .. 64S+62 at step j, a 128 x 64
// staircase V (column t nonzero in window rows t .. t+63) on the 128-row
// window that starts at row 64(S + j); blocks run S descending, j ascending,
// as above.  A block's operands are the
// nonzero 16 x 16 tiles of V^T (20: sweep tile mt meets window tiles mt ..
// mt+4) and of -V T (26: column t of V T reaches rows 0 .. t+63), 46 KiB in
// FP32 fragments, all held in registers while the block's products run.
constexpr int kB64 = 64;
constexpr int kBlockFrags64 = 46;

__host__ __device__ inline long wy64_sweep_blocks(long nsweeps) { return (nsweeps + kB64) / kB64; }

// V^T tiles of window row tile rt: sweep tiles vlo(rt) .. vhi(rt), stored
// from fragment voff(rt); -V T tiles of row tile rt: sweep tiles wlo(rt) .. 3,
// from fragment woff(rt)
// (closed forms, not recursion: the device code must fold them to constants)
__host__ __device__ constexpr int vlo64(int rt) { return rt > 4 ? rt - 4 : 0; }
__host__ __device__ constexpr int vhi64(int rt) { return rt < 3 ? rt : 3; }
__host__ __device__ constexpr int voff64(int rt) {
  return rt <= 4 ? rt * (rt + 1) / 2 : 10 + (rt - 4) * (13 - rt) / 2;  // 0 1 3 6 10 14 17 19
}
__host__ __device__ constexpr int wlo64(int rt) { return rt > 4 ? rt - 4 : 0; }
__host__ __device__ constexpr int woff64(int rt) {
  return rt <= 5 ? 20 + 4 * rt : 40 + (rt - 5) * (12 - rt) / 2;  // 20 24 .. 40 43 45
}
static_assert(voff64(1) == 1 && voff64(4) == 10 && voff64(5) == 14 && voff64(6) == 17 && voff64(7) == 19 &&
                  woff64(5) == 40 && woff64(6) == 43 && woff64(7) == 45,
              "fragment table");

// One workgroup per block (S, j): V (FP32, as stored), G = strict upper part
// of V^T V and T = (I + G)^-1 in FP64, then the fragments; -V T is formed per
// fragment element.  Fragment q of a block (f4 per lane, component i):
//   q < 20:  V^T tile (rt, mt): V[16 rt + 4 g + i][16 mt + m]
//   q >= 20: -V T tile (rt, st): -(V T)[16 rt + m][16 st + 4 g + i]
// (g = lane / 16, m = lane % 16), the operand layouts of apply_block64.
__global__ __launch_bounds__(256) void bcw64_prep_kernel(int n, const float *__restrict__ U, long ldu, int nsweeps,
                                                         f4 *__restrict__ F, long u_stride, long f_stride) {
  U += blockIdx.z * u_stride;
  F += blockIdx.z * f_stride;
  const long Jn = (n + kB64 - 1) / kB64;
  const int S = blockIdx.y, j = blockIdx.x;
  if (j >= wy_steps(Jn, S)) return;
  __shared__ float V[128][65];
  __shared__ double G[64][65];
  __shared__ double T[64][65];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < 128 * 64; idx += 256) {
    const int r = idx & 127, t = idx >> 7, i = r - t;
    const long sw = (long)S * kB64 - 1 + t;
    const long row = (long)S * kB64 + (long)j * kB64 + r;
    float v = 0.f;
    if (i >= 0 && i < kB64 && sw >= 0 && sw < nsweeps && row < n) v = U[row + sw * ldu];
    V[r][t] = v;
  }
  __syncthreads();
  // G[a][c] = v_a . v_c for a < c: the columns overlap in rows c .. a+63
  for (int idx = tid; idx < 64 * 64; idx += 256) {
    const int a = idx & 63, c = idx >> 6;
    double s = 0.0;
    if (a < c)
      for (int r = c; r < a + kB64; r++) s += (double)V[r][a] * (double)V[r][c];
    G[a][c] = s;
  }
  __syncthreads();
  // T = (I + G)^-1, column c by back substitution (one thread per column)
  if (tid < kB64) {
    const int c = tid;
    for (int i = 0; i < kB64; i++) T[i][c] = (i == c) ? 1.0 : 0.0;
    for (int i = c - 1; i >= 0; i--) {
      double s = 0.0;
      for (int q = i + 1; q <= c; q++) s += G[i][q] * T[q][c];
      T[i][c] = -s;
    }
  }
  __syncthreads();
  const long blk = wy_prefix(Jn, S) + j;
  f4 *out = F + blk * (long)(kBlockFrags64 * kFrag);
  for (int idx = tid; idx < kBlockFrags64 * kFrag; idx += 256) {
    const int q = idx >> 6, lane = idx & 63, g = lane >> 4, m = lane & 15;
    f4 v;
    if (q < 20) {
      int rt = 0;
      while (rt < 7 && voff64(rt + 1) <= q) rt++;
      const int mt = vlo64(rt) + q - voff64(rt);
#pragma unroll
      for (int i = 0; i < 4; i++) v[i] = V[16 * rt + 4 * g + i][16 * mt + m];
    } else {
      int rt = 0;
      while (rt < 7 && woff64(rt + 1) <= q) rt++;
      const int st = wlo64(rt) + q - woff64(rt);
      const int r = 16 * rt + m;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int t = 16 * st + 4 * g + i;
        double acc = 0.0;
        for (int p = 0; p <= t; p++) acc += (double)V[r][p] * T[p][t];
        v[i] = (float)-acc;
      }
    }
    out[idx] = v;
  }
}

// A workgroup of the b = 64 kernel owns kC64 columns: a half window is 64
// rows x kC64 columns, kNT64 column tiles of 16.
constexpr int kC64 = 64, kNT64 = kC64 / 16, kHalf64 = 4 * kNT64 * 64;  // f4 per half window

// Element (row r, column c) of a half window in accumulator order (tile
// kNT64 rt + nt, lane 16 g + c16, component i), as a float index.
__device__ __attribute__((always_inline)) inline int half64_index(int r, int c) {
  return (((r >> 4) * kNT64 + (c >> 4)) * 64 + ((r >> 2) & 3) * 16 + (c & 15)) * 4 + (r & 3);
}

__device__ __attribute__((always_inline)) inline void put_half64(f4 *slot, const f4 (&h)[4][kNT64], int lane) {
#pragma unroll
  for (int rt = 0; rt < 4; rt++)
#pragma unroll
    for (int nt = 0; nt < kNT64; nt++) slot[(rt * kNT64 + nt) * 64 + lane] = h[rt][nt];
}
__device__ __attribute__((always_inline)) inline void get_half64(f4 (&h)[4][kNT64], const f4 *slot, int lane) {
#pragma unroll
  for (int rt = 0; rt < 4; rt++)
#pragma unroll
    for (int nt = 0; nt < kNT64; nt++) h[rt][nt] = slot[(rt * kNT64 + nt) * 64 + lane];
}

// HBM traffic of the b = 64 kernel.  Half windows staged through LDS (wave
// 0's incoming rows, the last wave's finished rows) are spread over the 4
// waves: wave k moves columns kQ k .. kQ k + kQ - 1, lane = row, one
// instruction = one column of 64 rows.  Skipped rows and columns get an
// offset past the buffer (dropped stores, zero loads), so every step issues
// the same memory operations.  load_acc brings a wave's own first top half
// straight into its accumulator tiles.
struct HalfIO64 {
  static constexpr int kQ = kC64 / 4;
  __amdgpu_buffer_rsrc_t rs;
  int r, c;         // row of the lane, first column of the wave (within the workgroup's)
  unsigned cbytes;  // byte offset of the wave's first column
  int ncl;          // columns of this workgroup
  __device__ HalfIO64(float *Zc, long ldz, int n, int ncols_wg, int lane, int k)
      : rs(__builtin_amdgcn_make_buffer_rsrc(Zc, 0, (int)(((long)(ncols_wg - 1) * ldz + n) * 4), kBufWord3)),
        r(lane),
        c(kQ * k),
        cbytes((unsigned)(c * ldz * 4)),
        ncl(ncols_wg) {}
  __device__ __attribute__((always_inline)) inline unsigned off(bool on, long row0, int n, long ldz, int q) const {
    const long gr = row0 + r;
    return (on && gr < n && c + q < ncl) ? cbytes + (unsigned)((q * ldz + gr) * 4) : kOutside;
  }
  __device__ __attribute__((always_inline)) inline void load(float (&v)[kQ], bool on, long row0, int n,
                                                             long ldz) const {
#pragma unroll
    for (int q = 0; q < kQ; q++)
      v[q] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, off(on, row0, n, ldz, q), 0, 0));
  }
  __device__ __attribute__((always_inline)) inline void park(f4 *slot, const float (&v)[kQ]) const {
    float *s = reinterpret_cast<float *>(slot);
#pragma unroll
    for (int q = 0; q < kQ; q++) s[half64_index(r, c + q)] = v[q];
  }
  __device__ __attribute__((always_inline)) inline void store(const f4 *slot, bool on, long row0, int n,
                                                              long ldz) const {
    const float *s = reinterpret_cast<const float *>(slot);
#pragma unroll
    for (int q = 0; q < kQ; q++)
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, s[half64_index(r, c + q)]), rs,
                                            off(on, row0, n, ldz, q), 0, 0);
  }
  // rows row0 .. row0+63 of all the workgroup's columns in accumulator order:
  // tile (rt, nt), lane 16 g + c16, component i = row 16 rt + 4 g + i of
  // column 16 nt + c16
  __device__ __attribute__((always_inline)) inline void load_acc(f4 (&h)[4][kNT64], long row0, int n, long ldz,
                                                                 int lane) const {
    const int c16 = lane & 15, g = lane >> 4;
#pragma unroll
    for (int rt = 0; rt < 4; rt++)
#pragma unroll
      for (int nt = 0; nt < kNT64; nt++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const long gr = row0 + 16 * rt + 4 * g + i;
          const int col = 16 * nt + c16;
          const unsigned o = (gr < n && col < ncl) ? (unsigned)((col * ldz + gr) * 4) : kOutside;
          h[rt][nt][i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, o, 0, 0));
        }
  }
};

// fragments q0 .. q0 + N - 1 of a block into fr[0 .. N-1]
template <int N, int M>
__device__ __attribute__((always_inline)) inline void load_tiles(f4 (&fr)[M], const f4 *__restrict__ blk, int q0,
                                                                 int lane) {
#pragma unroll
  for (int q = 0; q < N; q++) fr[q] = blk[(q0 + q) * kFrag + lane];
}

// Zw <- (I - V T V^T) Zw for one block.  vt holds the block's 20 V^T tiles
// (loaded one block ahead).  The loads are placed so that few operand tiles
// are live at once: the -V T tiles of window row tiles 0..3 (16) load at
// the start, behind the X products, those of row tiles 4..7 (10) once X has
// consumed most of vt, and the next block's V^T tiles (nxt, may be null) once
// the first Zw half has consumed the first -V T part.  Every product loop runs
// the kNT64 = 4 column tiles as independent accumulators.
__device__ __attribute__((always_inline)) inline void apply_block64(f4 (&top)[4][kNT64], f4 (&bot)[4][kNT64],
                                                                    f4 (&vt)[20], const f4 *__restrict__ blk,
                                                                    const f4 *__restrict__ nxt, int lane) {
  f4 wa[16], wb[10];
  load_tiles<16>(wa, blk, woff64(0), lane);
  __builtin_amdgcn_sched_barrier(0);
  f4 x[4][kNT64];
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int nt = 0; nt < kNT64; nt++) x[mt][nt] = f4{0.f, 0.f, 0.f, 0.f};
  // X = V^T Zw, window row tile by row tile (every loop has a constant trip
  // count, so all indices fold after unrolling and the arrays stay in
  // registers)
#pragma unroll
  for (int rt = 0; rt < 8; rt++) {
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int nt = 0; nt < kNT64; nt++) {
          if (mt < vlo64(rt) || mt > vhi64(rt)) continue;
          x[mt][nt] = mma(vt[voff64(rt) + mt - vlo64(rt)][i], (rt < 4 ? top[rt & 3][nt] : bot[rt & 3][nt])[i],
                          x[mt][nt]);
        }
    if (rt == 4) {
      __builtin_amdgcn_sched_barrier(0);
      load_tiles<10>(wb, blk, woff64(4), lane);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  // Zw += (-V T) X, window row tile by row tile
#pragma unroll
  for (int rt = 0; rt < 8; rt++) {
    if (rt == 4) {
      __builtin_amdgcn_sched_barrier(0);
      if (nxt) load_tiles<20>(vt, nxt, 0, lane);
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int st = 0; st < 4; st++)
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int nt = 0; nt < kNT64; nt++) {
          if (st < wlo64(rt)) continue;
          const int q = woff64(rt) + st - wlo64(rt);
          const f4 &w = rt < 4 ? wa[(q - woff64(0)) & 15] : wb[(q - woff64(4)) % 10];
          if (rt < 4)
            top[rt & 3][nt] = mma(w[i], x[st][nt][i], top[rt & 3][nt]);
          else
            bot[rt & 3][nt] = mma(w[i], x[st][nt][i], bot[rt & 3][nt]);
        }
  }
}

// The b = 64 pipeline: as bcw_kernel (waves over consecutive sweep blocks,
// rows handed on through LDS), with kC64 = 64 columns per workgroup and 4
// waves, one per SIMD (a wave's window, X and a block's fragments take
// about 420 of its 512 registers).  Every workgroup streams all WY blocks,
// so the column width sets the fragment traffic: 64 columns read the 6.2 GB
// of blocks (n = 32768) 512 times, half what 32 columns did.  The LDS holds
// the hand-off slots and wave 0's incoming rows (160 KiB); each wave loads
// the top half of its first step of a group straight into its registers, two
// steps ahead.  Q is streamed from HBM once per 256 sweeps.
__global__ __launch_bounds__(256) void bcw64_kernel(int n, int ncols, float *__restrict__ Z, long ldz,
                                                    const f4 *__restrict__ F, int nsweeps, long z_stride,
                                                    long f_stride) {
  constexpr int NW = 4;
  constexpr int kQ = HalfIO64::kQ;
  Z += blockIdx.y * z_stride;
  F += blockIdx.y * f_stride;
  __shared__ f4 xch[NW * 2 * kHalf64];
  __shared__ f4 stage[2 * kHalf64];
  const int lane = threadIdx.x & 63;
  const int k = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int col0 = blockIdx.x * kC64;
  const int nblk = (int)wy64_sweep_blocks(nsweeps);
  const long Jn = (n + kB64 - 1) / kB64;
  const HalfIO64 io(Z + (long)col0 * ldz, ldz, n, min(kC64, ncols - col0), lane, k);
  f4 zt[4][kNT64], zb[4][kNT64];
  f4 vt[20];

  const int ngroups = (nblk + NW - 1) / NW;
  for (int gi = 0; gi < ngroups; gi++) {
    const int gtop = nblk - 1 - gi * NW;
    const int nact = min(NW, gtop + 1);
    const int bk = gtop - k;
    const int J = (int)wy_steps(Jn, gtop - nact + 1) + 1;
    const int T = J + nact - 1;
    const int kl = nact - 1;
    const long nst = wy_steps(Jn, bk);
    const f4 *fb = F + wy_prefix(Jn, bk) * (long)(kBlockFrags64 * kFrag);
    float lb[kQ];
    for (int tau = -2; tau <= T; tau++) {
      const int e = tau - k;
      const bool act = tau >= 0 && tau < T && k < nact && e >= 0 && e < J;
      const bool app = act && e < nst;
      const f4 *cur = fb + (long)e * (kBlockFrags64 * kFrag);
      if (app && e == 0) load_tiles<20>(vt, cur, 0, lane);
      // this wave's first top half, two steps ahead (no other wave of the
      // group touches those rows before then)
      if (tau == k - 2 && k < nact) io.load_acc(zt, (long)bk * kB64, n, ldz, lane);
      // wave 0's bottom half for step tau+1: in flight during this step's
      // products, parked at its end
      const bool stage_on = tau >= -1 && tau + 1 < J;
      io.load(lb, stage_on, (long)(gtop + tau + 2) * kB64, n, ldz);
      if (act) {
        if (e > 0) {
#pragma unroll
          for (int rt = 0; rt < 4; rt++)
#pragma unroll
            for (int nt = 0; nt < kNT64; nt++) zt[rt][nt] = zb[rt][nt];
        }
        get_half64(zb, k == 0 ? stage + (e & 1) * kHalf64 : xch + ((k - 1) * 2 + (e & 1)) * kHalf64, lane);
        if (app)
          apply_block64(zt, zb, vt, cur, (e + 1 < nst && e + 1 < J) ? cur + kBlockFrags64 * kFrag : nullptr, lane);
        put_half64(xch + (k * 2 + (e & 1)) * kHalf64, zt, lane);
      }
      // rows the last wave handed on at step tau-1 -> Q
      const int el = tau - 1 - kl;
      io.store(xch + (kl * 2 + (el & 1)) * kHalf64, el >= 0 && el < J, (long)(gtop - kl + el) * kB64, n, ldz);
      if (stage_on) io.park(stage + ((tau + 1) & 1) * kHalf64, lb);
      lds_barrier();
    }
    wait_mem();
    __syncthreads();
  }
}

}  // namespace

size_t bc_back_wy_bytes(long n) {
  if (n <= 2) return 0;
  return align256((size_t)wy_prefix((n + kB - 1) / kB, wy_sweep_blocks(n - 2)) * kBlockFrags * kFrag * sizeof(h8));
}

// the 64 columns of a workgroup must be addressable with 32-bit buffer offsets
bool bc_back_wy_fits(long n, long ldq) { return (63 * ldq + n) * 4 < (1L << 31); }

// Q <- H_1 .. H_K Q for b = 32 as WY blocks on the matrix cores: FP32 MFMA, or
// split FP16 (split16: columns of Q with 2-norm <= 8).  A batch of `count`
// same-size problems (Q + q q_stride, U + q u_stride) runs as one launch of
// each kernel.
int bc_back_wy_batched(evd_ctx *ctx, int count, long n, long ncols, float *Q, long ldq, long q_stride, const float *U,
                       long ldu, long u_stride, bool split16) {
  const int nsweeps = (int)std::max<long>(0, n - 2);
  if (nsweeps == 0 || ncols <= 0 || count <= 0) return 0;
  ArenaFrame frame(ctx);
  const long fbytes = (long)bc_back_wy_bytes(n);
  h8 *F = static_cast<h8 *>(arena_alloc(ctx, (size_t)fbytes * count));
  if (!F) return EVD_E_NOMEM;
  const long f_stride = fbytes / (long)sizeof(h8);
  const long Jn = (n + kB - 1) / kB, nblk = wy_sweep_blocks(nsweeps);
  const dim3 pg((unsigned)Jn, (unsigned)nblk, (unsigned)count);
  if (split16)
    bcw_prep_kernel<false><<<pg, 256, 0, ctx->stream>>>((int)n, U, ldu, nsweeps, F, ctx->ctrl + kCtrlError, u_stride,
                                                         f_stride);
  else
    bcw_prep_kernel<true><<<pg, 256, 0, ctx->stream>>>((int)n, U, ldu, nsweeps, F, ctx->ctrl + kCtrlError, u_stride,
                                                        f_stride);
  EVD_LAUNCHED(ctx);
  if (ctx->timing) hipEventRecord(ctx->kev[0], ctx->stream);
  constexpr int NW = 8;
  const dim3 grid((unsigned)((ncols + 63) / 64), (unsigned)count);
  if (split16)
    bcw_kernel<NW, false><<<grid, NW * 64, 0, ctx->stream>>>((int)n, (int)ncols, Q, ldq, F, nsweeps, q_stride,
                                                              f_stride);
  else
    bcw_kernel<NW, true><<<grid, NW * 64, 0, ctx->stream>>>((int)n, (int)ncols, Q, ldq, F, nsweeps, q_stride,
                                                             f_stride);
  EVD_LAUNCHED(ctx);
  if (ctx->timing) hipEventRecord(ctx->kev[1], ctx->stream);
  return 0;
}

size_t bc_back_wy64_bytes(long n) {
  if (n <= 2) return 0;
  return align256((size_t)wy_prefix((n + kB64 - 1) / kB64, wy64_sweep_blocks(n - 2)) * kBlockFrags64 * kFrag *
                  sizeof(f4));
}

// the columns of a workgroup must be addressable with 32-bit buffer offsets
bool bc_back_wy64_fits(long n, long ldq) { return ((kC64 - 1) * ldq + n) * 4 < (1L << 31); }

// Q <- H_1 .. H_K Q for b = 64 as WY blocks on FP32 MFMA (every precision
// form), a batch of `count` same-size problems per launch.
int bc_back_wy64_batched(evd_ctx *ctx, int count, long n, long ncols, float *Q, long ldq, long q_stride,
                         const float *U, long ldu, long u_stride) {
  const int nsweeps = (int)std::max<long>(0, n - 2);
  if (nsweeps == 0 || ncols <= 0 || count <= 0) return 0;
  ArenaFrame frame(ctx);
  const long fbytes = (long)bc_back_wy64_bytes(n);
  f4 *F = static_cast<f4 *>(arena_alloc(ctx, (size_t)fbytes * count));
  if (!F) return EVD_E_NOMEM;
  const long f_stride = fbytes / (long)sizeof(f4);
  const long Jn = (n + kB64 - 1) / kB64, nblk = wy64_sweep_blocks(nsweeps);
  bcw64_prep_kernel<<<dim3((unsigned)Jn, (unsigned)nblk, (unsigned)count), 256, 0, ctx->stream>>>(
      (int)n, U, ldu, nsweeps, F, u_stride, f_stride);
  EVD_LAUNCHED(ctx);
  if (ctx->timing) hipEventRecord(ctx->kev[0], ctx->stream);
  bcw64_kernel<<<dim3((unsigned)((ncols + kC64 - 1) / kC64), (unsigned)count), 256, 0, ctx->stream>>>(
      (int)n, (int)ncols, Q, ldq, F, nsweeps, q_stride, f_stride);
  EVD_LAUNCHED(ctx);
  if (ctx->timing) hipEventRecord(ctx->kev[1], ctx->stream);
  return 0;
}

int bc_back_wy(evd_ctx *ctx, long n, long ncols, float *Q, long ldq, const float *U, long ldu, bool split16) {
  return bc_back_wy_batched(ctx, 1, n, ncols, Q, ldq, 0, U, ldu, 0, split16);
}

}  // namespace evd
