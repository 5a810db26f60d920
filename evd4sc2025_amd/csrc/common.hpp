// Shared host/device infrastructure of the eigensolver library: the context
// (stream, zero-initialised arena, error text, timing events, tuning), status
// helpers and small device utilities.  Everything is gfx950 / wave64 only.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

#include "../../include/evd4sc.h"

typedef struct _rocblas_handle *rocblas_handle;

namespace evd {

constexpr int kWave = 64;

struct Tuning {
  int bc_sweeps = 0;       // sweeps aggregated per BC-back block (0 = b)
  int bc_waves = 0;        // pipelined waves per BC-back workgroup (0 = 4)
  int chase_groups = 0;    // persistent workgroups of the bulge chase (0 = auto)
  int sbr_back_block = 0;  // WY block width of the SBR back transformation (0 = 1024)
  int bc_valu = 0;         // 1: BC back transformation on the FP32 VALU kernel only
  int skinny_split = 0;    // split-K target (workgroups per CU) of long-K skinny products (0 = 4)
  int chase_depth = 0;     // sweeps per workgroup of the bulge chase (0 = 1: one sweep per workgroup)
  int panel_rows = 0;      // panel QR rows per workgroup for panels <= 32 wide (0 = 512; 256)
  int gemm_tile = 0;       // output tile of the large FP32 products (0 = auto, 1 = 128 x 128, 2 = 64 x 64)
  int own_gemm = 0;        // 1: large plain FP32 products on this library's MFMA kernels, not rocBLAS
  int panel_exchange = 0;  // 1: panels on the column-exchange kernel only (no Cholesky QR path)
};

// Merge plan of the tridiagonal divide and conquer for one order n (host
// level table + device copy of the merge descriptors, split flags and tile
// offsets), kept per context.
struct DcPlan {
  int nmerge = 0, H = 0;
  std::vector<int> level_off;    // merges of height h: [level_off[h], level_off[h + 1])
  std::vector<int> level_maxN;   // largest merge order of each height
  std::vector<int> level_tiles;  // grouped-GEMM tile budget of each height
  void *dev = nullptr;
  bool own = false;              // dev is its own allocation (not carved from the plan pool)
  std::vector<char> host;        // source of the stream-ordered upload, kept alive with the plan
};

}  // namespace evd

struct evd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int num_cus = 0;

  // arena: one device allocation, bump-allocated inside a top-level call
  char *arena = nullptr;
  size_t arena_bytes = 0;
  bool arena_external = false;  // the arena is lent by the owner (serving pool): never grown or freed here
  size_t arena_top = 0;
  size_t budget = 0;  // 0 = unlimited
  int depth = 0;      // nesting depth of public entry points

  // small always-resident control block: barrier words, error word, flags
  unsigned *ctrl = nullptr;  // kCtrlWords words, zeroed at creation
  size_t ctrl_words = 0;

  unsigned *tags = nullptr;  // kTagWords persistent exchange tags + kCqCountWords counters, zeroed at creation
  evd::Tuning tune;
  bool tune_cache = true;  // fill knobs left at 0 from the process-wide tuning cache
  int precision = EVD_PREC_FP32;

  std::map<long, evd::DcPlan> dc_plans;  // tridiagonal merge plans by order | (batch count << 32)
  // device pool the plans are carved from (allocated once: the first plan,
  // or evd::plan_pool_reserve), so a new order costs no device allocation
  char *plan_pool = nullptr;
  size_t plan_pool_bytes = 0, plan_pool_used = 0;
  bool capturing = false;                 // inside hipGraph capture: no host round trips
  int graphs = 0;                         // live captured graphs: the arena may not move
  long shape_div = 1;                     // GEMM launches carry this many problems' products (batched SBR)
  // host array (one entry per problem of the next batched solve, or null):
  // the input screen takes max |a_ij| over the leading amax_rows[q] rows and
  // columns only (mixed-size batches: the pad diagonal is left out)
  const long *amax_rows = nullptr;

  // library GEMM (rocBLAS) for the large plain FP32 products: handle on the
  // context stream, deterministic (no atomics), with its own workspace so a
  // call never allocates (also under capture)
  rocblas_handle blas = nullptr;
  void *blas_ws = nullptr;

  bool timing = false;
  hipEvent_t ev[EVD_STAGE_COUNT + 1] = {};
  hipEvent_t kev[2] = {};  // around the BC back-transformation kernel
  hipEvent_t cev[2] = {};  // around the bulge-chase kernel
  float stage_ms[EVD_TIMER_COUNT] = {};

  char err[1024] = {0};
};

namespace evd {

// control block layout (words)
enum : int {
  kCtrlError = 0,      // device-side error word (EVD_E_*), sticky until read
  kCtrlEpoch = 8,      // launch epoch of the panel factorisation's exchange
  kCtrlBarrier = 16,   // grid barrier: [count, generation] pairs, 64 slots
  kCtrlWordsTotal = 16 + 2 * 64,
  kTagWords = 4096,
  kCqCountWords = 4096,  // reduction counters of the Cholesky QR panel, after the tags
};

int set_error(evd_ctx *ctx, int code, const char *fmt, ...);

#define EVD_HIP(ctx, call)                                                                          \
  do {                                                                                              \
    hipError_t e_ = (call);                                                                         \
    if (e_ != hipSuccess)                                                                           \
      return ::evd::set_error((ctx), EVD_E_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                              __FILE__, __LINE__);                                                  \
  } while (0)

#define EVD_TRY(expr)        \
  do {                       \
    int s_ = (expr);         \
    if (s_ != 0) return s_;  \
  } while (0)

#define EVD_ARG(ctx, cond, k)                                                                   \
  do {                                                                                          \
    if (!(cond)) return ::evd::set_error((ctx), -(k), "argument %d invalid: %s", (k), #cond); \
  } while (0)

// Launch check: kernel launch errors surface through hipGetLastError.
#define EVD_LAUNCHED(ctx) EVD_HIP(ctx, hipGetLastError())

// RAII frame over the arena: sub-allocations made while the frame lives are
// released when it ends.
struct ArenaFrame {
  evd_ctx *ctx;
  size_t mark;
  explicit ArenaFrame(evd_ctx *c) : ctx(c), mark(c->arena_top) {}
  ~ArenaFrame() { ctx->arena_top = mark; }
};

// Ensure the arena holds at least `bytes` beyond the current top.  Growing
// reallocates, so it is only allowed when nothing is allocated (top == 0).
int arena_ensure(evd_ctx *ctx, size_t bytes);
// Bump allocation (256-byte aligned).  Returns nullptr and sets the error if
// the arena is too small.
void *arena_alloc(evd_ctx *ctx, size_t bytes);
template <typename T>
T *arena_alloc_n(evd_ctx *ctx, size_t n) {
  return static_cast<T *>(arena_alloc(ctx, n * sizeof(T)));
}

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// fills (stream-ordered kernels; never hipMemsetAsync so calls stay
// capturable in hipGraphs)
int fill_f32(evd_ctx *ctx, float *p, long rows, long cols, long ld, float v);
int fill_f64(evd_ctx *ctx, double *p, long rows, long cols, long ld, double v);
int fill_u32(evd_ctx *ctx, unsigned *p, size_t n, unsigned v);
int set_identity_f64(evd_ctx *ctx, double *p, long n, long ld);
int copy_f32(evd_ctx *ctx, const float *src, long lds, float *dst, long ldd, long rows, long cols);

// GEMM entry points (gemm.hip)
int sgemm(evd_ctx *ctx, char ta, char tb, long m, long n, long k, float alpha, const float *A, long lda,
          const float *B, long ldb, float beta, float *C, long ldc, long batch = 1, long sA = 0, long sB = 0,
          long sC = 0);
int sgemm_kcat(evd_ctx *ctx, char ta, char tb, long m, long n, long k1, long k2, float alpha, const float *A1,
               long lda1, const float *B1, long ldb1, const float *A2, long lda2, const float *B2, long ldb2, float beta,
               float *C, long ldc, long batch = 1, long sA1 = 0, long sB1 = 0, long sA2 = 0, long sB2 = 0,
               long sC = 0);
int dgemm(evd_ctx *ctx, char ta, char tb, long m, long n, long k, double alpha, const double *A, long lda,
          const double *B, long ldb, double beta, double *C, long ldc);
int syr2k_lower_mirror(evd_ctx *ctx, long n, long k, float alpha, const float *A, long lda, const float *B, long ldb,
                       float beta, float *C, long ldc);
int syr2k_lower_mirror_batched(evd_ctx *ctx, int count, long n, long k, float alpha, const float *A, long lda,
                               long sA, const float *B, long ldb, long sB, float beta, float *C, long ldc, long sC);

// grouped FP64 GEMM: C[:, colmap[j]] = A (m x k) * B (k x n) per problem
struct DGemmProblem {
  int m, n, k;
  int tile_begin;  // prefix of tiles
  const double *A;
  long lda;
  const double *B;
  long ldb;
  double *C;
  long ldc;
  const int *colmap;  // nullable: output column j goes to C column colmap[j]
};
int dgemm_grouped(evd_ctx *ctx, const DGemmProblem *dev_problems, int nproblems, int total_tiles);
constexpr int kDGroupTile = 64;

int max_coresident_blocks(evd_ctx *ctx, const void *kernel, int block, size_t dyn_smem, int *out);

// Lend the context an arena it does not own (nullptr = none); frees an own
// arena first.  Calls that need more than `bytes` fail with EVD_E_NOMEM.
int ctx_lend_arena(evd_ctx *ctx, void *p, size_t bytes);

// the D&C plan pool of a context (ctx.hip): allocated once with `bytes`
int plan_pool_reserve(evd_ctx *ctx, size_t bytes);
constexpr size_t kPlanPoolBytes = (size_t)4 << 20;  // a plan is ~100 KB at n = 16384

// error word readback (synchronises the stream)
int check_device_error(evd_ctx *ctx, const char *where);

// arena bytes of a solve (syevd.hip); sbr_knob = the sbr_back_block knob the
// solve will run with (0 = default block)
size_t syevd_workspace(long n, long b, long nb, char jobz, int precision, long sbr_knob = 0);
size_t syevd_batched_workspace(int count, long n, long b, long nb, char jobz, int precision, long sbr_knob = 0);

// what syevd_call / syevd_batched_call will reserve on this context
size_t syevd_call_workspace(const evd_ctx *ctx, long n, long b, long nb, char jobz);
size_t syevd_batched_call_workspace(const evd_ctx *ctx, int count, long n, long b, long nb, char jobz);

// evd_syevd_f32 after its argument checks (syevd.hip)
int syevd_call(evd_ctx *ctx, char jobz, long n, long b, long nb, float *A, long lda, double *w, float *Z, long ldz);
int syevd_batched_call(evd_ctx *ctx, char jobz, int count, long n, long b, long nb, float *A, long lda, long a_stride,
                       double *w, long w_stride, float *Z, long ldz, long z_stride);

// evd_syevd_mixed_f32's arguments (host arrays of per-problem values)
struct MixedArgs {
  int count;
  const long *n;
  float *const *A;
  const long *lda;
  double *const *w;
  float *const *Z;
  const long *ldz;
  const char *jobz;
  long b;
};
// mixed.hip: argument checks; bytes of the per-group device descriptors and
// their upload (synchronises the stream); arena bytes beyond the descriptors;
// the solve of every group over uploaded descriptors (enqueues only, so it can
// be captured)
int mixed_check(evd_ctx *ctx, const MixedArgs &a);
size_t mixed_desc_bytes(const MixedArgs &a);
int mixed_upload(evd_ctx *ctx, const MixedArgs &a, char *dev);
size_t mixed_stage_bytes(const evd_ctx *ctx, const MixedArgs &a);
int mixed_enqueue(evd_ctx *ctx, const MixedArgs &a, const char *dev);

}  // namespace evd

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
namespace evd {

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ inline unsigned ld_acquire(const unsigned *p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline int ld_acquire(const int *p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_release(int *p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Device-coherent (agent scope) relaxed accesses.  Data that one workgroup
// hands to another inside a persistent kernel goes through these instead of
// plain loads/stores + release/acquire fences: on a multi-XCD part an
// agent-scope fence writes back / invalidates the whole L2, while these
// accesses are coherent by themselves.  Ordering of such stores before a flag
// store is the writer's s_waitcnt (all stores acknowledged) + barrier.
template <typename T>
__device__ inline T ld_dev(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ inline void st_dev(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void wait_mem() { __builtin_amdgcn_s_waitcnt(0); }

// 100 MHz constant clock -> bounded waits (watchdog) in persistent kernels
__device__ inline unsigned long long clock100() { return wall_clock64(); }
constexpr unsigned long long kWatchdogTicks = 100ull * 1000ull * 1000ull * 20ull;  // 20 s

// Generation barrier over `nblocks` co-resident workgroups for kernels that
// exchange data only through ld_dev/st_dev.  words[0] = arrival count
// (returns to 0), words[1] = generation.  Every thread of every block must
// call it.  Returns false if the watchdog expired (the caller must then leave
// the kernel; the other blocks time out the same way).
__device__ inline bool grid_barrier(unsigned *words, unsigned nblocks, unsigned *err) {
  wait_mem();
  __syncthreads();
  __shared__ int s_ok;
  if (threadIdx.x == 0) {
    s_ok = 1;
    const unsigned gen = ld_dev(&words[1]);
    wait_mem();
    const unsigned arrived = __hip_atomic_fetch_add(&words[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1;
    if (arrived == nblocks) {
      st_dev(&words[0], 0u);
      wait_mem();
      __hip_atomic_fetch_add(&words[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      const unsigned long long t0 = clock100();
      while (ld_dev(&words[1]) == gen) {
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
}

}  // namespace evd
