// Context, arena and small stream-ordered utility kernels.
#include <map>
#include <mutex>
#include <tuple>

#include <rocblas/rocblas.h>

#include "common.hpp"

namespace evd {

int set_error(evd_ctx *ctx, int code, const char *fmt, ...) {
  if (ctx) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(ctx->err, sizeof(ctx->err), fmt, ap);
    va_end(ap);
  }
  return code;
}

int arena_ensure(evd_ctx *ctx, size_t bytes) {
  size_t need = ctx->arena_top + bytes;
  if (need <= ctx->arena_bytes) return 0;
  if (ctx->arena_top != 0)
    return set_error(ctx, EVD_E_NOMEM, "arena too small inside a call (need %zu, have %zu); reserve first", need,
                     ctx->arena_bytes);
  if (ctx->graphs > 0)
    return set_error(ctx, EVD_E_UNSUPPORTED,
                     "arena of %zu bytes cannot grow to %zu while captured graphs use it; destroy them first",
                     ctx->arena_bytes, need);
  if (ctx->arena_external)
    return set_error(ctx, EVD_E_NOMEM, "lent arena of %zu bytes is smaller than the %zu bytes the call needs",
                     ctx->arena_bytes, need);
  if (ctx->budget && need > ctx->budget)
    return set_error(ctx, EVD_E_NOMEM, "arena of %zu bytes exceeds the budget of %zu bytes", need, ctx->budget);
  if (ctx->arena) {
    EVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    EVD_HIP(ctx, hipFree(ctx->arena));
    ctx->arena = nullptr;
    ctx->arena_bytes = 0;
  }
  size_t sz = align256(need) + (1u << 20);
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, sz);
  if (e != hipSuccess) return set_error(ctx, EVD_E_NOMEM, "hipMalloc(%zu) failed: %s", sz, hipGetErrorString(e));
  ctx->arena = static_cast<char *>(p);
  ctx->arena_bytes = sz;
  return 0;
}

void *arena_alloc(evd_ctx *ctx, size_t bytes) {
  size_t off = align256(ctx->arena_top);
  if (off + bytes > ctx->arena_bytes) {
    set_error(ctx, EVD_E_NOMEM, "arena exhausted: need %zu more bytes at offset %zu of %zu", bytes, off,
              ctx->arena_bytes);
    return nullptr;
  }
  ctx->arena_top = off + bytes;
  return ctx->arena + off;
}

// ------------------------------------------------------------------ fills
template <typename T>
__global__ void fill2d_kernel(T *p, long rows, long cols, long ld, T v) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  long c = blockIdx.y;
  for (; c < cols; c += gridDim.y)
    for (long r = i; r < rows; r += (long)gridDim.x * blockDim.x) p[r + c * ld] = v;
}

template <typename T>
static int fill2d(evd_ctx *ctx, T *p, long rows, long cols, long ld, T v) {
  if (rows <= 0 || cols <= 0) return 0;
  unsigned gx = (unsigned)((rows + 255) / 256);
  if (gx > 64) gx = 64;
  unsigned gy = (unsigned)(cols < 4096 ? cols : 4096);
  fill2d_kernel<T><<<dim3(gx, gy), 256, 0, ctx->stream>>>(p, rows, cols, ld, v);
  EVD_LAUNCHED(ctx);
  return 0;
}

int fill_f32(evd_ctx *ctx, float *p, long rows, long cols, long ld, float v) {
  return fill2d<float>(ctx, p, rows, cols, ld, v);
}
int fill_f64(evd_ctx *ctx, double *p, long rows, long cols, long ld, double v) {
  return fill2d<double>(ctx, p, rows, cols, ld, v);
}

__global__ void fill_u32_kernel(unsigned *p, size_t n, unsigned v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
int fill_u32(evd_ctx *ctx, unsigned *p, size_t n, unsigned v) {
  if (n == 0) return 0;
  unsigned g = (unsigned)((n + 255) / 256);
  if (g > 1024) g = 1024;
  fill_u32_kernel<<<g, 256, 0, ctx->stream>>>(p, n, v);
  EVD_LAUNCHED(ctx);
  return 0;
}

__global__ void set_diag_f64_kernel(double *p, long n, long ld) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i + i * ld] = 1.0;
}
int set_identity_f64(evd_ctx *ctx, double *p, long n, long ld) {
  EVD_TRY(fill_f64(ctx, p, n, n, ld, 0.0));
  set_diag_f64_kernel<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(p, n, ld);
  EVD_LAUNCHED(ctx);
  return 0;
}

__global__ void copy2d_f32_kernel(const float *s, long lds, float *d, long ldd, long rows, long cols) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  for (long c = blockIdx.y; c < cols; c += gridDim.y)
    for (long r = i; r < rows; r += (long)gridDim.x * blockDim.x) d[r + c * ldd] = s[r + c * lds];
}
int copy_f32(evd_ctx *ctx, const float *src, long lds, float *dst, long ldd, long rows, long cols) {
  if (rows <= 0 || cols <= 0) return 0;
  unsigned gx = (unsigned)((rows + 255) / 256);
  if (gx > 64) gx = 64;
  unsigned gy = (unsigned)(cols < 4096 ? cols : 4096);
  copy2d_f32_kernel<<<dim3(gx, gy), 256, 0, ctx->stream>>>(src, lds, dst, ldd, rows, cols);
  EVD_LAUNCHED(ctx);
  return 0;
}

int max_coresident_blocks(evd_ctx *ctx, const void *kernel, int block, size_t dyn_smem, int *out) {
  // per (kernel, block, LDS, device): a property of the code object
  static std::mutex mu;
  static std::map<std::tuple<const void *, int, size_t, int>, int> known;
  const auto key = std::make_tuple(kernel, block, dyn_smem, ctx->device);
  int per_cu = -1;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = known.find(key);
    if (it != known.end()) per_cu = it->second;
  }
  if (per_cu < 0) {
    EVD_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, dyn_smem));
    std::lock_guard<std::mutex> lk(mu);
    known[key] = per_cu;
  }
  *out = per_cu * ctx->num_cus;
  if (*out <= 0) return set_error(ctx, EVD_E_UNSUPPORTED, "kernel cannot be resident (block %d)", block);
  return 0;
}

int ctx_lend_arena(evd_ctx *ctx, void *p, size_t bytes) {
  if (ctx->arena_top != 0) return set_error(ctx, EVD_E_UNSUPPORTED, "arena swap inside a call");
  if (ctx->arena && !ctx->arena_external) {
    EVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    EVD_HIP(ctx, hipFree(ctx->arena));
  }
  ctx->arena = static_cast<char *>(p);
  ctx->arena_bytes = p ? bytes : 0;
  ctx->arena_external = p != nullptr;
  return 0;
}

int plan_pool_reserve(evd_ctx *ctx, size_t bytes) {
  if (ctx->plan_pool) return 0;
  void *p = nullptr;
  const hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) return set_error(ctx, EVD_E_NOMEM, "plan pool of %zu bytes: %s", bytes, hipGetErrorString(e));
  ctx->plan_pool = static_cast<char *>(p);
  ctx->plan_pool_bytes = bytes;
  ctx->plan_pool_used = 0;
  return 0;
}

int check_device_error(evd_ctx *ctx, const char *where) {
  unsigned h = 0;
  EVD_HIP(ctx, hipMemcpyAsync(&h, ctx->ctrl + kCtrlError, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
  EVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (h) {
    unsigned z = 0;
    EVD_HIP(ctx, hipMemcpyAsync(ctx->ctrl + kCtrlError, &z, sizeof(unsigned), hipMemcpyHostToDevice, ctx->stream));
    EVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int code = (h & EVD_E_TIMEOUT) ? EVD_E_TIMEOUT : (h & EVD_E_NONFINITE) ? EVD_E_NONFINITE : (int)h;
    return set_error(ctx, code, "device reported error word 0x%x in %s", h, where);
  }
  return 0;
}

}  // namespace evd

using namespace evd;

// rocBLAS handle of a context (the large plain FP32 products): bound to the
// context stream, atomics off (deterministic results: a product of one shape
// always runs the same kernel the same way), host-side scalars and a fixed
// device workspace so that no call allocates.
static constexpr size_t kBlasWorkspace = (size_t)32 << 20;
static int blas_open(evd_ctx *ctx) {
  if (rocblas_create_handle(&ctx->blas) != rocblas_status_success) {
    ctx->blas = nullptr;
    return -1;
  }
  if (hipMalloc(&ctx->blas_ws, kBlasWorkspace) != hipSuccess) return -1;
  if (rocblas_set_stream(ctx->blas, ctx->stream) != rocblas_status_success ||
      rocblas_set_atomics_mode(ctx->blas, rocblas_atomics_not_allowed) != rocblas_status_success ||
      rocblas_set_pointer_mode(ctx->blas, rocblas_pointer_mode_host) != rocblas_status_success ||
      rocblas_set_workspace(ctx->blas, ctx->blas_ws, kBlasWorkspace) != rocblas_status_success)
    return -1;
  return 0;
}
static void blas_close(evd_ctx *ctx) {
  if (ctx->blas) rocblas_destroy_handle(ctx->blas);
  ctx->blas = nullptr;
  if (ctx->blas_ws) hipFree(ctx->blas_ws);
  ctx->blas_ws = nullptr;
}

extern "C" {

int evd_ctx_create(int device, void *stream, evd_ctx **out) {
  if (!out) return -3;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return EVD_E_HIP;
  if (device < 0 || device >= ndev) return -1;
  evd_ctx *ctx = new evd_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete ctx;
    return EVD_E_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    delete ctx;
    return EVD_E_HIP;
  }
  ctx->num_cus = prop.multiProcessorCount;
  if (stream) {
    ctx->stream = static_cast<hipStream_t>(stream);
  } else {
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
      delete ctx;
      return EVD_E_HIP;
    }
    ctx->own_stream = true;
  }
  ctx->ctrl_words = kCtrlWordsTotal;
  if (hipMalloc(&ctx->ctrl, ctx->ctrl_words * sizeof(unsigned)) != hipSuccess ||
      hipMemset(ctx->ctrl, 0, ctx->ctrl_words * sizeof(unsigned)) != hipSuccess ||
      hipMalloc(&ctx->tags, (kTagWords + kCqCountWords) * sizeof(unsigned)) != hipSuccess ||
      hipMemset(ctx->tags, 0, (kTagWords + kCqCountWords) * sizeof(unsigned)) != hipSuccess ||
      hipDeviceSynchronize() != hipSuccess) {
    delete ctx;
    return EVD_E_HIP;
  }
  for (auto &e : ctx->ev)
    if (hipEventCreate(&e) != hipSuccess) {
      delete ctx;
      return EVD_E_HIP;
    }
  for (auto &e : ctx->kev)
    if (hipEventCreate(&e) != hipSuccess) {
      delete ctx;
      return EVD_E_HIP;
    }
  for (auto &e : ctx->cev)
    if (hipEventCreate(&e) != hipSuccess) {
      delete ctx;
      return EVD_E_HIP;
    }
  if (blas_open(ctx) != 0) {
    evd_ctx_destroy(ctx);
    return EVD_E_HIP;
  }
  *out = ctx;
  return 0;
}

int evd_ctx_destroy(evd_ctx *ctx) {
  if (!ctx) return 0;
  if (ctx->graphs > 0) return set_error(ctx, EVD_E_UNSUPPORTED, "%d captured graphs still use this context", ctx->graphs);
  hipStreamSynchronize(ctx->stream);
  for (auto &kv : ctx->dc_plans)
    if (kv.second.dev && kv.second.own) hipFree(kv.second.dev);
  if (ctx->plan_pool) hipFree(ctx->plan_pool);
  if (ctx->arena && !ctx->arena_external) hipFree(ctx->arena);
  if (ctx->ctrl) hipFree(ctx->ctrl);
  if (ctx->tags) hipFree(ctx->tags);
  for (auto &e : ctx->ev)
    if (e) hipEventDestroy(e);
  for (auto &e : ctx->kev)
    if (e) hipEventDestroy(e);
  for (auto &e : ctx->cev)
    if (e) hipEventDestroy(e);
  blas_close(ctx);
  if (ctx->own_stream) hipStreamDestroy(ctx->stream);
  delete ctx;
  return 0;
}

int evd_ctx_set_stream(evd_ctx *ctx, void *stream) {
  if (!ctx) return -1;
  if (ctx->own_stream) {
    hipStreamSynchronize(ctx->stream);
    hipStreamDestroy(ctx->stream);
    ctx->own_stream = false;
  }
  ctx->stream = static_cast<hipStream_t>(stream);
  if (ctx->blas) rocblas_set_stream(ctx->blas, ctx->stream);
  return 0;
}

void *evd_ctx_get_stream(evd_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

const char *evd_ctx_last_error(const evd_ctx *ctx) { return ctx ? ctx->err : "null context"; }

int evd_ctx_reserve(evd_ctx *ctx, size_t bytes) {
  if (!ctx) return -1;
  if (ctx->arena_top != 0) return set_error(ctx, EVD_E_UNSUPPORTED, "reserve inside a call");
  return arena_ensure(ctx, bytes);
}

size_t evd_ctx_arena_bytes(const evd_ctx *ctx) { return ctx ? ctx->arena_bytes : 0; }

int evd_ctx_set_budget(evd_ctx *ctx, size_t bytes) {
  if (!ctx) return -1;
  ctx->budget = bytes;
  return 0;
}

int evd_ctx_set_tuning(evd_ctx *ctx, int bc_sweeps, int bc_waves, int chase_groups, int sbr_back_block) {
  if (!ctx) return -1;
  if (bc_sweeps < 0) return -2;
  if (bc_waves < 0 || bc_waves > 16) return -3;
  if (chase_groups < 0) return -4;
  if (sbr_back_block < 0) return -5;
  ctx->tune.bc_sweeps = bc_sweeps;
  ctx->tune.bc_waves = bc_waves;
  ctx->tune.chase_groups = chase_groups;
  ctx->tune.sbr_back_block = sbr_back_block;
  return 0;
}

int evd_ctx_set_knob(evd_ctx *ctx, const char *name, int value) {
  if (!ctx) return -1;
  if (!name) return -2;
  struct {
    const char *name;
    int *slot;
    int lo, hi;
  } knobs[] = {{"bc_sweeps", &ctx->tune.bc_sweeps, 0, 64},          {"bc_waves", &ctx->tune.bc_waves, 0, 16},
               {"chase_groups", &ctx->tune.chase_groups, 0, 1 << 20}, {"sbr_back_block", &ctx->tune.sbr_back_block, 0, 1 << 16},
               {"bc_valu", &ctx->tune.bc_valu, 0, 1},           {"skinny_split", &ctx->tune.skinny_split, 0, 4},
               {"chase_depth", &ctx->tune.chase_depth, 0, 6},   {"panel_rows", &ctx->tune.panel_rows, 0, 512},
               {"gemm_tile", &ctx->tune.gemm_tile, 0, 2},       {"own_gemm", &ctx->tune.own_gemm, 0, 1},
               {"panel_exchange", &ctx->tune.panel_exchange, 0, 1}};
  for (auto &kb : knobs)
    if (!strcmp(kb.name, name)) {
      if (value < kb.lo || value > kb.hi) return -3;
      if (kb.slot == &ctx->tune.panel_rows && value != 0 && value != 256 && value != 512) return -3;
      *kb.slot = value;
      return 0;
    }
  return -2;
}

int evd_ctx_use_tune_cache(evd_ctx *ctx, int on) {
  if (!ctx) return -1;
  ctx->tune_cache = on != 0;
  return 0;
}

int evd_ctx_set_precision(evd_ctx *ctx, int precision) {
  if (!ctx) return -1;
  if (precision < EVD_PREC_FP32 || precision > EVD_PREC_FP16) return -2;
  ctx->precision = precision;
  return 0;
}

int evd_ctx_enable_timing(evd_ctx *ctx, int on) {
  if (!ctx) return -1;
  ctx->timing = on != 0;
  return 0;
}

int evd_ctx_stage_ms(const evd_ctx *ctx, float *ms, int count) {
  if (!ctx) return -1;
  if (!ms) return -2;
  for (int i = 0; i < count && i < EVD_TIMER_COUNT; i++) ms[i] = ctx->stage_ms[i];
  return 0;
}

int evd_ctx_poison_arena(evd_ctx *ctx, unsigned byte) {
  if (!ctx) return -1;
  if (ctx->arena_top != 0) return set_error(ctx, EVD_E_UNSUPPORTED, "poison inside a call");
  if (!ctx->arena || ctx->arena_bytes == 0) return 0;
  const size_t words = ctx->arena_bytes / sizeof(unsigned);
  const unsigned b = byte & 0xFFu;
  EVD_TRY(fill_u32(ctx, reinterpret_cast<unsigned *>(ctx->arena), words, b | (b << 8) | (b << 16) | (b << 24)));
  return 0;
}

int evd_ctx_synchronize(evd_ctx *ctx) {
  if (!ctx) return -1;
  EVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return check_device_error(ctx, "evd_ctx_synchronize");
}

}  // extern "C"
