// Autotuner and tuning cache.  The reference fixes its tile shapes at compile
// time (template arguments such as chasing_kernel_one_timeV10<32>,
// util/inc/bugle_chasingV10.h:21, and the nb argument of myEVD,
// src/EVD/myEVD.cu:251); here the stage knobs are chosen per problem class:
//
//   nb              SBR block width (lazy trailing-update width), the panel
//                   width of the two-sided updates
//   chase_groups    persistent workgroups of the bulge chase (sweep depth in
//                   flight)
//   bc_valu         BC back transformation as WY blocks on the matrix cores
//                   (0) or on the FP32 VALU kernel (1)
//   sbr_back_block  WY block width of the SBR back transformation (MFMA tile
//                   of its GEMMs along k)
//   chase_depth     sweeps per workgroup of the bulge chase (sweep depth)
//   panel_rows      panel QR rows per workgroup (the panel's width in
//                   workgroups: 512 or 256 rows each)
//   gemm_tile       output tile of the large FP32 MFMA products (128 x 128 or
//                   64 x 64)
//   bc_sweeps       sweeps per block of the VALU BC back transformation
//   bc_waves        pipelined waves (sweep blocks in flight) of that kernel
//
// evd_autotune runs the eigensolver on a synthetic matrix of the requested
// order for candidate values, one knob at a time (coordinate descent, each
// knob judged by the stage it drives), and records the fastest setting under
// (n bucket = next power of two, b, precision).  evd_syevd_f32 fills every
// knob the caller left at 0 from that entry; the cache is process-wide and
// saved / loaded as text, one entry per line:
//   <n_bucket> <b> <precision> <nb> <chase_groups> <bc_valu> <sbr_back_block>
//   <chase_depth> <panel_rows> <gemm_tile> <bc_sweeps> <bc_waves>
// (lines with only the first seven fields load with the others at 0).
#include <map>
#include <mutex>
#include <tuple>

#include "common.hpp"

namespace evd {

namespace {

typedef std::tuple<long, long, int> Key;

// the knob vector of the C ABI, in its order (EVD_TUNE_KNOBS entries)
struct Entry {
  int nb = 0, chase_groups = 0, bc_valu = 0, sbr_back_block = 0, chase_depth = 0, panel_rows = 0, gemm_tile = 0,
      bc_sweeps = 0, bc_waves = 0;
  void to(int *k) const {
    const int v[EVD_TUNE_KNOBS] = {nb, chase_groups, bc_valu, sbr_back_block, chase_depth, panel_rows, gemm_tile,
                                   bc_sweeps, bc_waves};
    for (int i = 0; i < EVD_TUNE_KNOBS; i++) k[i] = v[i];
  }
  void from(const int *k) {
    nb = k[0];
    chase_groups = k[1];
    bc_valu = k[2];
    sbr_back_block = k[3];
    chase_depth = k[4];
    panel_rows = k[5];
    gemm_tile = k[6];
    bc_sweeps = k[7];
    bc_waves = k[8];
  }
  // the ranges evd_ctx_set_knob accepts; nb = 0 or a multiple of b
  bool valid(long b) const {
    return nb >= 0 && (nb == 0 || (nb >= b && nb % b == 0)) && chase_groups >= 0 && chase_groups <= (1 << 20) &&
           (bc_valu == 0 || bc_valu == 1) && sbr_back_block >= 0 && sbr_back_block <= (1 << 16) &&
           chase_depth >= 0 && chase_depth <= 6 && (panel_rows == 0 || panel_rows == 256 || panel_rows == 512) &&
           gemm_tile >= 0 && gemm_tile <= 2 && bc_sweeps >= 0 && bc_sweeps <= 64 && bc_waves >= 0 && bc_waves <= 16;
  }
};

std::mutex g_mu;
std::map<Key, Entry> g_cache;

long bucket(long n) {
  long p = 1;
  while (p < n) p <<= 1;
  return p;
}

}  // namespace

// knobs of (n, b, precision) from the cache, or false
bool tune_lookup(long n, long b, int precision, int *nb, Tuning *t) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_cache.find(Key(bucket(n), b, precision));
  if (it == g_cache.end()) return false;
  const Entry &e = it->second;
  if (nb) *nb = e.nb;
  if (t) {
    t->chase_groups = e.chase_groups;
    t->bc_valu = e.bc_valu;
    t->sbr_back_block = e.sbr_back_block;
    t->chase_depth = e.chase_depth;
    t->panel_rows = e.panel_rows;
    t->gemm_tile = e.gemm_tile;
    t->bc_sweeps = e.bc_sweeps;
    t->bc_waves = e.bc_waves;
  }
  return true;
}

}  // namespace evd

using namespace evd;

namespace {

// one timed solve: total ms and the stage times
int timed_solve(evd_ctx *ctx, long n, long b, long nb, const float *A0, float *A, double *w, float *Z, int reps,
                float *best_stage, float *best_total) {
  float st[EVD_TIMER_COUNT];
  *best_total = 1e30f;
  for (int s = 0; s < EVD_TIMER_COUNT; s++) best_stage[s] = 1e30f;
  for (int r = 0; r < reps; r++) {
    if (hipMemcpyAsync(A, A0, sizeof(float) * n * n, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
      return set_error(ctx, EVD_E_HIP, "autotune: copy failed");
    EVD_TRY(evd_syevd_f32(ctx, 'V', n, b, nb, A, n, w, Z, n));
    EVD_TRY(evd_ctx_stage_ms(ctx, st, EVD_TIMER_COUNT));
    float tot = 0.f;
    for (int s = 0; s < EVD_STAGE_COUNT; s++) tot += st[s];
    if (tot < *best_total) {
      *best_total = tot;
      for (int s = 0; s < EVD_TIMER_COUNT; s++) best_stage[s] = st[s];
    }
  }
  return 0;
}

}  // namespace

extern "C" int evd_autotune(evd_ctx *ctx, long n, long b, int reps, int *out_knobs) {
  if (!ctx) return -1;
  EVD_ARG(ctx, n >= 64, 2);
  EVD_ARG(ctx, b == 32 || b == 64, 3);
  EVD_ARG(ctx, reps >= 1, 4);
  const Tuning saved = ctx->tune;
  const bool saved_timing = ctx->timing, saved_cache = ctx->tune_cache;
  ctx->tune_cache = false;
  float *A0 = nullptr, *A = nullptr, *Z = nullptr;
  double *w = nullptr;
  int status = 0;
  Entry best;
  if (hipMalloc(&A0, sizeof(float) * n * n) != hipSuccess || hipMalloc(&A, sizeof(float) * n * n) != hipSuccess ||
      hipMalloc(&Z, sizeof(float) * n * n) != hipSuccess || hipMalloc(&w, sizeof(double) * n) != hipSuccess) {
    status = set_error(ctx, EVD_E_NOMEM, "autotune: device buffers for n = %ld", n);
  }
  if (status == 0) status = evd_gen_symmetric_f32(ctx, n, 7, 0, A0, n);
  if (status == 0) {
    ctx->timing = true;
    ctx->tune = Tuning();
    float stage[EVD_TIMER_COUNT], total;
    // each knob: candidates and the stage(s) it is judged by
    struct Knob {
      int *slot;
      std::vector<int> cand;
      int s0, s1;  // stage range [s0, s1) (EVD_STAGE_*), -1 = total
    };
    int nb = 1024 < n ? 1024 : (int)((n + b - 1) / b * b);
    std::vector<int> nbs;
    for (int v : {256, 512, 1024, 2048})
      if (v >= b && v <= std::max<long>(256, n)) nbs.push_back(v);
    Knob knobs[] = {
        {&nb, nbs, -1, -1},
        {&ctx->tune.panel_rows, {0, 256}, EVD_STAGE_SBR, EVD_STAGE_SBR + 1},
        {&ctx->tune.gemm_tile, {0, 1, 2}, -1, -1},
        {&ctx->tune.chase_depth, {0, 2, 3}, EVD_STAGE_BC, EVD_STAGE_BC + 1},
        {&ctx->tune.chase_groups, {0, (int)(n / 128) + 4, (int)(3 * n / 256) + 4, (int)(n / 32) + 4}, EVD_STAGE_BC,
         EVD_STAGE_BC + 1},
        {&ctx->tune.bc_valu, {0, 1}, EVD_STAGE_BCBACK, EVD_STAGE_BCBACK + 1},
        {&ctx->tune.bc_sweeps, {0, 32}, EVD_STAGE_BCBACK, EVD_STAGE_BCBACK + 1},
        {&ctx->tune.bc_waves, {0, 4}, EVD_STAGE_BCBACK, EVD_STAGE_BCBACK + 1},
        {&ctx->tune.sbr_back_block, {0, 512, 2048}, EVD_STAGE_SBRBACK, EVD_STAGE_SBRBACK + 1},
    };
    for (Knob &kb : knobs) {
      if (kb.slot == &ctx->tune.chase_depth && b > 32) continue;  // the grouped chase needs b <= 32
      // the VALU BC back transformation's shape, when that kernel runs (its
      // b = 64 form has one shape)
      const bool valu_bc = ctx->tune.bc_valu == 1 && b == 32;
      if ((kb.slot == &ctx->tune.bc_sweeps || kb.slot == &ctx->tune.bc_waves) && !valu_bc) continue;
      int best_v = *kb.slot;
      float best_t = 1e30f;
      for (int v : kb.cand) {
        *kb.slot = v;
        status = timed_solve(ctx, n, b, nb, A0, A, w, Z, reps, stage, &total);
        if (status) break;
        float t = total;
        if (kb.s0 >= 0) {
          t = 0.f;
          for (int s = kb.s0; s < kb.s1; s++) t += stage[s];
        }
        if (t < best_t) {
          best_t = t;
          best_v = v;
        }
      }
      if (status) break;
      *kb.slot = best_v;
    }
    best.nb = nb;
    best.chase_groups = ctx->tune.chase_groups;
    best.bc_valu = ctx->tune.bc_valu;
    best.sbr_back_block = ctx->tune.sbr_back_block;
    best.chase_depth = ctx->tune.chase_depth;
    best.panel_rows = ctx->tune.panel_rows;
    best.gemm_tile = ctx->tune.gemm_tile;
    best.bc_sweeps = ctx->tune.bc_sweeps;
    best.bc_waves = ctx->tune.bc_waves;
  }
  hipStreamSynchronize(ctx->stream);
  if (A0) hipFree(A0);
  if (A) hipFree(A);
  if (Z) hipFree(Z);
  if (w) hipFree(w);
  ctx->tune = saved;
  ctx->timing = saved_timing;
  ctx->tune_cache = saved_cache;
  if (status) return status;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache[Key(bucket(n), b, ctx->precision)] = best;
  }
  if (out_knobs) best.to(out_knobs);
  return 0;
}

extern "C" int evd_tune_cache_put(long n, long b, int precision, const int *knobs) {
  if (n < 1) return -1;
  if (b < 1) return -2;
  if (precision < EVD_PREC_FP32 || precision > EVD_PREC_FP16) return -3;
  if (!knobs) return -4;
  Entry e;
  e.from(knobs);
  if (!e.valid(b)) return -4;
  std::lock_guard<std::mutex> lk(g_mu);
  g_cache[Key(bucket(n), b, precision)] = e;
  return 0;
}

extern "C" int evd_tune_cache_get(long n, long b, int precision, int *knobs) {
  if (!knobs) return -4;
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_cache.find(Key(bucket(n), b, precision));
  if (it == g_cache.end()) return 1;
  it->second.to(knobs);
  return 0;
}

extern "C" int evd_tune_cache_clear(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_cache.clear();
  return 0;
}

extern "C" int evd_tune_cache_save(const char *path) {
  if (!path) return -1;
  FILE *f = fopen(path, "w");
  if (!f) return EVD_E_UNSUPPORTED;
  std::lock_guard<std::mutex> lk(g_mu);
  fprintf(f, "# n_bucket b precision nb chase_groups bc_valu sbr_back_block chase_depth panel_rows gemm_tile "
             "bc_sweeps bc_waves\n");
  for (auto &kv : g_cache) {
    const Entry &e = kv.second;
    fprintf(f, "%ld %ld %d %d %d %d %d %d %d %d %d %d\n", std::get<0>(kv.first), std::get<1>(kv.first),
            std::get<2>(kv.first), e.nb, e.chase_groups, e.bc_valu, e.sbr_back_block, e.chase_depth, e.panel_rows,
            e.gemm_tile, e.bc_sweeps, e.bc_waves);
  }
  fclose(f);
  return 0;
}

// Entries are merged into the cache; malformed lines make the call fail
// without changing the cache.
extern "C" int evd_tune_cache_load(const char *path) {
  if (!path) return -1;
  FILE *f = fopen(path, "r");
  if (!f) return EVD_E_UNSUPPORTED;
  std::vector<std::pair<Key, Entry>> rows;
  char line[512];
  int bad = 0;
  while (fgets(line, sizeof line, f)) {
    if (line[0] == '#' || line[0] == '\n') continue;
    long nbk, b;
    int prec;
    Entry e;
    const int got = sscanf(line, "%ld %ld %d %d %d %d %d %d %d %d %d %d", &nbk, &b, &prec, &e.nb, &e.chase_groups,
                           &e.bc_valu, &e.sbr_back_block, &e.chase_depth, &e.panel_rows, &e.gemm_tile, &e.bc_sweeps,
                           &e.bc_waves);
    if ((got != 7 && got != 12) || nbk < 1 || b < 1 || prec < EVD_PREC_FP32 || prec > EVD_PREC_FP16 || !e.valid(b)) {
      bad = 1;
      break;
    }
    rows.push_back({Key(bucket(nbk), b, prec), e});
  }
  fclose(f);
  if (bad) return EVD_E_UNSUPPORTED;
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto &r : rows) g_cache[r.first] = r.second;
  return 0;
}
