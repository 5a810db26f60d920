// Request loop for uneven-size eigenproblems (the serving form of the
// reference's myEVD driver, src/EVD/myEVD.cu:251-706, which solves one matrix
// per process run).
//
// Design: S slots, each a context with its own HIP stream and a worker
// thread.  With an HBM budget, the whole budget is one device allocation made
// when the server is created (the pool); an admitted request (or batch) gets
// a contiguous first-fit region of the pool, lent to its slot's context as
// the arena, and returns it when it completes, so no device memory is
// allocated or freed while requests are in flight (a hipFree would stall the
// other slots' streams).  A request whose workspace alone exceeds the budget
// is refused at submission; any other request fits once the slots are idle,
// so admission cannot wait forever.  Without a budget each slot's context
// grows its own arena.
//
// Admission order: the queued request of the largest order first (FIFO among
// equal orders), then the first one in that order that fits the pool.  The
// slowest requests of a mix are its largest orders; starting them as soon as
// a slot is free shortens the tail of the latency distribution, while the
// small requests fill the other slots.  evd_server_set_policy switches to
// plain FIFO.  Requests that arrive while others run are admitted as slots
// and memory free up, and their kernels overlap on the device.  Every kernel
// of the solver is deterministic with respect to co-scheduling (no
// floating-point atomics, a fixed launch configuration per size), so a
// request gives the same bits alone and in a mix.  Latency = completion -
// submission, per request.
//
// Same-size batching: a slot that takes a request also takes up to
// kBatchMax - 1 queued requests of the same (n, b, nb, jobz), stages them
// into strided buffers carved from the same region and solves them with the
// batched solver (every stage one launch for the group; bitwise the same
// results).  Large orders gain too: the group's chase, panels and D&C levels
// are latency-bound chains that then run side by side.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"

namespace {

using clk = std::chrono::steady_clock;

struct Request {
  long id;
  long n, b, nb;
  float *A;
  long lda;
  double *w;
  float *Z;
  long ldz;
  size_t ws;
  clk::time_point t_submit;
};

struct Done {
  int status;
  double latency_ms;
  clk::time_point t_done;
};

// first-fit allocator over [0, size) with coalescing free blocks
struct FreeList {
  std::map<size_t, size_t> blocks;  // offset -> bytes
  size_t used = 0, peak = 0;
  void init(size_t size) {
    blocks.clear();
    if (size) blocks[0] = size;
  }
  bool alloc(size_t bytes, size_t *off) {
    for (auto it = blocks.begin(); it != blocks.end(); ++it)
      if (it->second >= bytes) {
        *off = it->first;
        const size_t rest = it->second - bytes;
        const size_t at = it->first + bytes;
        blocks.erase(it);
        if (rest) blocks[at] = rest;
        used += bytes;
        peak = std::max(peak, used);
        return true;
      }
    return false;
  }
  void release(size_t off, size_t bytes) {
    used -= bytes;
    auto it = blocks.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != blocks.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      blocks.erase(nx);
    }
    if (it != blocks.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        blocks.erase(it);
      }
    }
  }
};

}  // namespace

constexpr int kBatchMax = 32;  // requests per batch

struct evd_server {
  int device = 0;
  size_t budget = 0;  // 0 = unlimited (no pool)
  char *pool = nullptr;
  FreeList free;
  std::vector<evd_ctx *> ctx;
  bool batching = true;
  int precision = EVD_PREC_FP32;  // GEMM precision of every slot (evd_ctx_set_precision)
  int policy = 0;                 // 0: largest order first, 1: FIFO
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request> queue;
  std::unordered_map<long, Done> done;
  long next_id = 0;
  long running = 0;
  bool stop = false;
  bool have_first = false;
  clk::time_point t_first, t_last;
  std::vector<double> lat;
};

static size_t batch_stage_bytes(int count, long n) {
  return evd::align256((size_t)count * n * n * sizeof(float)) * 2 + evd::align256((size_t)count * n * sizeof(double));
}

static bool same_class(const Request &a, const Request &b) {
  return a.n == b.n && a.b == b.b && a.nb == b.nb && (a.Z != nullptr) == (b.Z != nullptr);
}

static void finish(evd_server *s, const std::vector<Request> &batch, int st) {
  const clk::time_point t1 = clk::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (const Request &r : batch) {
      const double ms = std::chrono::duration<double, std::milli>(t1 - r.t_submit).count();
      s->done[r.id] = Done{st, ms, t1};
      s->lat.push_back(ms);
    }
    s->t_last = t1;
    s->running -= (long)batch.size();
  }
  s->cv_done.notify_all();
  s->cv_work.notify_all();
}

// A group of same-class requests through the batched solver: stage A into
// strided storage carved from the slot's arena, solve, copy w and Z back
// (all on the slot's stream).
static int run_batch(evd_server *s, int slot, const std::vector<Request> &batch) {
  evd_ctx *ctx = s->ctx[slot];
  const int count = (int)batch.size();
  const long n = batch[0].n;
  const bool vec = batch[0].Z != nullptr;
  evd::ArenaFrame frame(ctx);
  float *A = evd::arena_alloc_n<float>(ctx, (size_t)count * n * n);
  float *Z = evd::arena_alloc_n<float>(ctx, (size_t)count * n * n);
  double *w = evd::arena_alloc_n<double>(ctx, (size_t)count * n);
  if (!A || !Z || !w) return EVD_E_NOMEM;
  for (int q = 0; q < count; q++) EVD_TRY(evd::copy_f32(ctx, batch[q].A, batch[q].lda, A + (long)q * n * n, n, n, n));
  EVD_TRY(evd::syevd_batched_call(ctx, vec ? 'V' : 'N', count, n, batch[0].b, batch[0].nb, A, n, n * n, w, n, Z, n,
                                  n * n));
  for (int q = 0; q < count; q++) {
    if (hipMemcpyAsync(batch[q].w, w + (long)q * n, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream) !=
        hipSuccess)
      return EVD_E_HIP;
    if (vec) EVD_TRY(evd::copy_f32(ctx, Z + (long)q * n * n, n, batch[q].Z, batch[q].ldz, n, n));
  }
  return evd_ctx_synchronize(ctx);
}

// queue positions in admission order
static std::vector<size_t> admission_order(const evd_server *s) {
  std::vector<size_t> ord(s->queue.size());
  for (size_t i = 0; i < ord.size(); i++) ord[i] = i;
  if (s->policy == 0)
    std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return s->queue[a].n > s->queue[b].n; });
  return ord;
}

static void worker_main(evd_server *s, int slot) {
  hipSetDevice(s->device);
  evd_ctx *ctx = s->ctx[slot];
  for (;;) {
    std::vector<Request> batch;
    size_t off = 0, need = 0;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      for (;;) {
        if (s->stop && s->queue.empty()) return;
        bool taken = false;
        for (size_t hpos : admission_order(s)) {
          const Request h = s->queue[hpos];
          // same-class requests queued behind the head (queue order)
          std::vector<size_t> group{hpos};
          if (s->batching && h.n > 2)
            for (size_t i = 0; i < s->queue.size() && (int)group.size() < kBatchMax; i++)
              if (i != hpos && same_class(s->queue[i], h)) group.push_back(i);
          // largest group that fits (halving), down to the request alone
          for (int cnt = (int)group.size(); cnt >= 1; cnt = (cnt > 2) ? cnt / 2 : cnt - 1) {
            need = (cnt == 1) ? h.ws
                              : evd::syevd_batched_call_workspace(ctx, cnt, h.n, h.b, h.nb, h.Z ? 'V' : 'N') +
                                    batch_stage_bytes(cnt, h.n) + ((size_t)1 << 20);
            if (s->pool ? s->free.alloc(need, &off) : true) {
              std::vector<size_t> pos(group.begin(), group.begin() + cnt);
              for (size_t p : pos) batch.push_back(s->queue[p]);
              std::sort(pos.begin(), pos.end());
              for (auto it = pos.rbegin(); it != pos.rend(); ++it) s->queue.erase(s->queue.begin() + (long)*it);
              s->running += cnt;
              taken = true;
              break;
            }
          }
          if (taken) break;
        }
        if (taken) break;
        s->cv_work.wait(lk);
      }
    }
    int st = s->pool ? evd::ctx_lend_arena(ctx, s->pool + off, need) : evd_ctx_reserve(ctx, need);
    if (st == 0) {
      if (batch.size() == 1) {
        const Request &r = batch[0];
        st = evd_syevd_f32(ctx, r.Z ? 'V' : 'N', r.n, r.b, r.nb, r.A, r.lda, r.w, r.Z, r.ldz);
        if (st == 0) st = evd_ctx_synchronize(ctx);
      } else {
        st = run_batch(s, slot, batch);
      }
    }
    if (s->pool) {
      hipStreamSynchronize(ctx->stream);  // the region is free only when the slot's work is done
      std::lock_guard<std::mutex> lk(s->mu);
      s->free.release(off, need);
    }
    finish(s, batch, st);
  }
}

extern "C" {

int evd_server_create(int device, int slots, size_t hbm_budget, evd_server **out) {
  if (!out) return -4;
  *out = nullptr;
  if (slots < 1 || slots > 64) return -2;
  evd_server *s = new evd_server();
  s->device = device;
  s->budget = hbm_budget;
  auto fail = [&](int st) {
    for (evd_ctx *x : s->ctx) evd_ctx_destroy(x);
    if (s->pool) hipFree(s->pool);
    delete s;
    return st;
  };
  for (int i = 0; i < slots; i++) {
    evd_ctx *c = nullptr;
    const int st = evd_ctx_create(device, nullptr, &c);
    if (st) return fail(st);
    s->ctx.push_back(c);
    // D&C plans of every order come from this pool: no allocation per new order
    if (evd::plan_pool_reserve(c, evd::kPlanPoolBytes)) return fail(EVD_E_NOMEM);
  }
  if (hbm_budget) {
    // the plan pools count against the budget too
    const size_t plans = (size_t)slots * evd::kPlanPoolBytes;
    if (hbm_budget <= plans) return fail(EVD_E_NOMEM);
    const size_t bytes = (hbm_budget - plans) & ~(size_t)255;
    if (hipSetDevice(device) != hipSuccess || hipMalloc(&s->pool, bytes) != hipSuccess) {
      s->pool = nullptr;
      return fail(EVD_E_NOMEM);
    }
    s->free.init(bytes);
    s->budget = bytes;
  }
  for (int i = 0; i < slots; i++) s->workers.emplace_back(worker_main, s, i);
  *out = s;
  return 0;
}

int evd_server_set_batching(evd_server *s, int on) {
  if (!s) return -1;
  std::lock_guard<std::mutex> lk(s->mu);
  s->batching = on != 0;
  return 0;
}

int evd_server_set_precision(evd_server *s, int precision) {
  if (!s) return -1;
  if (precision < EVD_PREC_FP32 || precision > EVD_PREC_FP16) return -2;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->running != 0 || !s->queue.empty()) return EVD_E_UNSUPPORTED;  // only between workloads
  for (evd_ctx *c : s->ctx) EVD_TRY(evd_ctx_set_precision(c, precision));
  s->precision = precision;
  return 0;
}

int evd_server_set_policy(evd_server *s, int policy) {
  if (!s) return -1;
  if (policy != 0 && policy != 1) return -2;
  std::lock_guard<std::mutex> lk(s->mu);
  s->policy = policy;
  return 0;
}

int evd_server_destroy(evd_server *s) {
  if (!s) return 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto &t : s->workers) t.join();
  for (evd_ctx *c : s->ctx) {
    evd::ctx_lend_arena(c, nullptr, 0);
    evd_ctx_destroy(c);
  }
  if (s->pool) hipFree(s->pool);
  delete s;
  return 0;
}

int evd_server_submit(evd_server *s, long n, long b, long nb, float *A, long lda, double *w, float *Z, long ldz,
                      long *id) {
  if (!s) return -1;
  if (n < 0) return -2;
  if (b < 1 || b > 64) return -3;
  if (nb < b) return -4;
  if (!A && n > 0) return -5;
  if (lda < std::max<long>(1, n)) return -6;
  if (!w && n > 0) return -7;
  if (Z && ldz < std::max<long>(1, n)) return -9;
  if (!id) return -10;
  size_t ws;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    ws = evd::syevd_call_workspace(s->ctx[0], n, b, nb, Z ? 'V' : 'N') + ((size_t)1 << 20);
  }
  if (s->pool && ws > s->budget) return EVD_E_NOMEM;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    const clk::time_point t = clk::now();
    if (!s->have_first) {
      s->have_first = true;
      s->t_first = t;
    }
    *id = s->next_id++;
    s->queue.push_back(Request{*id, n, b, nb, A, lda, w, Z, ldz, ws, t});
  }
  s->cv_work.notify_one();
  return 0;
}

int evd_server_wait(evd_server *s, long id) {
  if (!s) return -1;
  std::unique_lock<std::mutex> lk(s->mu);
  if (id < 0 || id >= s->next_id) return -2;
  s->cv_done.wait(lk, [&] { return s->done.count(id) > 0; });
  return s->done[id].status;
}

int evd_server_wait_all(evd_server *s) {
  if (!s) return -1;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->queue.empty() && s->running == 0; });
  for (auto &kv : s->done)
    if (kv.second.status) return kv.second.status;
  return 0;
}

int evd_server_request_ms(evd_server *s, long id, double *latency_ms) {
  if (!s || !latency_ms) return -1;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->done.find(id);
  if (it == s->done.end()) return -2;
  *latency_ms = it->second.latency_ms;
  return 0;
}

int evd_server_stats(evd_server *s, long *completed, double *solves_per_s, double *p50_ms, double *p99_ms,
                     size_t *arena_bytes) {
  if (!s) return -1;
  std::lock_guard<std::mutex> lk(s->mu);
  std::vector<double> v = s->lat;
  std::sort(v.begin(), v.end());
  auto pct = [&](double q) -> double {
    if (v.empty()) return 0.0;
    const size_t i = std::min(v.size() - 1, (size_t)(q * (double)(v.size() - 1) + 0.5));
    return v[i];
  };
  const double span = s->have_first ? std::chrono::duration<double>(s->t_last - s->t_first).count() : 0.0;
  if (completed) *completed = (long)v.size();
  if (solves_per_s) *solves_per_s = span > 0 ? (double)v.size() / span : 0.0;
  if (p50_ms) *p50_ms = pct(0.50);
  if (p99_ms) *p99_ms = pct(0.99);
  if (arena_bytes) {
    // peak pool bytes lent to slots, or the slots' own arenas without a budget
    size_t t = s->free.peak;
    if (!s->pool)
      for (const evd_ctx *c : s->ctx) t += evd_ctx_arena_bytes(c);
    *arena_bytes = t;
  }
  return 0;
}

}  // extern "C"
