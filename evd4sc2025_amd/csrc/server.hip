// Request loop for uneven-size eigenproblems (the serving form of the
// reference's myEVD driver, src/EVD/myEVD.cu:251-706, which solves one matrix
// per process run).
//
// Design: S slots, each a context with its own HIP stream and device-resident
// workspace arena, and a worker thread.  Requests (device-resident A, w, Z
// owned by the caller) enter a FIFO queue; a worker takes the head request when
// the slot arenas, grown to the request's workspace, stay within the HBM
// budget.  Requests that arrive while others run are admitted as slots and
// memory free up, and their kernels overlap on the device (continuous
// batching at request granularity).  Every kernel of the solver is
// deterministic with respect to co-scheduling (no floating-point atomics, a
// fixed launch configuration per size), so a request gives the same bits alone
// and in a mix.  Latency = completion - submission, per request.
//
// Same-size batching: when a slot takes a request of order n <= kBatchMaxN,
// it also takes up to kBatchMax - 1 queued requests of the same (n, b, nb,
// jobz), stages them into strided buffers and solves them with
// evd_syevd_batched_f32 (every stage one launch for the group; bitwise the
// same results).  Large orders gain too: the group's chase, panels and D&C
// levels are latency-bound chains that then run side by side.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"

namespace evd {
}

extern "C" int evd_syevd_batched_f32(evd_ctx *ctx, char jobz, int count, long n, long b, long nb, float *A, long lda,
                                     long a_stride, double *w, long w_stride, float *Z, long ldz, long z_stride);

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

}  // namespace

constexpr long kBatchMaxN = 16384;  // orders that are batched
constexpr int kBatchMax = 32;      // requests per batch

struct evd_server {
  int device = 0;
  size_t budget = 0;
  std::vector<evd_ctx *> ctx;
  std::vector<size_t> arena;  // bytes each slot's arena (and batch staging) holds (grows only)
  std::vector<char *> stage;  // per-slot staging for batched requests
  std::vector<size_t> stage_bytes;
  bool batching = true;
  int precision = EVD_PREC_FP32;  // GEMM precision of every slot (evd_ctx_set_precision)
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

static size_t arena_total(const evd_server *s) {
  size_t t = 0;
  for (size_t a : s->arena) t += a;
  return t;
}

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
// strided storage, solve, copy w and Z back (all on the slot's stream).
static int run_batch(evd_server *s, int slot, const std::vector<Request> &batch, size_t ws) {
  evd_ctx *ctx = s->ctx[slot];
  const int count = (int)batch.size();
  const long n = batch[0].n;
  const bool vec = batch[0].Z != nullptr;
  const size_t need = batch_stage_bytes(count, n);
  if (s->stage_bytes[slot] < need) {
    if (s->stage[slot]) {
      hipStreamSynchronize(ctx->stream);
      hipFree(s->stage[slot]);
      s->stage[slot] = nullptr;
      s->stage_bytes[slot] = 0;
    }
    if (hipMalloc(&s->stage[slot], need) != hipSuccess) return EVD_E_NOMEM;
    s->stage_bytes[slot] = need;
  }
  float *A = reinterpret_cast<float *>(s->stage[slot]);
  float *Z = reinterpret_cast<float *>(s->stage[slot] + evd::align256((size_t)count * n * n * sizeof(float)));
  double *w = reinterpret_cast<double *>(s->stage[slot] + 2 * evd::align256((size_t)count * n * n * sizeof(float)));
  for (int q = 0; q < count; q++) EVD_TRY(evd::copy_f32(ctx, batch[q].A, batch[q].lda, A + (long)q * n * n, n, n, n));
  EVD_TRY(evd_ctx_reserve(ctx, ws));
  EVD_TRY(evd_syevd_batched_f32(ctx, vec ? 'V' : 'N', count, n, batch[0].b, batch[0].nb, A, n, n * n, w, n, Z, n,
                                n * n));
  for (int q = 0; q < count; q++) {
    if (hipMemcpyAsync(batch[q].w, w + (long)q * n, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream) !=
        hipSuccess)
      return EVD_E_HIP;
    if (vec) EVD_TRY(evd::copy_f32(ctx, Z + (long)q * n * n, n, batch[q].Z, batch[q].ldz, n, n));
  }
  return evd_ctx_synchronize(ctx);
}

static void worker_main(evd_server *s, int slot) {
  hipSetDevice(s->device);
  evd_ctx *ctx = s->ctx[slot];
  for (;;) {
    std::vector<Request> batch;
    size_t bws = 0;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      for (;;) {
        if (s->stop && s->queue.empty()) return;
        if (!s->queue.empty()) {
          const Request h = s->queue.front();
          // same-class requests queued behind the head (FIFO among them)
          std::vector<size_t> more;  // queue positions (erasing invalidates deque iterators)
          if (s->batching && h.n <= kBatchMaxN && h.n > 2)
            for (size_t i = 1; i < s->queue.size() && (int)more.size() + 1 < kBatchMax; i++)
              if (same_class(s->queue[i], h)) more.push_back(i);
          // largest group whose workspace fits the budget (a single request as before)
          int cnt = 1 + (int)more.size();
          size_t need = 0;
          for (; cnt >= 1; cnt = (cnt > 2) ? cnt / 2 : cnt - 1) {
            const size_t ws = (cnt == 1) ? h.ws
                                         : evd::syevd_batched_workspace(cnt, h.n, h.b, h.nb, h.Z ? 'V' : 'N',
                                                                        s->precision) +
                                               batch_stage_bytes(cnt, h.n);
            need = std::max(ws, s->arena[slot]);
            if (arena_total(s) - s->arena[slot] + need <= s->budget) {
              bws = (cnt == 1) ? h.ws : ws - batch_stage_bytes(cnt, h.n);
              break;
            }
          }
          if (cnt >= 1) {
            batch.push_back(h);
            for (int i = 0; i + 1 < cnt; i++) batch.push_back(s->queue[more[i]]);
            for (int i = cnt - 2; i >= 0; i--) s->queue.erase(s->queue.begin() + (long)more[i]);
            s->queue.pop_front();
            s->arena[slot] = need;
            s->running += cnt;
            break;
          }
        }
        s->cv_work.wait(lk);
      }
    }
    int st;
    if (batch.size() == 1) {
      const Request &r = batch[0];
      st = evd_ctx_reserve(ctx, r.ws);
      if (st == 0) st = evd_syevd_f32(ctx, r.Z ? 'V' : 'N', r.n, r.b, r.nb, r.A, r.lda, r.w, r.Z, r.ldz);
      if (st == 0) st = evd_ctx_synchronize(ctx);
    } else {
      st = run_batch(s, slot, batch, bws);
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
  s->budget = hbm_budget ? hbm_budget : ~(size_t)0;
  for (int i = 0; i < slots; i++) {
    evd_ctx *c = nullptr;
    const int st = evd_ctx_create(device, nullptr, &c);
    if (st) {
      for (evd_ctx *x : s->ctx) evd_ctx_destroy(x);
      delete s;
      return st;
    }
    s->ctx.push_back(c);
    s->arena.push_back(0);
    s->stage.push_back(nullptr);
    s->stage_bytes.push_back(0);
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

int evd_server_destroy(evd_server *s) {
  if (!s) return 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto &t : s->workers) t.join();
  for (char *p : s->stage)
    if (p) hipFree(p);
  for (evd_ctx *c : s->ctx) evd_ctx_destroy(c);
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
    ws = evd::syevd_workspace(n, b, nb, Z ? 'V' : 'N', s->precision);
  }
  if (ws > s->budget) return EVD_E_NOMEM;
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
  if (arena_bytes) *arena_bytes = arena_total(s);
  return 0;
}

}  // extern "C"
