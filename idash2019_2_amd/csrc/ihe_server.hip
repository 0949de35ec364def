// Serving loop: encrypt -> cloud -> decrypt requests against resident model
// bands, with continuous admission by HBM budget.
//
// A request is one cohort: genotypes of a model's tag SNPs for NS samples, a
// binary key and a seed, and a host buffer for the decrypted scores of every
// output and sample. Requests above 1024 samples run as chunks of at most
// 1024 samples (one packing layout each); the host buffers are chunk-major
// (chunk k: genotypes [num_tags][ns_k], scores [num_outputs][ns_k]), so each
// chunk moves with one contiguous copy each way. A worker thread owns the
// device side: every chunk takes one slot (a context with its own stream,
// default priority) and one block of a device arena sized by the HBM budget,
// and runs its whole pipeline there (H2D of its genotypes, encrypt, cloud on
// the resident band, decrypt, D2H of its scores). Chunks are admitted in
// arrival order as soon as a slot and arena space are free, so requests that
// arrive while others run join the device at once and their stages overlap
// across streams. The scores (4 bytes per output and sample, against about a
// thousandth of that in device time) cross the link one chunk at a time in
// admission order: the worker enqueues a chunk's D2H copy, on the chunk's own
// stream, once its decrypt has ended and the previous chunk's copy has, so
// transfers run back to back, never several at once, while later chunks
// compute. A chunk's result depends only on its request (seed + chunk index
// selects the random streams), so it is the same alone and in the mix.
//
// The worker holds the server's mutex only to update the shared state
// (requests, queue, arena, slot ownership): it polls the slots' events and
// enqueues a chunk's copies and kernels with the mutex released, so submit,
// poll and wait never wait behind a chunk's enqueue.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ihe_band.h"

using namespace ihe;
using Clock = std::chrono::steady_clock;

namespace {

constexpr uint32_t kChunkSamples = 1024;
constexpr size_t kAlign = 256;

size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

// First-fit allocator over one device allocation (offsets, coalescing frees).
struct Arena {
  size_t size = 0;
  std::map<size_t, size_t> free;  // offset -> length
  size_t used = 0, peak = 0;
  void init(size_t n) {
    size = n;
    free.clear();
    if (n) free[0] = n;
  }
  bool alloc(size_t n, size_t &off) {
    for (auto it = free.begin(); it != free.end(); ++it) {
      if (it->second < n) continue;
      off = it->first;
      const size_t rest = it->second - n;
      free.erase(it);
      if (rest) free[off + n] = rest;
      used += n;
      peak = std::max(peak, used);
      return true;
    }
    return false;
  }
  void release(size_t off, size_t n) {
    used -= n;
    auto next = free.lower_bound(off);
    if (next != free.end() && off + n == next->first) {
      n += next->second;
      next = free.erase(next);
    }
    if (next != free.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        prev->second += n;
        return;
      }
    }
    free[off] = n;
  }
};

struct ModelRec {
  const ihe_band *band;
  uint32_t num_tags, num_out;
};

struct Request {
  ihe_request r{};
  uint32_t chunks = 0, issued = 0, finished = 0;
  int rc = IHE_OK;
  std::string err;
  Clock::time_point t_arrive, t_done;
  bool done = false;
};

// A chunk taken from the queue, to be enqueued with the mutex released.
struct Issue {
  size_t slot;
  uint64_t id;
  ihe_request r;
  uint32_t k;
};

// Device bytes of one chunk: genotypes, input ciphertexts and variances,
// output ciphertexts, decrypted scores.
struct ChunkPlan {
  ihe_layout L;
  uint32_t num_in;
  size_t o_geno, o_in, o_var, o_out, o_dec, bytes;
};

ChunkPlan plan_chunk(const ModelRec &m, uint32_t ns) {
  ChunkPlan p;
  const uint32_t R = kN / ns;
  p.L = {ns, R, kN / R};
  p.num_in = (3 * m.num_tags + R - 1) / R;
  p.o_geno = 0;
  p.o_in = p.o_geno + align_up((size_t)m.num_tags * ns);
  p.o_var = p.o_in + align_up((size_t)p.num_in * kCtWords * 4);
  p.o_out = p.o_var + align_up((size_t)p.num_in * 8);
  p.o_dec = p.o_out + align_up((size_t)m.num_out * kCtWords * 4);
  p.bytes = p.o_dec + align_up((size_t)m.num_out * ns * 4);
  return p;
}

// A slot is busy from the moment a chunk is assigned to it until the worker
// has seen the chunk's copy complete. Only the worker changes a slot: `busy`
// (and the arena block) under the mutex, which get_stats reads it under too;
// the rest, which no other thread reads, without it.
struct Slot {
  ihe_ctx *ctx = nullptr;
  hipEvent_t ready = nullptr;  // the chunk's decrypt has ended
  hipEvent_t ev = nullptr;     // the chunk's scores are on the host
  bool busy = false;
  bool copying = false;   // its D2H copy and `ev` are on the stream
  void *dst = nullptr;    // the copy: host destination, device source, bytes
  const void *src = nullptr;
  size_t out_bytes = 0;
  uint64_t req = 0;
  size_t off = 0, bytes = 0;
  ChunkPlan plan{};
  ihe_key *key = nullptr;
};

}  // namespace

struct ihe_server {
  ihe_ctx *ctx = nullptr;
  ihe_server_config cfg{};
  void *arena_base = nullptr;
  Arena arena;
  std::vector<Slot> slots;
  // slots whose device work is enqueued, in admission order; the first is the
  // one whose copy runs or runs next (the worker's own: no mutex)
  std::deque<size_t> order;
  std::vector<ModelRec> models;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<uint64_t> pending;  // requests with chunks left to issue, arrival order
  std::unordered_map<uint64_t, Request> reqs;
  uint64_t next_id = 1, completed = 0;
  bool stop = false;
  std::thread worker;

  // Enqueues chunk `k` of a request on slot `s` (mutex released; the slot and
  // its arena block are this chunk's).
  int issue(Slot &s, const ihe_request &r, const ModelRec &m, uint32_t k) {
    const ChunkPlan &p = s.plan;
    const uint32_t col0 = k * kChunkSamples, ns = p.L.num_samples;
    char *base = static_cast<char *>(arena_base) + s.off;
    int8_t *d_geno = reinterpret_cast<int8_t *>(base + p.o_geno);
    int32_t *d_in = reinterpret_cast<int32_t *>(base + p.o_in);
    double *d_var = reinterpret_cast<double *>(base + p.o_var);
    int32_t *d_out = reinterpret_cast<int32_t *>(base + p.o_out);
    float *d_dec = reinterpret_cast<float *>(base + p.o_dec);
    hipStream_t st = s.ctx->stream;
    if (int rc = make_current(s.ctx)) return rc;
    IHE_HIP(hipMemcpyAsync(d_geno, r.geno + (size_t)m.num_tags * col0, (size_t)m.num_tags * ns,
                           hipMemcpyHostToDevice, st));
    if (int rc = ihe_key_create(s.ctx, r.key, &s.key)) return rc;
    if (int rc = ihe_encrypt(s.ctx, s.key, &p.L, m.num_tags, d_geno, r.seed + k, d_in, d_var))
      return rc;
    if (int rc = ihe_cloud_score_layout(s.ctx, m.band, &p.L, p.num_in, d_in, nullptr, d_out, nullptr))
      return rc;
    if (int rc = ihe_decrypt(s.ctx, s.key, ns, d_out, m.num_out, d_dec)) return rc;
    IHE_HIP(hipEventRecord(s.ready, st));
    s.dst = r.out + (size_t)m.num_out * col0;
    s.src = d_dec;
    s.out_bytes = (size_t)m.num_out * ns * 4;
    return IHE_OK;
  }

  // Walks the copy order: starts the first chunk's D2H once its decrypt has
  // ended, and when that copy has ended, reports the chunk and goes on to the
  // next. Returns the chunks that ended (slot, status).
  void advance_copies(std::vector<std::pair<size_t, hipError_t>> &ended) {
    while (!order.empty()) {
      const size_t i = order.front();
      Slot &s = slots[i];
      if (!s.copying) {
        hipError_t q = hipEventQuery(s.ready);
        if (q == hipErrorNotReady) return;
        if (q == hipSuccess && make_current(s.ctx) != IHE_OK) q = hipErrorInvalidDevice;
        if (q == hipSuccess)
          q = hipMemcpyAsync(s.dst, s.src, s.out_bytes, hipMemcpyDeviceToHost, s.ctx->stream);
        if (q == hipSuccess) q = hipEventRecord(s.ev, s.ctx->stream);
        if (q != hipSuccess) {
          (void)hipStreamSynchronize(s.ctx->stream);
          ended.push_back({i, q});
          order.pop_front();
          continue;
        }
        s.copying = true;
      }
      const hipError_t q = hipEventQuery(s.ev);
      if (q == hipErrorNotReady) return;
      ended.push_back({i, q});
      order.pop_front();
    }
  }

  // The chunk on slot `s` has ended (mutex held).
  void finish_chunk(Slot &s, int rc, const std::string &err) {
    arena.release(s.off, s.bytes);
    s.busy = false;
    s.copying = false;
    auto it = reqs.find(s.req);
    if (it == reqs.end()) return;
    Request &R = it->second;
    if (rc && !R.rc) R.rc = rc, R.err = err;
    if (++R.finished == R.chunks) {
      R.done = true;
      R.t_done = Clock::now();
      ++completed;
      cv_done.notify_all();
    }
  }

  void fail_request(uint64_t id, int rc, const std::string &err) {
    Request &R = reqs[id];
    R.rc = rc, R.err = err;
    R.finished += R.chunks - R.issued;  // never issued
    R.issued = R.chunks;
    if (R.finished == R.chunks) {
      R.done = true;
      R.t_done = Clock::now();
      ++completed;
      cv_done.notify_all();
    }
  }

  // Assigns queued chunks to free slots and arena blocks, in arrival order
  // (mutex held); the chunks are enqueued by the caller afterwards.
  bool admit(std::vector<Issue> &todo) {
    bool progress = false;
    while (!pending.empty()) {
      const uint64_t id = pending.front();
      Request &R = reqs[id];
      size_t free_slot = slots.size();
      for (size_t i = 0; i < slots.size(); ++i)
        if (!slots[i].busy) {
          free_slot = i;
          break;
        }
      if (free_slot == slots.size()) break;
      const uint32_t k = R.issued;
      const uint32_t ns = std::min(kChunkSamples, R.r.num_samples - k * kChunkSamples);
      const ChunkPlan p = plan_chunk(models[R.r.model], ns);
      if (p.bytes > arena.size) {
        fail_request(id, IHE_ERR_NOMEM, "a chunk of this request exceeds the HBM budget");
        pending.pop_front();
        progress = true;
        continue;
      }
      size_t off;
      if (!arena.alloc(p.bytes, off)) break;  // wait for memory, in arrival order
      Slot &s = slots[free_slot];
      s.busy = true;
      s.req = id;
      s.off = off;
      s.bytes = p.bytes;
      s.plan = p;
      todo.push_back({free_slot, id, R.r, k});
      if (++R.issued == R.chunks) pending.pop_front();
      progress = true;
    }
    return progress;
  }

  void run() {
    std::vector<Issue> todo;
    std::vector<std::pair<size_t, hipError_t>> ended;
    for (;;) {
      // copies to start and chunks that have ended (event queries without the mutex)
      ended.clear();
      advance_copies(ended);
      for (const auto &e : ended) {
        Slot &s = slots[e.first];
        if (s.key) ihe_key_destroy(s.key);
        s.key = nullptr;
      }
      todo.clear();
      bool busy = false, progress = !ended.empty();
      {
        std::unique_lock<std::mutex> lk(mu);
        for (const auto &e : ended)
          finish_chunk(slots[e.first], e.second == hipSuccess ? IHE_OK : IHE_ERR_HIP,
                       e.second == hipSuccess
                           ? ""
                           : std::string("request failed on the device: ") + hipGetErrorString(e.second));
        progress |= admit(todo);
        for (const Slot &s : slots) busy |= s.busy;
        if (stop && !busy && pending.empty()) break;
        if (!progress)
          cv_work.wait_for(lk, std::chrono::microseconds(busy ? 20 : 2000));
      }
      // enqueue the admitted chunks with the mutex released
      for (const Issue &t : todo) {
        Slot &s = slots[t.slot];
        const int rc = issue(s, t.r, models_snapshot(t.r.model), t.k);
        if (rc) {
          // nothing of this chunk reached the device queue after the failure
          const std::string err = ihe_last_error();
          (void)hipStreamSynchronize(s.ctx->stream);
          if (s.key) ihe_key_destroy(s.key);
          s.key = nullptr;
          std::lock_guard<std::mutex> lk(mu);
          finish_chunk(s, rc, err);
        } else {
          order.push_back(t.slot);
        }
      }
    }
  }

  // A model record, read under the mutex (add_model may grow the table).
  ModelRec models_snapshot(uint32_t m) {
    std::lock_guard<std::mutex> lk(mu);
    return models[m];
  }
};

extern "C" {

int ihe_server_create(ihe_ctx *ctx, const ihe_server_config *cfg, ihe_server **out) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!cfg || !out) return fail(IHE_ERR_ARG, "null argument");
  if (cfg->num_slots == 0 || cfg->num_slots > 64) return fail(IHE_ERR_ARG, "num_slots must be in [1, 64]");
  if (cfg->hbm_budget == 0) return fail(IHE_ERR_ARG, "zero HBM budget");
  *out = nullptr;
  ihe_server *s = new ihe_server();
  s->ctx = ctx;
  s->cfg = *cfg;
  if (s->cfg.queue_capacity == 0) s->cfg.queue_capacity = 1024;
  if (int rc = ihe_malloc(ctx, align_up(cfg->hbm_budget), &s->arena_base)) {
    delete s;
    return rc;
  }
  s->arena.init(align_up(cfg->hbm_budget));
  s->slots.resize(cfg->num_slots);
  for (Slot &sl : s->slots) {
    int rc = ihe_ctx_create(ctx->device, &sl.ctx);
    if (!rc && (hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&sl.ready, hipEventDisableTiming) != hipSuccess))
      rc = fail(IHE_ERR_HIP, "event creation failed");
    if (rc) {
      for (Slot &o : s->slots) {
        if (o.ev) (void)hipEventDestroy(o.ev);
        if (o.ready) (void)hipEventDestroy(o.ready);
        if (o.ctx) ihe_ctx_destroy(o.ctx);
      }
      (void)ihe_free(ctx, s->arena_base);
      delete s;
      return rc;
    }
  }
  s->worker = std::thread([s] { s->run(); });
  *out = s;
  return IHE_OK;
}

int ihe_server_add_model(ihe_server *s, const ihe_band *band, uint32_t num_tags, uint32_t *model) {
  if (!s || !band || !model) return fail(IHE_ERR_ARG, "null argument");
  if (band->ctx->device != s->ctx->device) return fail(IHE_ERR_ARG, "band is not on the server's device");
  if (band->mode != kModeMfma)
    return fail(IHE_ERR_UNSUPPORTED, "served models need MFMA bands (any sample count)");
  if (band->has_cols && band->max_bidx >= 3 * num_tags)
    return fail(IHE_ERR_RANGE, "the model references tags beyond num_tags");
  std::lock_guard<std::mutex> lk(s->mu);
  s->models.push_back({band, num_tags, band->num_out});
  band->users++;
  *model = (uint32_t)s->models.size() - 1;
  return IHE_OK;
}

int ihe_server_submit(ihe_server *s, const ihe_request *req, uint64_t *ticket) {
  if (!s || !req || !ticket) return fail(IHE_ERR_ARG, "null argument");
  if (req->num_samples == 0 || !req->geno || !req->key || !req->out)
    return fail(IHE_ERR_ARG, "request needs samples, genotypes, a key and an output buffer");
  for (uint32_t i = 0; i < kN; ++i)
    if (req->key[i] != 0 && req->key[i] != 1) return fail(IHE_ERR_ARG, "TLWE key must be binary");
  std::lock_guard<std::mutex> lk(s->mu);
  if (req->model >= s->models.size()) return fail(IHE_ERR_ARG, "unknown model");
  if (s->pending.size() >= s->cfg.queue_capacity) return fail(IHE_ERR_BUSY, "request queue is full");
  const uint64_t id = s->next_id++;
  Request &R = s->reqs[id];
  R.r = *req;
  R.chunks = (req->num_samples + kChunkSamples - 1) / kChunkSamples;
  // steady_clock is CLOCK_MONOTONIC, the clock arrival_ns is read from
  R.t_arrive = req->arrival_ns
                   ? Clock::time_point(std::chrono::nanoseconds((int64_t)req->arrival_ns))
                   : Clock::now();
  s->pending.push_back(id);
  *ticket = id;
  s->cv_work.notify_one();
  return IHE_OK;
}

static int collect(ihe_server *s, std::unordered_map<uint64_t, Request>::iterator it,
                   double *latency_ms) {
  const Request R = it->second;
  s->reqs.erase(it);
  if (latency_ms)
    *latency_ms = std::chrono::duration<double, std::milli>(R.t_done - R.t_arrive).count();
  if (R.rc) return fail(R.rc, R.err);
  return IHE_OK;
}

int ihe_server_poll(ihe_server *s, uint64_t ticket, int *done, double *latency_ms) {
  if (!s || !done) return fail(IHE_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return fail(IHE_ERR_ARG, "unknown ticket");
  *done = it->second.done ? 1 : 0;
  if (!it->second.done) return IHE_OK;
  return collect(s, it, latency_ms);
}

int ihe_server_wait(ihe_server *s, uint64_t ticket, double *latency_ms) {
  if (!s) return fail(IHE_ERR_ARG, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return fail(IHE_ERR_ARG, "unknown ticket");
  // another thread's poll or wait may collect the ticket meanwhile
  s->cv_done.wait(lk, [&] {
    auto f = s->reqs.find(ticket);
    return f == s->reqs.end() || f->second.done;
  });
  it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return fail(IHE_ERR_ARG, "ticket collected by another call");
  return collect(s, it, latency_ms);
}

int ihe_server_get_stats(ihe_server *s, ihe_server_stats *st) {
  if (!s || !st) return fail(IHE_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  st->completed = s->completed;
  st->pending = (uint32_t)s->pending.size();
  uint32_t busy = 0;
  for (const Slot &sl : s->slots) busy += sl.busy ? 1u : 0u;
  st->running = busy;
  st->arena_bytes = s->arena.size;
  st->arena_used = s->arena.used;
  st->arena_peak = s->arena.peak;
  return IHE_OK;
}

int ihe_server_destroy(ihe_server *s) {
  if (!s) return IHE_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  for (Slot &sl : s->slots) {
    if (sl.ev) (void)hipEventDestroy(sl.ev);
    if (sl.ready) (void)hipEventDestroy(sl.ready);
    if (sl.ctx) ihe_ctx_destroy(sl.ctx);
  }
  (void)ihe_free(s->ctx, s->arena_base);
  for (const ModelRec &m : s->models) m.band->users--;
  delete s;
  return IHE_OK;
}

}  // extern "C"
