// server.cpp -- the request server (heu_server_* in include/heu_paillier.h).
//
// One scheduler thread per server.  Each turn it retires finished batches
// (stream events), then, while a stream slot is free and requests are queued,
// admits the longest prefix of the queue whose batch plan fits the largest
// free block of the workspace arena (and the unit cap), plans it into that
// block, copies its work lists up from pinned staging and enqueues it on the
// slot's stream.  Requests arriving meanwhile queue for the next free slot.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine_internal.hpp"

using namespace heu_engine;
using Clock = std::chrono::steady_clock;

namespace {

enum ReqState { kQueued = 0, kRunning = 1, kDone = 2, kFailed = 3 };

struct Request {
  heu_task task;
  std::vector<int64_t> offsets;  // owned copy for SUM / DOT
  Clock::time_point t_submit, t_done;
  int state = kQueued;
  std::string error;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  size_t off = 0, bytes = 0;
  std::vector<uint64_t> tickets;
  char* pinned = nullptr;  // staging of the batch's work lists
  size_t pinned_cap = 0;
};

// First-fit allocator over [0, size) with coalescing.
struct Arena {
  std::map<size_t, size_t> free_;  // offset -> bytes
  void init(size_t size) {
    free_.clear();
    free_[0] = size;
  }
  size_t largest() const {
    size_t m = 0;
    for (auto& kv : free_) m = std::max(m, kv.second);
    return m;
  }
  bool take(size_t bytes, size_t* off) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes;
      const size_t at = it->first + bytes;
      free_.erase(it);
      if (rest) free_[at] = rest;
      return true;
    }
    return false;
  }
  void give(size_t off, size_t bytes) {
    auto it = free_.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
};

size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

struct heu_server {
  heu_ctx* ctx = nullptr;
  heu_server_config cfg{};
  DevBuf ws;
  Arena arena;
  size_t in_use = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<uint64_t> pending;
  std::unordered_map<uint64_t, Request> reqs;
  uint64_t next_ticket = 1;
  std::vector<Slot> slots;
  std::thread worker;
  bool stop = false;
  heu_server_counters st{};

  void run();
  bool retire();
  bool admit(std::unique_lock<std::mutex>& lk);
  void fail_request(uint64_t t, const std::string& why) {
    Request& r = reqs[t];
    r.state = kFailed;
    r.error = why;
    r.t_done = Clock::now();
    ++st.failed;
  }
};

// Retires finished slots; returns true if any finished.
bool heu_server::retire() {
  bool any = false;
  for (Slot& s : slots) {
    if (!s.busy || hipEventQuery(s.done) != hipSuccess) continue;
    std::lock_guard<std::mutex> lk(mu);
    const auto now = Clock::now();
    for (uint64_t t : s.tickets) {
      auto it = reqs.find(t);
      if (it == reqs.end()) continue;
      it->second.state = kDone;
      it->second.t_done = now;
      ++st.completed;
    }
    arena.give(s.off, s.bytes);
    in_use -= s.bytes;
    s.busy = false;
    s.tickets.clear();
    any = true;
  }
  if (any) cv_done.notify_all();
  return any;
}

// Admits one batch into a free slot if possible (called with `lk` held; may
// release it while planning).  Returns true if a batch was started or a
// request was failed.
bool heu_server::admit(std::unique_lock<std::mutex>& lk) {
  Slot* slot = nullptr;
  for (Slot& s : slots)
    if (!s.busy) {
      slot = &s;
      break;
    }
  if (!slot || pending.empty()) return false;
  const size_t n = std::min<size_t>(pending.size(), cfg.max_batch_tasks);
  std::vector<uint64_t> tickets(pending.begin(), pending.begin() + n);
  std::vector<heu_task> tasks;
  for (uint64_t t : tickets) tasks.push_back(reqs[t].task);
  const size_t room = arena.largest(), total = ws.bytes;
  lk.unlock();
  // longest admissible prefix: bytes fit the largest free block, units the cap
  auto size_of = [&](size_t k, PlanOut* p) {
    std::lock_guard<std::mutex> cl(ctx->mu);
    return plan_batch(ctx, tasks.data(), k, nullptr, p);
  };
  auto fits = [&](const PlanOut& p, size_t k) {
    return round256(p.bytes) <= room && (cfg.max_batch_units == 0 || p.units <= cfg.max_batch_units || k == 1);
  };
  size_t k = n;
  PlanOut p;
  int rc = size_of(k, &p);
  if (rc == HEU_OK && !fits(p, k)) {
    size_t lo = 0, hi = n;  // fits(lo) known true (or lo == 0), fits(hi) false
    while (hi - lo > 1) {
      const size_t mid = (lo + hi) / 2;
      PlanOut q;
      if (size_of(mid, &q) == HEU_OK && fits(q, mid)) lo = mid;
      else hi = mid;
    }
    k = lo;
    if (k) rc = size_of(k, &p);
  }
  lk.lock();
  if (rc != HEU_OK || k == 0) {
    // the head request alone: fail it if it can never fit, else wait for room
    PlanOut one;
    lk.unlock();
    const int rc1 = size_of(1, &one);
    lk.lock();
    if (rc1 != HEU_OK || round256(one.bytes) > total) {
      fail_request(tickets[0], rc1 != HEU_OK ? heu_last_error() : "request needs more workspace than the server has");
      pending.pop_front();
      cv_done.notify_all();
      return true;
    }
    return false;
  }
  size_t off = 0;
  const size_t bytes = round256(std::max<size_t>(p.bytes, 256));
  if (!arena.take(bytes, &off)) return false;
  in_use += bytes;
  st.peak_workspace_bytes = std::max(st.peak_workspace_bytes, in_use);
  for (size_t i = 0; i < k; ++i) {
    pending.pop_front();
    reqs[tickets[i]].state = kRunning;
  }
  slot->busy = true;
  slot->off = off;
  slot->bytes = bytes;
  slot->tickets.assign(tickets.begin(), tickets.begin() + k);
  ++st.batches;
  st.max_batch_requests = std::max<uint64_t>(st.max_batch_requests, k);
  lk.unlock();
  PlanOut real;
  char* base = ws.as<char>() + off;
  {
    std::lock_guard<std::mutex> cl(ctx->mu);
    rc = plan_batch(ctx, tasks.data(), k, base, &real);
  }
  size_t meta_bytes = 0;
  for (auto& m : real.metas) meta_bytes += m.second.size();
  if (rc == HEU_OK && meta_bytes > slot->pinned_cap) {
    if (slot->pinned) (void)hipHostFree(slot->pinned);
    slot->pinned = nullptr;
    slot->pinned_cap = 0;
    if (hipHostMalloc(reinterpret_cast<void**>(&slot->pinned), meta_bytes * 2) == hipSuccess)
      slot->pinned_cap = meta_bytes * 2;
    else rc = HEU_ENOMEM;
  }
  if (rc == HEU_OK) {
    size_t at = 0;
    for (auto& m : real.metas) {
      std::memcpy(slot->pinned + at, m.second.data(), m.second.size());
      if (hipMemcpyAsync(base + m.first, slot->pinned + at, m.second.size(), hipMemcpyHostToDevice, slot->stream) !=
          hipSuccess)
        rc = HEU_EDEVICE;
      at += m.second.size();
    }
    for (auto& s : real.steps) s(slot->stream);
    if (hipGetLastError() != hipSuccess) rc = HEU_EDEVICE;
  }
  if (hipEventRecord(slot->done, slot->stream) != hipSuccess) rc = HEU_EDEVICE;
  lk.lock();
  if (rc != HEU_OK) {
    // nothing (or not all) was enqueued: the requests fail; the slot retires
    // through its event as usual
    for (uint64_t t : slot->tickets) fail_request(t, "could not enqueue the batch");
    slot->tickets.clear();
    cv_done.notify_all();
  }
  return true;
}

void heu_server::run() {
  (void)hipSetDevice(ctx->device);
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    lk.unlock();
    retire();
    lk.lock();
    bool busy = false;
    for (Slot& s : slots) busy |= s.busy;
    if (stop && pending.empty() && !busy) break;
    if (admit(lk)) continue;
    // nothing admissible now: wait for a submission, or poll running batches
    if (busy) cv_work.wait_for(lk, std::chrono::microseconds(50));
    else cv_work.wait(lk, [&] { return stop || !pending.empty(); });
  }
}

extern "C" {

int heu_server_create(heu_ctx* ctx, const heu_server_config* cfg, heu_server** out) {
  if (!ctx || !out) return fail(HEU_EINVAL, "null argument");
  auto srv = std::make_unique<heu_server>();
  srv->ctx = ctx;
  if (cfg) srv->cfg = *cfg;
  if (!srv->cfg.workspace_bytes) srv->cfg.workspace_bytes = (size_t)8 << 30;
  if (!srv->cfg.max_inflight) srv->cfg.max_inflight = 2;
  if (!srv->cfg.max_batch_tasks) srv->cfg.max_batch_tasks = 512;
  if (srv->cfg.max_inflight > 8) return fail(HEU_EINVAL, "max_inflight must be at most 8");
  HIP_TRY(hipSetDevice(ctx->device));
  int rc = srv->ws.alloc(round256(srv->cfg.workspace_bytes));
  if (rc) return rc;
  srv->arena.init(srv->ws.bytes);
  srv->slots.resize(srv->cfg.max_inflight);
  for (Slot& s : srv->slots) {
    HIP_TRY(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
  }
  heu_server* raw = srv.release();
  raw->worker = std::thread([raw] { raw->run(); });
  *out = raw;
  return HEU_OK;
}

int heu_server_submit(heu_server* srv, const heu_task* task, uint64_t* ticket) {
  if (!srv || !task || !ticket) return fail(HEU_EINVAL, "null argument");
  if (task->op == HEU_OP_BARRIER) return fail(HEU_EINVAL, "a request cannot be a barrier");
  Request r;
  r.task = *task;
  if ((task->op == HEU_OP_SUM || task->op == HEU_OP_DOT_I64) && task->segments) {
    if (!task->offsets) return fail(HEU_EINVAL, "null offsets");
    r.offsets.assign(task->offsets, task->offsets + task->segments + 1);
    r.task.offsets = r.offsets.data();
  }
  {
    PlanOut chk;  // argument checks (sizing pass only)
    std::lock_guard<std::mutex> cl(srv->ctx->mu);
    int rc = plan_batch(srv->ctx, &r.task, 1, nullptr, &chk);
    if (rc) return rc;
  }
  std::lock_guard<std::mutex> lk(srv->mu);
  if (srv->stop) return fail(HEU_EINVAL, "server is stopping");
  const uint64_t t = srv->next_ticket++;
  r.t_submit = Clock::now();
  auto& slot = srv->reqs[t];
  slot = std::move(r);
  if (!slot.offsets.empty()) slot.task.offsets = slot.offsets.data();  // the vector moved
  srv->pending.push_back(t);
  ++srv->st.submitted;
  *ticket = t;
  srv->cv_work.notify_one();
  return HEU_OK;
}

int heu_server_wait(heu_server* srv, uint64_t ticket, double timeout_ms, double* latency_ms) {
  if (!srv) return fail(HEU_EINVAL, "null server");
  std::unique_lock<std::mutex> lk(srv->mu);
  auto it = srv->reqs.find(ticket);
  if (it == srv->reqs.end()) return fail(HEU_EINVAL, "unknown ticket %llu", (unsigned long long)ticket);
  auto ready = [&] {
    auto i = srv->reqs.find(ticket);
    return i == srv->reqs.end() || i->second.state == kDone || i->second.state == kFailed;
  };
  if (timeout_ms < 0) srv->cv_done.wait(lk, ready);
  else if (!srv->cv_done.wait_for(lk, std::chrono::microseconds((long long)(timeout_ms * 1e3)), ready))
    return fail(HEU_EDEVICE, "request %llu: timeout", (unsigned long long)ticket);
  it = srv->reqs.find(ticket);
  if (it == srv->reqs.end()) return fail(HEU_EINVAL, "ticket %llu was waited elsewhere", (unsigned long long)ticket);
  const Request& r = it->second;
  const bool ok = r.state == kDone;
  const std::string err = r.error;
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r.t_done - r.t_submit).count();
  srv->reqs.erase(it);
  if (!ok) return fail(HEU_EDEVICE, "request %llu failed: %s", (unsigned long long)ticket, err.c_str());
  return HEU_OK;
}

int heu_server_stats(heu_server* srv, heu_server_counters* stats) {
  if (!srv || !stats) return fail(HEU_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(srv->mu);
  *stats = srv->st;
  return HEU_OK;
}

int heu_server_destroy(heu_server* srv) {
  if (!srv) return HEU_OK;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stop = true;
  }
  srv->cv_work.notify_all();
  if (srv->worker.joinable()) srv->worker.join();
  (void)hipSetDevice(srv->ctx->device);
  for (Slot& s : srv->slots) {
    if (s.stream) (void)hipStreamSynchronize(s.stream);
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.stream) (void)hipStreamDestroy(s.stream);
    if (s.pinned) (void)hipHostFree(s.pinned);
  }
  delete srv;
  return HEU_OK;
}

}  // extern "C"
