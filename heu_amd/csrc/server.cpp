// server.cpp -- the request server (heu_server_* in include/heu_paillier.h).
//
// One scheduler thread per server.  Work is budgeted in device rounds: a
// round is the number of work-list units a kernel shape runs at once, and a
// unit's latency is fixed by its op, so a batch holding less than a round of a
// shape leaves the chip idle for as long as a full one takes.  Each turn the
// scheduler retires finished batches (stream events), then, while a stream
// slot is free, builds the next batch:
//   - order the active requests by remaining work, shortest first, aged (the
//     key shrinks the longer a request waits); requests that have waited
//     longer than urgent_ms go before all others, oldest first, so a long
//     request is not held back for good by a stream of short ones;
//   - the batch takes the kernel shape of the first request whose shape has
//     room (each shape may hold batch_rounds rounds in flight across all slots,
//     decryption only a share of that, which keeps CUs free for the short ops
//     that arrive during a decryption's long units);
//   - round-robin over the requests of that shape, each pass carving the next
//     contiguous range of up to a quantum (1/parts_per_round of a round) from
//     every request, until the room is used (a sum or dot is taken whole); an
//     urgent request at the head of the order takes all the room it can use;
//   - trim to the longest prefix whose plan fits the largest free block of the
//     workspace arena (first growing the arena by a new device segment while
//     hipMemGetInfo shows HBM to spare), plan it into that block, copy its work
//     lists up from pinned staging and enqueue it on the slot's stream.
// A long request therefore advances a slice per batch while short requests
// that arrive meanwhile join the next batch; it completes when its last
// element does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
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

// Waiting this long halves a request's remaining-work priority key.
constexpr double kAgingMs = 100.0;
// Share of a round that decryption batches may hold at once: a decryption unit
// runs for tens of milliseconds, so the rest of the chip is kept for the short
// ops that arrive meanwhile.
constexpr double kDecRoundShare = 0.875;

struct Request {
  heu_task task;
  std::vector<int64_t> offsets;                   // owned copy for SUM / DOT
  const heu_hip::WorklistShape* shape = nullptr;  // the kernel shape its units run on
  long round = 1;                                 // units that shape runs at once
  long gpu = 1;                                   // elements per unit
  long factor = 1;                                // units per group of elements (2 CRT halves)
  bool whole = false;                             // SUM / DOT: one part
  long whole_units = 0;
  long n = 0, admitted = 0, finished = 0;         // elements (1 for a whole request)
  long inflight = 0;                              // parts admitted to a slot and not yet retired
  double elem_cost = 1.0;                         // Montgomery products mod n^2 per element (whole: per request)
  size_t sa = 0, sb = 0, so = 0;                  // bytes per element of a, b and out
  Clock::time_point t_submit, t_done;
  int state = kQueued;
  std::string error;

  long units_for(long elems) const { return whole ? whole_units : (elems + gpu - 1) / gpu * factor; }
  long remaining_units() const { return whole ? (admitted ? 0 : whole_units) : units_for(n - admitted); }
  // Work left, in Montgomery products mod a 2048-bit key's n^2: what orders
  // requests.  Work units differ in cost by three orders of magnitude between
  // ops (a round of ct+ct units takes ~0.1 ms, one of decryption units ~30 ms),
  // so units or rounds would rank a long cheap vector behind short expensive ones.
  double remaining_cost() const { return whole ? (admitted ? 0.0 : elem_cost) : (double)(n - admitted) * elem_cost; }
  heu_task part(long begin, long count) const {
    if (whole) return task;
    heu_task p = task;
    p.count = (size_t)count;
    p.a = task.a ? static_cast<const char*>(task.a) + begin * sa : nullptr;
    p.b = task.b ? static_cast<const char*>(task.b) + begin * sb : nullptr;
    p.out = static_cast<char*>(task.out) + begin * so;
    return p;
  }
};

// Montgomery products mod n^2 per element of each op, for the key's table
// window: encryption and Randomize take one table product per window and one
// more; ct+ct and ct+pt two; ct x int32 about 46 (30 squarings, 16 products)
// with the inversion of negative scalars amortized; a dot term that and one
// more; a sum term one.  A decryption runs two exponentiations of about 1210
// products mod p^2, each a quarter of one mod n^2.
double op_cost(const heu_key* k, int op) {
  switch (op) {
    case HEU_OP_ENCRYPT_I64:
    case HEU_OP_RANDOMIZE: return (double)k->windows + 1.0;
    case HEU_OP_ADD:
    case HEU_OP_ADD_I64: return 2.0;
    case HEU_OP_MUL_I64: return 46.0;
    case HEU_OP_DOT_I64: return 47.0;
    case HEU_OP_SUM: return 1.0;
    case HEU_OP_DECRYPT: return 2.0 * 1210.0 / 4.0;
    default: return 1.0;
  }
}

// Fills in the request's shape and geometry (caller holds ctx->mu).
void describe(heu_ctx* ctx, Request& r) {
  const heu_task& t = r.task;
  const bool dec = t.op == HEU_OP_DECRYPT;
  // products mod a larger n^2 cost the square of the size ratio
  const double size = (double)t.key->ct_words / 128.0;
  r.elem_cost = op_cost(t.key, t.op) * size * size;
  r.shape = dec ? &t.key->dshape().wl : &t.key->ops->wl_n;
  r.round = std::max<long>(1, worklist_round_units(ctx, r.shape));
  r.gpu = r.shape->groups_per_unit;
  r.factor = dec ? 2 : 1;
  const bool seg = t.op == HEU_OP_SUM || t.op == HEU_OP_DOT_I64;
  r.whole = seg || t.count == 0;
  if (r.whole) {
    const long n = seg && t.segments ? (long)(t.offsets[t.segments] - t.offsets[0]) : 0;
    const long chunk = reduce_chunk0(ctx, t.key->ops, t.op == HEU_OP_DOT_I64, n);  // level 0 of the reduction
    r.whole_units = (n / chunk + (long)t.segments + r.gpu - 1) / r.gpu;
    r.n = 1;
    r.elem_cost *= (double)std::max(1L, n);
    return;
  }
  r.n = (long)t.count;
  const size_t ct = (size_t)t.key->ct_words * 4;
  r.sa = r.sb = r.so = ct;
  switch (t.op) {
    case HEU_OP_ENCRYPT_I64: r.sa = 8; r.sb = (size_t)t.aux * 4; break;
    case HEU_OP_RANDOMIZE: r.sb = (size_t)t.aux * 4; break;
    case HEU_OP_ADD_I64:
    case HEU_OP_MUL_I64: r.sb = 8; break;
    case HEU_OP_DECRYPT: r.sb = 0; r.so = (size_t)t.aux * 4; break;
    default: break;
  }
}

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  const heu_hip::WorklistShape* shape = nullptr;  // every batch runs one kernel shape
  long units = 0;
  size_t off = 0, bytes = 0;
  std::vector<std::pair<uint64_t, long>> parts;  // (ticket, elements)
  char* pinned = nullptr;                         // staging of the batch's work lists
  size_t pinned_cap = 0;
  bool drained = false;                           // the stream was synchronized: retire without the event
};

// First-fit allocator over device segments (separate allocations) with
// coalescing.  Offsets are virtual: segment i covers [base_i, base_i + size_i),
// and a gap between segments keeps free blocks from merging across them, so a
// block always lies inside one allocation.
struct Arena {
  static constexpr size_t kGap = 256;
  std::map<size_t, size_t> free_;  // offset -> bytes
  std::map<size_t, std::unique_ptr<DevBuf>> segs_;  // base -> allocation
  size_t end_ = 0, capacity = 0, largest_seg = 0;
  int add(size_t bytes) {
    auto seg = std::make_unique<DevBuf>();
    const int rc = seg->alloc(bytes);
    if (rc) return rc;
    const size_t base = end_;
    free_[base] = bytes;
    segs_[base] = std::move(seg);
    end_ = base + bytes + kGap;
    capacity += bytes;
    largest_seg = std::max(largest_seg, bytes);
    return HEU_OK;
  }
  char* ptr(size_t off) const {
    auto it = std::prev(segs_.upper_bound(off));
    return it->second->as<char>() + (off - it->first);
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

// Device memory the server leaves free when it grows its workspace: 5% of the
// device, at least 2 GiB (keys, tables and other users of the device).
size_t hbm_reserve(size_t total) { return std::max<size_t>((size_t)2 << 30, total / 20); }

size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

// One part of a batch under construction.
struct Pick {
  uint64_t ticket;
  long begin, count;  // elements (whole requests: 0, 1)
};

}  // namespace

struct heu_server {
  heu_ctx* ctx = nullptr;
  heu_server_config cfg{};
  Arena arena;
  size_t in_use = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<uint64_t> pending;  // requests with elements not yet admitted
  std::unordered_map<uint64_t, Request> reqs;
  uint64_t next_ticket = 1;
  std::vector<Slot> slots;
  std::thread worker;
  bool stop = false;
  heu_server_counters st{};
  // Failure injection for tests (HEU_SERVER_FAULT_BATCH=k): the k-th batch is
  // reported as not enqueued after its launches were issued, so its requests
  // fail while their kernels still run.
  uint64_t fault_batch = 0;

  void run();
  bool grow(size_t need);
  bool retire();
  bool admit(std::unique_lock<std::mutex>& lk);
  bool select(std::vector<Pick>* picks, const heu_hip::WorklistShape** shape);
  // A failed request is reported (and forgotten by heu_server_wait) only once
  // none of its parts is still on the device: until then the caller's buffers
  // may still be read or written by a running batch.
  static bool settled(const Request& r) { return r.state == kDone || (r.state == kFailed && r.inflight == 0); }
  void fail_request(uint64_t t, const std::string& why) {
    Request& r = reqs[t];
    if (r.state == kFailed) return;
    r.state = kFailed;
    r.error = why;
    r.t_done = Clock::now();
    ++st.failed;
    pending.erase(std::remove(pending.begin(), pending.end(), t), pending.end());
  }
};

// Retires finished slots; returns true if any finished.
bool heu_server::retire() {
  bool any = false;
  for (Slot& s : slots) {
    if (!s.busy || (!s.drained && hipEventQuery(s.done) != hipSuccess)) continue;
    std::lock_guard<std::mutex> lk(mu);
    const auto now = Clock::now();
    for (const auto& pt : s.parts) {
      auto it = reqs.find(pt.first);
      if (it == reqs.end()) continue;
      Request& r = it->second;
      --r.inflight;
      if (r.state == kFailed) continue;
      r.finished += pt.second;
      if (r.finished == r.n) {
        r.state = kDone;
        r.t_done = now;
        ++st.completed;
      }
    }
    arena.give(s.off, s.bytes);
    in_use -= s.bytes;
    s.busy = false;
    s.drained = false;
    s.parts.clear();
    any = true;
  }
  if (any) cv_done.notify_all();
  return any;
}

// Picks the parts of the next batch (called with mu held).  False if no
// request's shape has room.
bool heu_server::select(std::vector<Pick>* picks, const heu_hip::WorklistShape** shape_out) {
  // key: remaining rounds, aged; urgent requests sort first by their (negated) age
  std::vector<std::pair<double, uint64_t>> order;
  const auto now = Clock::now();
  for (uint64_t t : pending) {
    const Request& r = reqs[t];
    const double age_ms = std::chrono::duration<double, std::milli>(now - r.t_submit).count();
    if (age_ms > (double)cfg.urgent_ms)
      order.emplace_back(-1.0 - age_ms, t);
    else
      order.emplace_back(r.remaining_cost() / (1.0 + age_ms / kAgingMs), t);
  }
  std::stable_sort(order.begin(), order.end(),
                   [](const std::pair<double, uint64_t>& x, const std::pair<double, uint64_t>& y) {
                     return x.first < y.first;
                   });
  auto inflight = [&](const heu_hip::WorklistShape* sh) {
    long u = 0;
    for (const Slot& s : slots)
      if (s.busy && s.shape == sh) u += s.units;
    return u;
  };
  auto cap = [&](const Request& r) {
    const bool dec = r.task.op == HEU_OP_DECRYPT;
    return std::max<long>(1, (long)((double)r.round * cfg.batch_rounds * (dec ? kDecRoundShare : 1.0)));
  };
  const heu_hip::WorklistShape* shape = nullptr;
  long room = 0, quantum = 1;
  bool idle = false;
  for (const auto& ot : order) {
    const Request& r = reqs[ot.second];
    const long busy = inflight(r.shape);
    if (busy < cap(r)) {
      shape = r.shape;
      room = cap(r) - busy;
      idle = busy == 0;
      quantum = std::max<long>(r.factor, r.round / (long)cfg.parts_per_round);
      break;
    }
  }
  if (!shape) return false;
  std::unordered_map<uint64_t, size_t> mine;  // ticket -> index in picks (one merged range each)
  long used = 0;
  for (bool added = true; added && picks->size() < cfg.max_batch_tasks;) {
    added = false;
    for (const auto& ot : order) {
      const uint64_t t = ot.second;
      const Request& r = reqs[t];
      if (r.shape != shape) continue;
      auto at = mine.find(t);
      const bool first = picks->empty() && idle;
      const bool urgent = picks->empty() && ot.first < 0.0;
      if (r.whole) {
        if (r.admitted || at != mine.end()) continue;
        if (!first && used + r.whole_units > room) continue;
        mine[t] = picks->size();
        picks->push_back(Pick{t, 0, 1});
        used += r.whole_units;
        added = true;
      } else {
        const long have = at == mine.end() ? 0 : (*picks)[at->second].count;
        const long rem = r.n - r.admitted - have;
        if (rem <= 0) continue;
        const long avail = urgent  ? (first ? std::max(quantum, room - used) : room - used)
                           : first ? quantum
                                   : std::min(quantum, room - used);
        const long elems = std::min(rem, avail / r.factor * r.gpu);
        if (elems <= 0) continue;
        const long before = r.units_for(have);
        if (at == mine.end()) {
          mine[t] = picks->size();
          picks->push_back(Pick{t, r.admitted, elems});
        } else {
          (*picks)[at->second].count += elems;
        }
        used += r.units_for(have + elems) - before;
        added = true;
      }
      if (picks->size() >= cfg.max_batch_tasks) break;
    }
  }
  *shape_out = shape;
  return !picks->empty();
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
  std::vector<Pick> picks;
  const heu_hip::WorklistShape* shape = nullptr;
  if (!select(&picks, &shape)) return false;
  std::vector<heu_task> tasks;
  for (const Pick& p : picks) tasks.push_back(reqs[p.ticket].part(p.begin, p.count));
  const size_t n = tasks.size();
  size_t free_bytes = arena.largest();
  lk.unlock();
  // longest admissible prefix: its plan fits the largest free block
  auto size_of = [&](size_t k, PlanOut* p) {
    std::lock_guard<std::mutex> cl(ctx->mu);
    return plan_batch(ctx, tasks.data(), k, nullptr, p);
  };
  auto fits = [&](const PlanOut& p) { return round256(p.bytes) <= free_bytes; };
  size_t k = n;
  PlanOut p;
  int rc = size_of(k, &p);
  if (rc == HEU_OK && !fits(p)) {
    // admission by free device memory: a batch that does not fit grows the
    // workspace while the device has HBM to spare
    lk.lock();
    if (grow(round256(p.bytes))) free_bytes = arena.largest();
    lk.unlock();
  }
  if (rc == HEU_OK && !fits(p)) {
    size_t lo = 0, hi = n;  // fits(lo) holds (or lo == 0), fits(hi) does not
    while (hi - lo > 1) {
      const size_t mid = (lo + hi) / 2;
      PlanOut q;
      if (size_of(mid, &q) == HEU_OK && fits(q)) lo = mid;
      else hi = mid;
    }
    k = lo;
    if (k) rc = size_of(k, &p);
  }
  lk.lock();
  if (rc != HEU_OK || k == 0) {
    // the first part alone: fail its request if it can never fit, else wait for room
    PlanOut one;
    lk.unlock();
    const int rc1 = size_of(1, &one);
    lk.lock();
    if (rc1 == HEU_OK && round256(one.bytes) > arena.largest_seg) (void)grow(round256(one.bytes));
    if (rc1 != HEU_OK || round256(one.bytes) > arena.largest_seg) {
      fail_request(picks[0].ticket,
                   rc1 != HEU_OK ? heu_last_error() : "request needs more workspace than the server has");
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
  long units = 0;
  slot->parts.clear();
  for (size_t i = 0; i < k; ++i) {
    Request& r = reqs[picks[i].ticket];
    r.state = kRunning;
    units += r.units_for(picks[i].count);
    r.admitted += picks[i].count;
    ++r.inflight;
    slot->parts.emplace_back(picks[i].ticket, picks[i].count);
    if (r.admitted == r.n) pending.erase(std::find(pending.begin(), pending.end(), picks[i].ticket));
  }
  slot->busy = true;
  slot->shape = shape;
  slot->units = units;
  slot->off = off;
  slot->bytes = bytes;
  const uint64_t batch_no = ++st.batches;
  st.max_batch_requests = std::max<uint64_t>(st.max_batch_requests, k);
  lk.unlock();
  PlanOut real;
  char* base = arena.ptr(off);
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
    (void)hipGetLastError();  // only this batch's launches count below
    for (auto& s : real.steps) s(slot->stream);
    if (hipGetLastError() != hipSuccess) rc = HEU_EDEVICE;
    if (fault_batch && batch_no == fault_batch) rc = HEU_EDEVICE;
  }
  const bool recorded = hipEventRecord(slot->done, slot->stream) == hipSuccess;
  if (!recorded) {
    // without the event, retire() cannot see when the device is done with the
    // parts: wait for the stream here instead
    rc = HEU_EDEVICE;
    (void)hipStreamSynchronize(slot->stream);
  }
  lk.lock();
  if (rc != HEU_OK) {
    // nothing (or not all) was enqueued: the requests fail.  The slot keeps
    // its parts and retires as usual, which releases each part's in-flight
    // count; heu_server_wait reports a failed request only after that.
    for (const auto& pt : slot->parts) fail_request(pt.first, "could not enqueue the batch");
    slot->drained = !recorded;
  }
  return true;
}

// Adds a workspace segment of at least `need` bytes (called with mu held):
// half the initial workspace when the device can spare it, else exactly
// `need`, and only while hipMemGetInfo leaves hbm_reserve() free and the
// workspace stays within max_workspace_bytes.
bool heu_server::grow(size_t need) {
  size_t freeb = 0, totalb = 0;
  if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) return false;
  const size_t reserve = hbm_reserve(totalb);
  const size_t cap = cfg.max_workspace_bytes;
  const size_t room = !cap ? SIZE_MAX : (cap > arena.capacity ? (cap - arena.capacity) & ~(size_t)255 : 0);
  size_t bytes = std::min(round256(std::max(need, cfg.workspace_bytes / 2)), room);
  if (bytes < need || freeb < reserve + bytes) bytes = need;
  if (bytes > room || freeb < reserve + bytes) return false;
  if (arena.add(bytes) != HEU_OK) return false;
  st.workspace_bytes = arena.capacity;
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

// Waits for and frees every slot's stream, event and pinned staging.
static void release_slots(heu_server* srv) {
  for (Slot& s : srv->slots) {
    if (s.stream) (void)hipStreamSynchronize(s.stream);
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.stream) (void)hipStreamDestroy(s.stream);
    if (s.pinned) (void)hipHostFree(s.pinned);
    s.stream = nullptr;
    s.done = nullptr;
    s.pinned = nullptr;
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
  if (!srv->cfg.batch_rounds) srv->cfg.batch_rounds = 1;
  if (!srv->cfg.parts_per_round) srv->cfg.parts_per_round = 8;
  if (!srv->cfg.urgent_ms) srv->cfg.urgent_ms = 150;
  if (srv->cfg.max_inflight > 8) return fail(HEU_EINVAL, "max_inflight must be at most 8");
  if (const char* f = std::getenv("HEU_SERVER_FAULT_BATCH")) srv->fault_batch = std::strtoull(f, nullptr, 10);
  HIP_TRY(hipSetDevice(ctx->device));
  int rc = srv->arena.add(round256(srv->cfg.workspace_bytes));
  if (rc) return rc;
  srv->st.workspace_bytes = srv->arena.capacity;
  srv->slots.resize(srv->cfg.max_inflight);
  for (Slot& s : srv->slots) {
    hipError_t e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
    if (e != hipSuccess) {
      release_slots(srv.get());
      return fail(HEU_EDEVICE, "server stream or event: %s", hipGetErrorString(e));
    }
  }
  ctx_retain(ctx);  // the server holds the context
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
    describe(srv->ctx, r);
  }
  std::lock_guard<std::mutex> lk(srv->mu);
  if (srv->stop) return fail(HEU_EINVAL, "server is stopping");
  const uint64_t t = srv->next_ticket++;
  r.t_submit = Clock::now();
  Request& stored = srv->reqs[t];
  stored = std::move(r);
  if (!stored.offsets.empty()) stored.task.offsets = stored.offsets.data();  // the vector moved
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
    return i == srv->reqs.end() || heu_server::settled(i->second);
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
  release_slots(srv);
  heu_ctx* ctx = srv->ctx;
  delete srv;
  ctx_release(ctx);
  return HEU_OK;
}

}  // extern "C"
