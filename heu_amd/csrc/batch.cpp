// batch.cpp -- multi-task batches (heu_batch_* in include/heu_paillier.h).
//
// A batch is planned once: every task's elementwise work becomes items of a
// persistent work-list launch (one launch per key size and modulus shape),
// and the dependent work that follows (CRT recombination, batch inversion of
// negative-scalar powers, upper reduction levels, segment finishes) becomes a
// list of steps in task order.  The planner runs twice: a sizing pass that only
// counts workspace bytes, then, with the workspace allocated, the pass that
// produces device pointers, uploads the work lists and records the steps.
// Within a stage, each work-list launch and the steps that read its results
// form a lane; lanes of one stage run on streams of their own (forked from the
// context's stream and joined back), so one shape's tail of long units
// overlaps another shape's work.  A run replays the stages, or launches the
// hipGraph they were captured into (the fork and join become graph edges).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <utility>
#include <vector>

#include "engine_internal.hpp"

using heu_hip::Chunk;
using heu_hip::WorkItem;
using heu_hip::WorklistShape;
using namespace heu_engine;

struct heu_batch {
  heu_ctx* ctx = nullptr;
  DevBuf ws;  // workspace: work lists, chunk lists, partials, window tables
  std::vector<PlanStage> stages;
  uint32_t flags = 0;
  uint64_t launches = 0, units = 0;
  std::vector<hipStream_t> side;   // lanes 1.. of a stage (lane 0 runs on the main stream)
  std::vector<hipEvent_t> events;  // per multi-lane stage: one fork, one join per side lane
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

typedef std::function<void(hipStream_t)> Step;

// One pass of the planner.  alloc() hands out 256-byte aligned pieces of the
// workspace; in the sizing pass base is null and nothing is recorded.
struct Builder {
  char* base = nullptr;
  size_t used = 0;
  std::vector<std::pair<size_t, std::vector<char>>> metas;  // host images to upload
  std::vector<Step>* cur = nullptr;                         // the lane a task's steps join
  std::vector<Step> follow;                                 // steps after every lane of the stage
  std::vector<Step> steps;                                  // every stage, in order
  std::vector<PlanStage> stages;
  uint64_t launches = 0;

  template <class T>
  T* alloc(size_t bytes) {
    const size_t off = used;
    used += (bytes + 255) & ~(size_t)255;
    return reinterpret_cast<T*>(reinterpret_cast<uintptr_t>(base) + off);
  }
  template <class T>
  T* meta(const std::vector<T>& v) {
    const size_t off = used;
    T* p = alloc<T>(v.size() * sizeof(T));
    if (base && !v.empty()) {
      const char* c = reinterpret_cast<const char*>(v.data());
      metas.emplace_back(off, std::vector<char>(c, c + v.size() * sizeof(T)));
    }
    return p;
  }
  void then(Step s, uint64_t nlaunch = 1) {
    launches += nlaunch;
    if (base) (cur ? *cur : follow).push_back(std::move(s));
  }
};

// Items of one persistent launch, with a per-unit cost used to hand out the
// expensive units first.
struct WorkList {
  const WorklistShape* shape = nullptr;
  int key_bits = 0;
  bool dec = false;
  std::vector<std::pair<double, WorkItem>> items;
  std::vector<Step> follow;  // steps reading this launch's results, in task order
};

// Montgomery products per unit (one element or chunk per group), for ordering.
double unit_cost(int op) {
  switch (op) {
    case heu_hip::kWorkDot0: return 8 * 45;
    case heu_hip::kWorkDecHalf: return 1210;
    case heu_hip::kWorkMul: return 80;
    case heu_hip::kWorkSum0: return 32;
    case heu_hip::kWorkEncrypt:
    case heu_hip::kWorkRandomize: return 28;
    default: return 2;
  }
}

WorkItem make_item(int op, long count, const heu_key* key) {
  WorkItem it;
  std::memset(&it, 0, sizeof(it));
  it.op = op;
  it.count = count;
  it.pub = key->pub();
  it.err = key->ctx->d_err;
  return it;
}

// Batch inversion (Montgomery's trick) of the elements of `in` into `out`
// (may alias), as steps: all of them, or (`selected`) those flagged in `sel`.
// The choice is the caller's, not sel's value: in the sizing pass workspace
// pointers are offsets, and the first one is null.
void plan_invert(Builder& b, heu_key* key, const uint32_t* in, const uint8_t* sel, bool selected, long count,
                 uint32_t* out) {
  if (count <= 0) return;
  const Ops* o = key->ops;
  const ModArgs mod = key->mod_n2();
  const size_t NLB = (size_t)key->nl_n * 4;
  const long chunk0 = inv_chunk0(key->ctx, o, count);
  const std::vector<long> cnt = inv_levels(count, chunk0);
  const int levels = (int)cnt.size() - 1;
  // selected elements (in place) run from a packed index list
  heu_hip::InvSel s0{selected ? sel : nullptr, nullptr, nullptr};
  if (selected && in == out) {
    uint32_t* idx = b.alloc<uint32_t>((size_t)count * 4);
    uint32_t* nsel = b.alloc<uint32_t>(4);
    s0 = heu_hip::InvSel{nullptr, idx, nsel};
    b.then([=](hipStream_t st) { o->inv_compact(st, sel, count, idx, nsel); }, 2);
  }
  const heu_hip::InvSel none{nullptr, nullptr, nullptr};
  std::vector<uint32_t*> pre(levels + 1, nullptr), tot(levels + 1), inv(levels + 1);
  for (int l = 0; l <= levels; ++l) {
    if (l < levels) pre[l] = b.alloc<uint32_t>(cnt[l] * NLB);
    tot[l] = b.alloc<uint32_t>(cnt[l] * NLB);
    inv[l] = b.alloc<uint32_t>(cnt[l] * NLB);
  }
  for (int l = 0; l < levels; ++l) {
    const uint32_t* src = l == 0 ? in : tot[l];
    const heu_hip::InvSel s = l == 0 ? s0 : none;
    uint32_t *p = pre[l], *t = tot[l + 1];
    const long n = cnt[l];
    const int chunk = inv_chunk(l, chunk0);
    b.then([=](hipStream_t st) { o->inv_prefix(st, mod, l == 0, src, s, n, chunk, p, t); });
  }
  const uint32_t* n2w = key->n2_words.as<uint32_t>();
  uint32_t *top_inv = inv[levels], *top_tot = tot[levels], *err = key->ctx->d_err;
  b.then([=](hipStream_t st) {
    (void)hipMemcpyAsync(top_inv, top_tot, NLB, hipMemcpyDeviceToDevice, st);
    o->inv_single(st, mod, n2w, top_inv, err);
  }, 2);
  for (int l = levels - 1; l >= 0; --l) {
    const uint32_t* src = l == 0 ? in : tot[l];
    const heu_hip::InvSel s = l == 0 ? s0 : none;
    uint32_t* dst = l == 0 ? out : inv[l];
    uint32_t *p = pre[l], *ti = inv[l + 1];
    const long n = cnt[l];
    const int chunk = inv_chunk(l, chunk0);
    b.then([=](hipStream_t st) { o->inv_down(st, mod, l == 0, src, s, n, chunk, p, ti, dst); });
  }
}

// Segmented SUM / DOT: level 0 as work items, upper levels and the finish as
// steps (numpy/evaluator.cc:355-373 Sum, :209-292 MatMul dot).
void plan_segmented(Builder& b, heu_key* key, const heu_task& t, WorkList& wl) {
  const Ops* o = key->ops;
  const ModArgs mod = key->mod_n2();
  const size_t NLB = (size_t)key->nl_n * 4;
  const bool dot = t.op == HEU_OP_DOT_I64;
  const size_t segs = t.segments;
  std::vector<long> start(segs), len(segs);
  long total = 0;
  for (size_t s = 0; s < segs; ++s) {
    start[s] = t.offsets[s];
    len[s] = t.offsets[s + 1] - t.offsets[s];
    total += len[s];
  }
  struct Level {
    std::vector<Chunk> ch;
    std::vector<long> first, nch;
    Chunk* d_ch = nullptr;
    uint32_t *pos = nullptr, *neg = nullptr;
  };
  std::vector<Level> lv(1);
  lv[0].ch = make_chunks(start, len, reduce_chunk0(key->ctx, o, dot, total), &lv[0].first, &lv[0].nch);
  for (;;) {
    const Level& prev = lv.back();
    bool done = true;
    for (size_t s = 0; s < segs; ++s) done &= prev.nch[s] == 1;
    if (done) break;
    Level n;
    n.ch = make_chunks(prev.first, prev.nch, kReduceChunkUpper, &n.first, &n.nch);
    lv.push_back(std::move(n));
  }
  for (auto& l : lv) {
    l.d_ch = b.meta(l.ch);
    l.pos = b.alloc<uint32_t>(l.ch.size() * NLB);
    if (dot) l.neg = b.alloc<uint32_t>(l.ch.size() * NLB);
  }
  WorkItem it = make_item(dot ? heu_hip::kWorkDot0 : heu_hip::kWorkSum0, (long)lv[0].ch.size(), key);
  it.a = t.a;
  it.b = t.b;
  it.c = lv[0].d_ch;
  it.out = lv[0].pos;
  it.out2 = lv[0].neg;
  wl.items.emplace_back(unit_cost(it.op), it);
  for (size_t i = 1; i < lv.size(); ++i) {
    const uint32_t *ppos = lv[i - 1].pos, *pneg = lv[i - 1].neg;
    uint32_t *cpos = lv[i].pos, *cneg = lv[i].neg;
    const Chunk* ch = lv[i].d_ch;
    const long nch = (long)lv[i].ch.size();
    if (dot) {
      b.then([=](hipStream_t st) {
        o->reduce_mont(st, mod, ppos, ch, nch, cpos);
        o->reduce_mont(st, mod, pneg, ch, nch, cneg);
      }, 2);
    } else {
      b.then([=](hipStream_t st) { o->reduce_sum(st, mod, false, ppos, ch, nch, cpos, nullptr); });
    }
  }
  // the last level holds one partial per segment, in segment order
  const Level& top = lv.back();
  uint32_t* out = static_cast<uint32_t*>(t.out);
  if (!dot) {
    uint32_t* pos = top.pos;
    b.then([=](hipStream_t st) { o->seg_finish(st, mod, pos, (long)segs, out); });
  } else {
    uint32_t* nz = b.alloc<uint32_t>(segs * (size_t)key->ct_words * 4);
    uint32_t *pos = top.pos, *neg = top.neg;
    b.then([=](hipStream_t st) { o->seg_finish(st, mod, neg, (long)segs, nz); });
    plan_invert(b, key, nz, nullptr, false, (long)segs, nz);
    b.then([=](hipStream_t st) { o->mont_times_z(st, mod, pos, nz, (long)segs, out); });
  }
}

int check_task(const heu_ctx* ctx, const heu_task& t, size_t i) {
  if (t.op == HEU_OP_BARRIER) return HEU_OK;
  if (!t.key || !t.key->ops) return fail(HEU_EINVAL, "task %zu: null key", i);
  if (t.key->ctx != ctx) return fail(HEU_EINVAL, "task %zu: key belongs to another context", i);
  const bool seg = t.op == HEU_OP_SUM || t.op == HEU_OP_DOT_I64;
  if (t.op < HEU_OP_ENCRYPT_I64 || t.op > HEU_OP_DOT_I64) return fail(HEU_EINVAL, "task %zu: unknown op %d", i, t.op);
  if (seg) {
    if (t.segments == 0) return HEU_OK;
    if (!t.offsets || !t.out) return fail(HEU_EINVAL, "task %zu: null offsets or output", i);
    for (size_t s = 0; s < t.segments; ++s)
      if (t.offsets[s + 1] < t.offsets[s]) return fail(HEU_EINVAL, "task %zu: offsets must be non-decreasing", i);
    if (t.offsets[t.segments] > t.offsets[0] && !t.a) return fail(HEU_EINVAL, "task %zu: null input", i);
    if (t.op == HEU_OP_DOT_I64 && t.offsets[t.segments] > t.offsets[0] && !t.b)
      return fail(HEU_EINVAL, "task %zu: null weights", i);
    return HEU_OK;
  }
  if (t.count == 0) return HEU_OK;
  if (!t.out) return fail(HEU_EINVAL, "task %zu: null output", i);
  const bool need_a = t.op != HEU_OP_RANDOMIZE, need_b = t.op != HEU_OP_DECRYPT;
  if ((need_a && !t.a) || (need_b && !t.b)) return fail(HEU_EINVAL, "task %zu: null input", i);
  if ((t.op == HEU_OP_ENCRYPT_I64 || t.op == HEU_OP_RANDOMIZE || t.op == HEU_OP_DECRYPT) && t.aux == 0)
    return fail(HEU_EINVAL, "task %zu: aux (limbs) must be positive", i);
  if (t.op == HEU_OP_DECRYPT && !t.key->has_sk) return fail(HEU_EKEY, "task %zu: key has no secret part", i);
  return HEU_OK;
}

struct ShapeKey {
  const WorklistShape* shape;
  bool operator<(const ShapeKey& o) const { return shape < o.shape; }
};

// Resident workgroups per CU of a work-list kernel (occupancy query, cached).
int resident_per_cu(heu_ctx* ctx, const WorklistShape* s) {
  auto it = ctx->wl_resident.find(s);
  if (it != ctx->wl_resident.end()) return it->second;
  const int r = s->resident();
  ctx->wl_resident[s] = r;
  return r;
}

// One stage: the tasks [begin, end), none reading another's output.
int plan_stage(Builder& b, heu_ctx* ctx, const heu_task* tasks, size_t ntasks, uint64_t* units_out) {
  std::map<ShapeKey, WorkList> lists;
  auto list_for = [&](heu_key* key, bool dec) -> WorkList& {
    const WorklistShape* s = dec ? &key->dshape().wl : &key->ops->wl_n;
    WorkList& wl = lists[ShapeKey{s}];
    wl.shape = s;
    wl.key_bits = key->key_bits;
    wl.dec = dec;
    return wl;
  };
  for (size_t ti = 0; ti < ntasks; ++ti) {
    const heu_task& t = tasks[ti];
    heu_key* key = t.key;
    const long n = (long)t.count;
    const Ops* o = key->ops;
    switch (t.op) {
      case HEU_OP_ENCRYPT_I64:
      case HEU_OP_RANDOMIZE:
      case HEU_OP_ADD:
      case HEU_OP_ADD_I64: {
        if (n == 0) break;
        static const int kOp[] = {0, heu_hip::kWorkEncrypt, heu_hip::kWorkRandomize, heu_hip::kWorkAdd,
                                  heu_hip::kWorkAddPt};
        WorkItem it = make_item(kOp[t.op], n, key);
        it.a = t.a;
        it.b = t.b;
        it.aux = (int)t.aux;
        it.out = t.out;
        list_for(key, false).items.emplace_back(unit_cost(it.op), it);
        break;
      }
      case HEU_OP_MUL_I64: {
        if (n == 0) break;
        uint8_t* neg = b.alloc<uint8_t>((size_t)n);
        WorkItem it = make_item(heu_hip::kWorkMul, n, key);
        it.a = t.a;
        it.b = t.b;
        it.out = t.out;
        it.neg = neg;
        WorkList& wl = list_for(key, false);
        wl.items.emplace_back(unit_cost(it.op), it);
        uint32_t* out = static_cast<uint32_t*>(t.out);
        b.cur = &wl.follow;
        plan_invert(b, key, out, neg, true, n, out);
        b.cur = nullptr;
        break;
      }
      case HEU_OP_DECRYPT: {
        if (n == 0) break;
        const size_t HW = (size_t)key->ct_words / 2;
        uint32_t* halves = b.alloc<uint32_t>(2 * (size_t)n * HW * 4);
        const heu_hip::DecArgs D = key->dec();
        for (int h = 0; h < 2; ++h) {
          WorkItem it = make_item(heu_hip::kWorkDecHalf, n, key);
          it.pub.n2 = h ? D.q2 : D.p2;
          it.a = t.a;
          it.c = h ? D.sched_q : D.sched_p;
          it.aux = h ? D.nsched_q : D.nsched_p;
          it.out = halves + (size_t)h * n * HW;
          list_for(key, true).items.emplace_back(unit_cost(it.op), it);
        }
        const heu_hip::CombArgs C = key->comb();
        uint32_t* out = static_cast<uint32_t*>(t.out);
        const int ow = (int)t.aux;
        b.cur = &list_for(key, true).follow;
        b.then([=](hipStream_t st) { o->combine(st, C, halves, n, out, ow, nullptr); });
        b.cur = nullptr;
        break;
      }
      case HEU_OP_SUM:
      case HEU_OP_DOT_I64: {
        if (t.segments == 0) break;
        WorkList& wl = list_for(key, false);
        b.cur = &wl.follow;
        plan_segmented(b, key, t, wl);
        b.cur = nullptr;
        break;
      }
      default:
        break;
    }
  }
  // one persistent launch per shape, expensive units first
  int cus = 0;
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  std::vector<std::pair<double, std::vector<Step>>> lanes;  // (estimated work, steps)
  uint64_t units = 0;
  for (auto& kv : lists) {
    WorkList& wl = kv.second;
    if (wl.items.empty()) continue;
    std::stable_sort(wl.items.begin(), wl.items.end(),
                     [](const std::pair<double, WorkItem>& x, const std::pair<double, WorkItem>& y) {
                       return x.first > y.first;
                     });
    const long gpu = wl.shape->groups_per_unit;
    std::vector<WorkItem> items;
    long u0 = 0;
    double work = 0;
    // unit costs count products mod n^2 (mod p^2 for decryption, a quarter of
    // the work); a 3072-bit product is 2.25 times a 2048-bit one
    const double scale = (wl.key_bits == 3072 ? 2.25 : 1.0) * (wl.dec ? 0.25 : 1.0);
    for (auto& ci : wl.items) {
      WorkItem it = ci.second;
      it.unit0 = u0;
      const long nu = (it.count + gpu - 1) / gpu;
      u0 += nu;
      work += ci.first * (double)nu * (double)wl.shape->groups_per_unit * scale;
      items.push_back(it);
    }
    if (u0 >= (long)UINT32_MAX) return fail(HEU_EINVAL, "batch too large for one work list");
    const uint32_t total = (uint32_t)u0;
    const uint32_t mult = ctx->grid_mult[wl.key_bits == 3072][wl.dec];
    const long cap = (long)resident_per_cu(ctx, wl.shape) * cus * (long)std::max<uint32_t>(mult, 1);
    const unsigned grid = (unsigned)std::min<long>(total, std::max<long>(cap, 1));
    const WorkItem* d_items = b.meta(items);
    uint32_t* ctr = b.meta(std::vector<uint32_t>{0u, 0u});
    uint32_t* slot_tab = b.alloc<uint32_t>((size_t)grid * gpu * wl.shape->slot_table_words * 4);
    const heu_hip::WorklistFn fn = wl.shape->launch;
    const int nitems = (int)items.size();
    b.launches += 1;
    if (b.base) {
      std::vector<Step> lane{[=](hipStream_t st) { fn(st, d_items, nitems, total, ctr, slot_tab, grid); }};
      lane.insert(lane.end(), wl.follow.begin(), wl.follow.end());
      lanes.emplace_back(work, std::move(lane));
    }
    units += total;
  }
  if (b.base) {
    // the most work first: its launch is issued first and takes the chip, and
    // the lighter lanes fill in as its workgroups drain
    std::stable_sort(lanes.begin(), lanes.end(),
                     [](const std::pair<double, std::vector<Step>>& x, const std::pair<double, std::vector<Step>>& y) {
                       return x.first > y.first;
                     });
    PlanStage stage;
    for (auto& l : lanes) {
      b.steps.insert(b.steps.end(), l.second.begin(), l.second.end());
      stage.lanes.push_back(std::move(l.second));
    }
    b.steps.insert(b.steps.end(), b.follow.begin(), b.follow.end());
    stage.after = std::move(b.follow);
    b.follow.clear();
    if (!stage.lanes.empty() || !stage.after.empty()) b.stages.push_back(std::move(stage));
  }
  *units_out += units;
  return HEU_OK;
}

int plan(Builder& b, heu_ctx* ctx, const std::vector<heu_task>& tasks, uint64_t* units_out) {
  *units_out = 0;
  size_t begin = 0;
  for (size_t i = 0; i <= tasks.size(); ++i) {
    if (i < tasks.size() && tasks[i].op != HEU_OP_BARRIER) continue;
    int rc = plan_stage(b, ctx, tasks.data() + begin, i - begin, units_out);
    if (rc) return rc;
    begin = i + 1;
  }
  return HEU_OK;
}

}  // namespace

namespace heu_engine {

long worklist_round_units(heu_ctx* ctx, const WorklistShape* shape) {
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || cus <= 0)
    cus = 1;
  return (long)resident_per_cu(ctx, shape) * cus;
}

int plan_batch(heu_ctx* ctx, const heu_task* tasks, size_t ntasks, char* base, PlanOut* out) {
  std::vector<heu_task> ts(tasks, tasks + ntasks);
  for (size_t i = 0; i < ntasks; ++i) {
    int rc = check_task(ctx, ts[i], i);
    if (rc) return rc;
  }
  Builder b;
  b.base = base;
  uint64_t units = 0;
  int rc = plan(b, ctx, ts, &units);
  if (rc) return rc;
  out->bytes = b.used;
  out->launches = b.launches;
  out->units = units;
  out->metas = std::move(b.metas);
  out->steps = std::move(b.steps);
  out->stages = std::move(b.stages);
  return HEU_OK;
}

}  // namespace heu_engine

namespace {

void release_streams(heu_batch* b) {
  for (hipStream_t s : b->side) (void)hipStreamDestroy(s);
  for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
  b->side.clear();
  b->events.clear();
}

// Enqueues one run on `main`: stage by stage; a stage with several lanes forks
// them onto the side streams after an event on `main`, and `main` waits for
// each side lane's event before the stage's `after` steps and the next stage.
hipError_t enqueue(heu_batch* b, hipStream_t main) {
  size_t ev = 0;
  hipError_t e;
  for (const PlanStage& st : b->stages) {
    const size_t nl = st.lanes.size();
    if (nl > 1) {
      hipEvent_t fork = b->events[ev++];
      if ((e = hipEventRecord(fork, main)) != hipSuccess) return e;
      for (size_t i = 1; i < nl; ++i)
        if ((e = hipStreamWaitEvent(b->side[i - 1], fork, 0)) != hipSuccess) return e;
    }
    for (size_t i = 0; i < nl; ++i)
      for (const PlanStep& s : st.lanes[i]) s(i ? b->side[i - 1] : main);
    for (size_t i = 1; i < nl; ++i) {
      hipEvent_t join = b->events[ev++];
      if ((e = hipEventRecord(join, b->side[i - 1])) != hipSuccess) return e;
      if ((e = hipStreamWaitEvent(main, join, 0)) != hipSuccess) return e;
    }
    for (const PlanStep& s : st.after) s(main);
  }
  return hipSuccess;
}

}  // namespace

extern "C" {

int heu_batch_create(heu_ctx* ctx, const heu_task* tasks, size_t ntasks, uint32_t flags, heu_batch** out) {
  if (!ctx || !out || (!tasks && ntasks)) return fail(HEU_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  auto batch = std::make_unique<heu_batch>();
  batch->ctx = ctx;
  batch->flags = flags;
  PlanOut sizing, p;
  int rc = plan_batch(ctx, tasks, ntasks, nullptr, &sizing);
  if (rc) return rc;
  if ((rc = batch->ws.alloc(std::max<size_t>(sizing.bytes, 256)))) return rc;
  if ((rc = plan_batch(ctx, tasks, ntasks, batch->ws.as<char>(), &p))) return rc;
  if (p.bytes != sizing.bytes) return fail(HEU_EINVAL, "internal: batch plan is not deterministic");
  for (auto& m : p.metas)
    HIP_TRY(hipMemcpyAsync(batch->ws.as<char>() + m.first, m.second.data(), m.second.size(), hipMemcpyHostToDevice,
                           ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));  // the host images die with `p`
  batch->stages = std::move(p.stages);
  batch->launches = p.launches;
  batch->units = p.units;
  size_t lanes = 1, nev = 0;
  for (const PlanStage& st : batch->stages) {
    lanes = std::max(lanes, st.lanes.size());
    if (st.lanes.size() > 1) nev += st.lanes.size();
  }
  rc = HEU_OK;
  for (size_t i = 1; i < lanes && !rc; ++i) {
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) rc = fail(HEU_EDEVICE, "hipStreamCreate failed");
    else batch->side.push_back(s);
  }
  for (size_t i = 0; i < nev && !rc; ++i) {
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) rc = fail(HEU_EDEVICE, "hipEventCreate failed");
    else batch->events.push_back(e);
  }
  if (rc) {
    release_streams(batch.get());
    return rc;
  }
  ctx->refs += 1;  // under ctx->mu: the batch holds the context
  *out = batch.release();
  return HEU_OK;
}

int heu_batch_run(heu_batch* batch) {
  if (!batch) return fail(HEU_EINVAL, "null batch");
  heu_ctx* ctx = batch->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  (void)hipGetLastError();  // an error left by other code is not this run's
  if (!(batch->flags & HEU_BATCH_GRAPH)) {
    HIP_TRY(enqueue(batch, ctx->stream));
    HIP_TRY(hipGetLastError());
    return HEU_OK;
  }
  if (!batch->exec) {
    // capture on a private stream (the context's stream may be the null stream)
    if (!batch->cap_stream) HIP_TRY(hipStreamCreateWithFlags(&batch->cap_stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamBeginCapture(batch->cap_stream, hipStreamCaptureModeThreadLocal));
    const hipError_t qe = enqueue(batch, batch->cap_stream);
    hipGraph_t g = nullptr;
    const hipError_t le = hipGetLastError();
    const hipError_t ce = hipStreamEndCapture(batch->cap_stream, &g);
    if (g && (qe != hipSuccess || ce != hipSuccess || le != hipSuccess)) (void)hipGraphDestroy(g);
    HIP_TRY(qe);
    HIP_TRY(ce);
    HIP_TRY(le);
    const hipError_t ie = hipGraphInstantiate(&batch->exec, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
      (void)hipGraphDestroy(g);
      batch->exec = nullptr;
      HIP_TRY(ie);
    }
    batch->graph = g;
  }
  HIP_TRY(hipGraphLaunch(batch->exec, ctx->stream));
  return HEU_OK;
}

int heu_batch_destroy(heu_batch* batch) {
  if (!batch) return HEU_OK;
  heu_ctx* ctx = batch->ctx;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (batch->exec) (void)hipGraphExecDestroy(batch->exec);
    if (batch->graph) (void)hipGraphDestroy(batch->graph);
    if (batch->cap_stream) (void)hipStreamDestroy(batch->cap_stream);
    release_streams(batch);
    delete batch;
  }
  ctx_release(ctx);
  return HEU_OK;
}

int heu_batch_stats(const heu_batch* batch, uint64_t* launches, uint64_t* units, size_t* device_bytes) {
  if (!batch) return fail(HEU_EINVAL, "null batch");
  if (launches) *launches = batch->launches;
  if (units) *units = batch->units;
  if (device_bytes) *device_bytes = batch->ws.bytes;
  return HEU_OK;
}

int heu_ctx_set_grid_multiplier(heu_ctx* ctx, uint32_t key_bits, int decrypt_shape, uint32_t multiplier) {
  if (!ctx) return fail(HEU_EINVAL, "null ctx");
  if (key_bits != 2048 && key_bits != 3072) return fail(HEU_EUNSUPPORTED, "no kernels for %u-bit keys", key_bits);
  if (multiplier < 1 || multiplier > 64) return fail(HEU_EINVAL, "multiplier must be in [1, 64]");
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->grid_mult[key_bits == 3072][decrypt_shape ? 1 : 0] = multiplier;
  return HEU_OK;
}

}  // extern "C"
