// engine_internal.hpp -- engine state shared by the C ABI translation units
// (engine.cpp: contexts, keys and one-op entry points; batch.cpp: multi-task
// batches and their graphs).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/heu_paillier.h"
#include "kernels.hpp"

namespace heu_engine {

// Per-key budget of the fixed-base h_s table: HEU_TABLE_BYTES, default 4 GiB.
// At 4 GiB a 2048-bit key takes 16-bit windows for r of 1,024 bits (2.4 GB)
// as well as for r of 256 bits (0.6 GB); at 1 GiB the first fell to 14 bits,
// with 15% slower encryption.
inline size_t table_budget_bytes() {
  size_t budget = (size_t)4 << 30;
  if (const char* b = std::getenv("HEU_TABLE_BYTES")) budget = (size_t)std::strtoull(b, nullptr, 10);
  return budget;
}

// Records the message of a failing call (heu_last_error) and returns `code`.
int fail(int code, const char* fmt, ...);

#define HIP_TRY(expr)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      return ::heu_engine::fail(e_ == hipErrorOutOfMemory ? HEU_ENOMEM : HEU_EDEVICE, "%s: %s", #expr, \
                                hipGetErrorString(e_));                                            \
  } while (0)

// ---------------------------------------------------------------------------
// Device buffers
// ---------------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t n) {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (n == 0) return HEU_OK;
    HIP_TRY(hipMalloc(&p, n));
    bytes = n;
    return HEU_OK;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Batch inversion (Montgomery's trick): level 0 multiplies the inputs in
// chunks of up to kInvChunk0, which fills the chip for large batches.  A batch
// that would not fill the chip about twice over gets shorter level-0 chunks
// (inv_chunk0, at least kInvChunk0Min), so a short batch is a short chain of
// products rather than a long chain on a few groups.  Upper levels see few
// values, where a level's time is its chunk's chain of sequential products, so
// they multiply pairs (kInvChunkUpper), as the reductions' upper levels do.
constexpr long kInvChunk0 = 32, kInvChunk0Min = 4, kInvChunkUpper = 2;
long inv_chunk0(heu_ctx* ctx, const heu_hip::Ops* o, long count);
inline int inv_chunk(size_t level, long chunk0) { return (int)(level == 0 ? chunk0 : kInvChunkUpper); }

// Element counts per level of a batch inversion (up to one value).
std::vector<long> inv_levels(long count, long chunk0);

// Chunking of segments for a reduction level.  Level 0 reads the inputs in
// chunks of up to kReduceChunk0 (kReduceChunkDot0 for dot terms, each a power),
// which fills the chip for large inputs.  Upper levels see few partials, so a
// level takes as long as its chunk's sequential products while most of the chip
// idles, and a level costs about one product's latency per product in its
// chain plus a launch: pairs (kReduceChunkUpper = 2) take 128 partials per
// segment to one in 7 levels of one product each, 214 us for a 4096 x 256
// axis-0 sum against 257 us in 4 levels of chunks of 4 (3 + 3 + 3 + 1).
constexpr long kReduceChunk0 = 32, kReduceChunkDot0 = 8, kReduceChunkUpper = 2;
static_assert(kReduceChunk0 < heu_hip::kRpowEntries, "a Z-space level-0 chunk needs its R power");
// Level 0's chunk for a reduction of `total` terms: the full chunk when the
// terms fill the chip about twice over, shorter for fewer, so a short vector's
// level 0 is a short chain of products rather than a long chain on a few groups.
// The one-op path, the batch planner and the server's work estimate share it.
long reduce_chunk0(heu_ctx* ctx, const heu_hip::Ops* o, bool dot, long total);
std::vector<heu_hip::Chunk> make_chunks(const std::vector<long>& seg_start, const std::vector<long>& seg_len, long chunk,
                                        std::vector<long>* seg_first_chunk, std::vector<long>* seg_nchunks);

// A planned batch (batch.cpp): with base == nullptr only `bytes` is computed
// (the sizing pass); otherwise every device pointer lies in [base, base +
// bytes), `metas` are host images to copy to base + offset before the first
// run, and `steps` enqueue one run on a stream.  `stages` hold the same steps
// with their dependencies: stages run in order; within a stage each lane (one
// work-list launch and the steps that read its results) is independent of the
// other lanes, so lanes may run on streams of their own, and `after` runs once
// every lane is done.
typedef std::function<void(hipStream_t)> PlanStep;
struct PlanStage {
  std::vector<std::vector<PlanStep>> lanes;
  std::vector<PlanStep> after;
};
struct PlanOut {
  size_t bytes = 0;
  uint64_t launches = 0, units = 0;
  std::vector<std::pair<size_t, std::vector<char>>> metas;
  std::vector<PlanStep> steps;
  std::vector<PlanStage> stages;
};
// Context lifetime: objects created on a context hold a reference; the
// context is freed by heu_ctx_destroy or by the last release after it.
void ctx_retain(heu_ctx* ctx);
void ctx_release(heu_ctx* ctx);

// Checks the tasks and plans them (caller holds ctx->mu).
int plan_batch(heu_ctx* ctx, const heu_task* tasks, size_t ntasks, char* base, PlanOut* out);
// heu_sum's segment product with every scratch part carved from the caller's
// fixed device buffer [buf, buf + bytes) instead of the context's pool; fails
// with HEU_ERANGE when the call needs more (engine.cpp).
int sum_fixed_scratch(heu_key* key, const uint32_t* ct, const int64_t* offsets, size_t segments, uint32_t* out,
                      char* buf, size_t bytes);
// Looks up this device's address of heu_sum_noalloc's fixed scratch array,
// once per device, with the device current (noalloc.hip); returns the HIP
// status, which heu_sum_noalloc reports if the lookup failed.
hipError_t noalloc_setup(int device);

// Work-list units of `shape` that run at once on the whole device: resident
// workgroups per CU x CUs.  A grid multiplier above 1 launches more workgroups
// than that, but they queue; it does not change what runs at once (caller
// holds ctx->mu).
long worklist_round_units(heu_ctx* ctx, const heu_hip::WorklistShape* shape);

}  // namespace heu_engine

using heu_hip::CombArgs;
using heu_hip::DecArgs;
using heu_hip::ModArgs;
using heu_hip::Ops;
using heu_hip::PubArgs;

struct heu_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  uint32_t* d_err = nullptr;
  std::mutex mu;
  // scratch pool: one growable buffer carved per call
  heu_engine::DevBuf scratch;
  size_t scratch_used = 0;
  // small host->device staging for chunk lists etc.
  heu_engine::DevBuf meta;
  // optional per-kernel timing: HIP events recorded on the work stream
  struct ProfRec {
    const char* name;
    hipEvent_t start, stop;
  };
  // keys, batches and servers created on this context: it is freed when the
  // last of them goes after heu_ctx_destroy (whatever order callers release in)
  int refs = 0;
  bool closing = false;
  // persistent work-list launches: grid multiplier per [key size is 3072][decryption shape]
  // and the occupancy of each work-list kernel (resident workgroups per CU)
  uint32_t grid_mult[2][2] = {{1, 1}, {1, 1}};
  std::map<const heu_hip::WorklistShape*, int> wl_resident;
  bool prof = false;
  std::vector<ProfRec> prof_recs;
  std::vector<hipEvent_t> prof_pool;
  hipEvent_t prof_event() {
    if (!prof_pool.empty()) {
      hipEvent_t e = prof_pool.back();
      prof_pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
  void prof_clear() {
    for (auto& r : prof_recs) {
      prof_pool.push_back(r.start);
      prof_pool.push_back(r.stop);
    }
    prof_recs.clear();
  }
};

// Launches `call` and, when profiling is on, brackets it with events.
#define LAUNCH(ctx, name, call)                             \
  do {                                                      \
    hipEvent_t a_ = nullptr;                                \
    if ((ctx)->prof) {                                      \
      a_ = (ctx)->prof_event();                             \
      (void)hipEventRecord(a_, (ctx)->stream);              \
    }                                                       \
    call;                                                   \
    if ((ctx)->prof) {                                      \
      hipEvent_t b_ = (ctx)->prof_event();                  \
      (void)hipEventRecord(b_, (ctx)->stream);              \
      (ctx)->prof_recs.push_back({(name), a_, b_});         \
    }                                                       \
  } while (0)

struct heu_key {
  heu_ctx* ctx = nullptr;
  const Ops* ops = nullptr;
  int key_bits = 0, n_words = 0, ct_words = 0, prime_words = 0, has_sk = 0;
  int nl_n = 0;
  uint32_t r_bits = 0;  // randomness bits the table covers
  // public side
  heu_engine::DevBuf n2_m, n2_one, n2_r2, n2_rpow, n29, table, n2_words;
  uint32_t n2_mprime = 0;
  int windows = 0, unit = 0;
  // secret side: p^2 / q^2 constants for every decryption shape, one active
  struct Half {
    heu_engine::DevBuf m, one, r2, r3;
    uint32_t mprime = 0;
    ModArgs args() const {  // no R-power table: only sums mod n^2 read one
      return ModArgs{m.as<uint32_t>(), one.as<uint32_t>(), r2.as<uint32_t>(), r3.as<uint32_t>(), mprime, nullptr};
    }
  };
  Half p2[heu_hip::kDecShapes], q2[heu_hip::kDecShapes];
  int dec_shape = 0;
  heu_engine::DevBuf sched_p, sched_q;
  int nsched_p = 0, nsched_q = 0;
  heu_engine::DevBuf c_p, c_q, c_pinv, c_qinv, c_hpR, c_hqR, c_pinvqR, c_n, c_nhalf;
  uint32_t c_pprime = 0, c_qprime = 0;
  // host copies for argument checks and conversions
  std::vector<uint32_t> n_host, n2_host, nhalf_host, hs_host;
  size_t device_bytes = 0;

  ModArgs mod_n2() const {
    return ModArgs{n2_m.as<uint32_t>(), n2_one.as<uint32_t>(), n2_r2.as<uint32_t>(), n2_r2.as<uint32_t>(), n2_mprime,
                   n2_rpow.as<uint32_t>()};
  }
  PubArgs pub() const { return PubArgs{mod_n2(), n29.as<uint32_t>(), table.as<uint32_t>(), windows, unit}; }
  const heu_hip::DecShape& dshape() const { return ops->dec[dec_shape]; }
  int nl_p() const { return dshape().L * dshape().K; }
  DecArgs dec() const {
    return DecArgs{p2[dec_shape].args(), q2[dec_shape].args(), sched_p.as<uint32_t>(), sched_q.as<uint32_t>(),
                   nsched_p, nsched_q};
  }
  CombArgs comb() const {
    return CombArgs{c_p.as<uint32_t>(),   c_q.as<uint32_t>(),    c_pinv.as<uint32_t>(), c_qinv.as<uint32_t>(),
                    c_hpR.as<uint32_t>(), c_hqR.as<uint32_t>(),  c_pinvqR.as<uint32_t>(), c_n.as<uint32_t>(),
                    c_nhalf.as<uint32_t>(), c_pprime, c_qprime};
  }
};
