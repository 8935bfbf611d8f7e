// paillier_hip.cc -- the vectorized-SPI classes of paillier_hip.h over the
// engine's C ABI (include/heu_paillier.h).
//
// Every batch call packs its operands into flat limb arrays, stages them in a
// device arena on the engine context's stream, runs one engine entry point and
// copies the results back.  Calls from several threads share the process-wide
// runtime (one context, one stream, one device lock).  Elementwise calls go
// through a combining queue: a thread that finds no other thread running the
// queue takes every queued call, and runs each group of calls with the same
// operation, key and parameters as one staged engine call.
//
// Host side of a call: operands are packed into, and results read from,
// pooled page-locked buffers (no fresh allocation or page faults per call,
// DMA-speed copies), and the per-element conversions (validation, limb
// copies, BigInt construction, randomness) are split over the usable cores.
#include "paillier_hip.h"

#include <hip/hip_runtime_api.h>
#include <openssl/bn.h>
#include <openssl/crypto.h>
#include <openssl/rand.h>
#include <sched.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <condition_variable>
#include <exception>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <stdexcept>

#include "heu_paillier.h"

namespace heu_amd {

// ------------------------------------------------------------------ BigInt
namespace {

BN_CTX* tls_ctx() {
  thread_local std::unique_ptr<BN_CTX, void (*)(BN_CTX*)> c(BN_CTX_new(), BN_CTX_free);
  return c.get();
}

[[noreturn]] void throw_msg(const std::string& m) { throw std::runtime_error(m); }

void bn_check(int ok, const char* what) {
  if (!ok) throw_msg(std::string("BIGNUM ") + what + " failed");
}

}  // namespace

BigInt::BigInt() : b_(BN_new()) {
  if (!b_) throw std::bad_alloc();
}

BigInt::BigInt(int64_t v) : BigInt() {
  uint64_t mag = v < 0 ? 0ull - (uint64_t)v : (uint64_t)v;
  bn_check(BN_set_word(b_, mag), "set_word");
  BN_set_negative(b_, v < 0);
}

BigInt::BigInt(const BigInt& o) : b_(BN_dup(o.b_)) {
  if (!b_) throw std::bad_alloc();
}

BigInt::BigInt(BigInt&& o) noexcept : b_(o.b_) { o.b_ = nullptr; }

BigInt& BigInt::operator=(const BigInt& o) {
  if (this != &o) {
    if (!b_) b_ = BN_new();
    bn_check(BN_copy(b_, o.b_) != nullptr, "copy");
  }
  return *this;
}

BigInt& BigInt::operator=(BigInt&& o) noexcept {
  std::swap(b_, o.b_);
  return *this;
}

BigInt::~BigInt() { BN_free(b_); }

BigInt BigInt::FromString(const std::string& s, int base) {
  BigInt r;
  BIGNUM* p = r.b_;
  int n = base == 16 ? BN_hex2bn(&p, s.c_str()) : BN_dec2bn(&p, s.c_str());
  if (n == 0 || (size_t)n + (s[0] == '-') < s.size()) throw_msg("BigInt: cannot parse '" + s + "'");
  return r;
}

BigInt BigInt::FromWords(const uint32_t* w, size_t words, bool twos_complement) {
  BigInt r;
  if (words == 0) return r;
  bn_check(BN_lebin2bn(reinterpret_cast<const unsigned char*>(w), (int)(words * 4), r.b_) != nullptr, "lebin2bn");
  if (twos_complement && (w[words - 1] >> 31)) {
    BigInt mod;
    bn_check(BN_lshift(mod.b_, BN_value_one(), (int)(32 * words)), "lshift");
    bn_check(BN_sub(r.b_, r.b_, mod.b_), "sub");
  }
  return r;
}

void BigInt::ToWords(uint32_t* w, size_t words) const {
  // fits iff -2^(cap-1) <= v < 2^(cap-1), cap = 32 words
  const size_t cap = 32 * words, bits = BitCount();
  bool fits = bits < cap;
  if (!fits && IsNegative() && bits == cap) {  // only v = -2^(cap-1) has a cap-bit magnitude and fits
    BigInt t(*this);
    BN_clear_bit(t.b_, (int)cap - 1);
    fits = t.IsZero();
  }
  if (!fits) throw_msg("BigInt: value does not fit " + std::to_string(words) + " words");
  if (!IsNegative()) {
    bn_check(BN_bn2lebinpad(b_, reinterpret_cast<unsigned char*>(w), (int)(words * 4)) >= 0, "bn2lebinpad");
    return;
  }
  BigInt t;
  bn_check(BN_lshift(t.b_, BN_value_one(), (int)(32 * words)), "lshift");
  bn_check(BN_add(t.b_, t.b_, b_), "add");  // 2^(32 words) - |v|
  bn_check(BN_bn2lebinpad(t.b_, reinterpret_cast<unsigned char*>(w), (int)(words * 4)) >= 0, "bn2lebinpad");
}

std::string BigInt::ToString() const {
  char* s = BN_bn2dec(b_);
  std::string out(s);
  OPENSSL_free(s);
  return out;
}

std::string BigInt::ToHexString() const {
  char* s = BN_bn2hex(b_);
  std::string out(s);
  OPENSSL_free(s);
  return out;
}

size_t BigInt::BitCount() const { return (size_t)BN_num_bits(b_); }
bool BigInt::IsNegative() const { return BN_is_negative(b_) != 0; }
bool BigInt::IsZero() const { return BN_is_zero(b_) != 0; }
int BigInt::Compare(const BigInt& o) const { return BN_cmp(b_, o.b_); }
int BigInt::CompareAbs(const BigInt& o) const { return BN_ucmp(b_, o.b_); }

bool BigInt::FitsInt64() const {
  size_t bits = BitCount();
  if (bits <= 63) return true;
  // INT64_MIN = -2^63
  return bits == 64 && IsNegative() && BN_get_word(b_) == (1ull << 63);
}

int64_t BigInt::ToInt64() const {
  if (!FitsInt64()) throw_msg("BigInt: value does not fit int64");
  uint64_t mag = BN_get_word(b_);
  return IsNegative() ? (int64_t)(0ull - mag) : (int64_t)mag;
}

BigInt BigInt::operator+(const BigInt& o) const {
  BigInt r;
  bn_check(BN_add(r.b_, b_, o.b_), "add");
  return r;
}

BigInt BigInt::operator-(const BigInt& o) const {
  BigInt r;
  bn_check(BN_sub(r.b_, b_, o.b_), "sub");
  return r;
}

BigInt BigInt::operator*(const BigInt& o) const {
  BigInt r;
  bn_check(BN_mul(r.b_, b_, o.b_, tls_ctx()), "mul");
  return r;
}

BigInt BigInt::operator-() const {
  BigInt r(*this);
  if (!r.IsZero()) BN_set_negative(r.b_, !IsNegative());
  return r;
}

BigInt BigInt::operator>>(int bits) const {
  BigInt r;
  bn_check(BN_rshift(r.b_, b_, bits), "rshift");
  return r;
}

namespace paillier_hip {

// ------------------------------------------------------------------ runtime
namespace {

void ck(int rc) {
  if (rc != HEU_OK) {
    const char* m = heu_last_error();
    throw_msg(std::string("heu engine error ") + std::to_string(rc) + ": " + (m ? m : ""));
  }
}

void hip_ck(hipError_t e, const char* what) {
  if (e != hipSuccess) throw_msg(std::string(what) + ": " + hipGetErrorString(e));
}

// An elementwise call waiting in the combining queue: `count` elements, each
// input array holding in_elem[i] bytes per element, the output out_elem.
struct Job {
  int op = 0;
  heu_key* key = nullptr;
  uint32_t param = 0;
  size_t count = 0;
  int nin = 0;
  const void* in[2] = {nullptr, nullptr};
  size_t in_elem[2] = {0, 0};
  void* out = nullptr;
  size_t out_elem = 0;
  bool ran = false;   // set by the thread that ran it
  bool done = false;  // set under the queue lock once its results are final
  std::exception_ptr err;
};

// One engine context per process (device from HEU_DEVICE, default 0) and a
// growable device arena for staging batch operands.  Deliberately never
// destroyed: HIP's own teardown at exit must not race a late hipFree.
struct Runtime {
  heu_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  std::mutex mu;  // the device: arena, stream, engine calls
  char* arena = nullptr;
  size_t cap = 0;
  // combining queue
  std::mutex qmu;
  std::condition_variable qcv;
  std::vector<Job*> pending;
  bool running = false, hold = false;
  SpiStats stats;

  static Runtime& get() {
    static Runtime* rt = [] {
      auto* r = new Runtime;
      const char* d = std::getenv("HEU_DEVICE");
      ck(heu_ctx_create(d ? std::atoi(d) : 0, &r->ctx));
      r->stream = static_cast<hipStream_t>(heu_ctx_stream(r->ctx));
      return r;
    }();
    return *rt;
  }

  char* reserve(size_t bytes) {
    if (bytes > cap) {
      // grow geometrically, so a run of slightly larger calls reallocates rarely
      const size_t want = std::max(bytes, cap * 2);
      hip_ck(hipStreamSynchronize(stream), "hipStreamSynchronize");
      if (arena) hip_ck(hipFree(arena), "hipFree");
      arena = nullptr;
      cap = 0;
      hip_ck(hipMalloc(reinterpret_cast<void**>(&arena), want), "hipMalloc");
      cap = want;
    }
    return arena;
  }
};

// Stages one call's operands: inputs are copied up, outputs copied back after
// the engine call, then the stream is synchronized and device flags checked.
class DeviceCall {
 public:
  DeviceCall() : rt_(Runtime::get()), lock_(rt_.mu) {}

  size_t in(const void* host, size_t bytes) {
    size_t h = slot(bytes);
    ups_.push_back({host, bytes, h});
    return h;
  }
  size_t out(size_t bytes) { return slot(bytes); }
  // a piece of a slot: upload `bytes` from host at byte offset `off`
  void put(size_t h, size_t off, const void* host, size_t bytes) { pieces_.push_back({host, bytes, h, off}); }

  template <class T>
  T* ptr(size_t h) {
    return reinterpret_cast<T*>(base_ + offs_[h]);
  }

  // An exception between upload() and finish() leaves copies in flight that
  // read or write the callers' host buffers: wait for them before those go.
  ~DeviceCall() {
    if (started_ && !finished_) (void)hipStreamSynchronize(rt_.stream);
  }

  void upload() {
    started_ = true;
    base_ = rt_.reserve(total_ ? total_ : 256);
    for (auto& u : ups_)
      if (u.bytes)
        hip_ck(hipMemcpyAsync(base_ + offs_[u.h], u.host, u.bytes, hipMemcpyHostToDevice, rt_.stream),
               "hipMemcpyAsync H2D");
    for (auto& u : pieces_)
      if (u.bytes)
        hip_ck(hipMemcpyAsync(base_ + offs_[u.h] + u.off, u.host, u.bytes, hipMemcpyHostToDevice, rt_.stream),
               "hipMemcpyAsync H2D");
  }

  void download(size_t h, void* host, size_t bytes, size_t off = 0) {
    if (bytes)
      hip_ck(hipMemcpyAsync(host, base_ + offs_[h] + off, bytes, hipMemcpyDeviceToHost, rt_.stream),
             "hipMemcpyAsync D2H");
  }

  void finish() {
    finished_ = true;
    ck(heu_ctx_sync(rt_.ctx));
  }

 private:
  struct Up {
    const void* host;
    size_t bytes;
    size_t h;
  };
  struct Piece {
    const void* host;
    size_t bytes;
    size_t h, off;
  };
  size_t slot(size_t bytes) {
    offs_.push_back(total_);
    total_ += (bytes + 255) & ~size_t(255);
    return offs_.size() - 1;
  }

  Runtime& rt_;
  std::lock_guard<std::mutex> lock_;
  std::vector<size_t> offs_;
  std::vector<Up> ups_;
  std::vector<Piece> pieces_;
  size_t total_ = 0;
  char* base_ = nullptr;
  bool started_ = false, finished_ = false;
};

// ---- host staging and threads

// Threads for one call's per-element host work: the usable cores (the
// affinity mask), at most 16.
size_t host_threads() {
  static const size_t n = [] {
    cpu_set_t set;
    size_t c = 1;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) c = (size_t)CPU_COUNT(&set);
    return std::max<size_t>(1, std::min<size_t>(c, 16));
  }();
  return n;
}

// Elements per thread below which a loop stays on the calling thread.
constexpr size_t kGrain = 2048;

// Helper threads running now, across all calls: concurrent calls (HEU's
// worker threads each converting a chunk) share host_threads() - 1 helpers
// instead of each starting its own.
std::atomic<size_t> g_helpers{0};

// f(begin, end) over [0, n), split into contiguous chunks on up to
// host_threads() threads, the caller's among them.  An exception ends its
// chunk; once every chunk ended, the exception of the lowest chunk is
// rethrown, which is the first failing element in order, as a serial loop
// would report it.
template <class F>
void par_for(size_t n, size_t grain, const F& f) {
  const size_t want = std::min(host_threads(), n / std::max<size_t>(grain, 1));
  size_t got = 0;
  if (want > 1) {
    const size_t budget = host_threads() - 1;
    size_t cur = g_helpers.load();
    do {
      got = std::min(want - 1, cur < budget ? budget - cur : 0);
    } while (!g_helpers.compare_exchange_weak(cur, cur + got));
  }
  const size_t t = got + 1;
  if (t <= 1) return f(size_t(0), n);
  struct Release {
    size_t k;
    ~Release() { g_helpers -= k; }
  } release{got};
  std::vector<std::exception_ptr> err(t);
  auto chunk = [&](size_t i) {
    try {
      f(n * i / t, n * (i + 1) / t);
    } catch (...) {
      err[i] = std::current_exception();
    }
  };
  std::vector<std::thread> th;
  th.reserve(t - 1);
  size_t next = 1;
  try {
    for (; next < t; ++next) th.emplace_back(chunk, next);
  } catch (...) {
    // no more threads: the chunks left run here
  }
  chunk(0);
  for (size_t i = next; i < t; ++i) chunk(i);
  for (auto& x : th) x.join();
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
}

// Page-locked host buffers, pooled for reuse.  A call takes one per operand
// or result array and gives it back when it returns.  Capacities are powers
// of two; a request reuses an idle buffer of at most 4x its size.  At most
// kKeep bytes stay idle, the rest is freed.  Never destroyed, like Runtime.
struct HostPool {
  static constexpr size_t kKeep = size_t(2) << 30;
  std::mutex mu;
  std::multimap<size_t, void*> idle;
  size_t idle_bytes = 0;

  static HostPool& get() {
    static HostPool* p = new HostPool;
    return *p;
  }

  void* take(size_t bytes, size_t* cap) {
    {
      std::lock_guard<std::mutex> lk(mu);
      auto it = idle.lower_bound(bytes);
      if (it != idle.end() && it->first / 4 <= bytes) {
        void* p = it->second;
        *cap = it->first;
        idle_bytes -= it->first;
        idle.erase(it);
        return p;
      }
    }
    size_t c = size_t(1) << 16;
    while (c < bytes) c <<= 1;
    Runtime::get();  // the device the context uses is current
    void* p = nullptr;
    hip_ck(hipHostMalloc(&p, c, hipHostMallocDefault), "hipHostMalloc");
    *cap = c;
    return p;
  }

  void give(void* p, size_t cap) {
    {
      std::lock_guard<std::mutex> lk(mu);
      if (idle_bytes + cap <= kKeep) {
        idle.emplace(cap, p);
        idle_bytes += cap;
        return;
      }
    }
    (void)hipHostFree(p);
  }
};

// n values of a trivially copyable T in a pooled page-locked buffer
// (contents undefined until written).
template <class T>
class Staged {
 public:
  Staged() = default;
  explicit Staged(size_t n) : n_(n) {
    if (n) p_ = static_cast<T*>(HostPool::get().take(n * sizeof(T), &cap_));
  }
  Staged(Staged&& o) noexcept : p_(o.p_), n_(o.n_), cap_(o.cap_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  Staged& operator=(Staged&& o) noexcept {
    std::swap(p_, o.p_);
    std::swap(n_, o.n_);
    std::swap(cap_, o.cap_);
    return *this;
  }
  Staged(const Staged&) = delete;
  ~Staged() {
    if (p_) HostPool::get().give(p_, cap_);
  }
  T* data() { return p_; }
  const T* data() const { return p_; }
  size_t size() const { return n_; }
  T& operator[](size_t i) { return p_[i]; }

 private:
  T* p_ = nullptr;
  size_t n_ = 0, cap_ = 0;
};

// Operations the combining queue runs (elementwise, fixed bytes per element).
enum JobOp { kJobEncI64, kJobEncZero, kJobRandomize, kJobDecrypt, kJobAdd, kJobSub, kJobAddI64, kJobMulI64, kJobNegate };

int engine_call(const Job& j, void* const* in, size_t n, void* out) {
  auto u32 = [](void* p) { return static_cast<uint32_t*>(p); };
  auto i64 = [](void* p) { return static_cast<int64_t*>(p); };
  switch (j.op) {
    case kJobEncI64: return heu_encrypt_i64(j.key, i64(in[0]), u32(in[1]), j.param, n, u32(out));
    case kJobEncZero: return heu_randomize(j.key, nullptr, u32(in[0]), j.param, n, u32(out));
    case kJobRandomize: return heu_randomize(j.key, u32(in[0]), u32(in[1]), j.param, n, u32(out));
    case kJobDecrypt: return heu_decrypt(j.key, u32(in[0]), n, u32(out), j.param, nullptr);
    case kJobAdd: return heu_add(j.key, u32(in[0]), u32(in[1]), n, u32(out));
    case kJobSub: return heu_sub(j.key, u32(in[0]), u32(in[1]), n, u32(out));
    case kJobAddI64: return heu_add_i64(j.key, u32(in[0]), i64(in[1]), n, u32(out));
    case kJobMulI64: return heu_mul_i64(j.key, u32(in[0]), i64(in[1]), n, u32(out));
    case kJobNegate: return heu_negate(j.key, u32(in[0]), n, u32(out));
    default: return HEU_EINVAL;
  }
}

// Runs a group of jobs with the same op, key and parameter as one staged
// engine call.  If it fails, the jobs run one by one, so an error reaches only
// the call whose inputs caused it.
void run_group(const std::vector<Job*>& g) {
  const Job& j0 = *g[0];
  size_t total = 0;
  for (const Job* j : g) total += j->count;
  try {
    DeviceCall call;
    size_t hin[2] = {0, 0};
    for (int i = 0; i < j0.nin; ++i) {
      hin[i] = call.out(total * j0.in_elem[i]);
      size_t off = 0;
      for (const Job* j : g) {
        call.put(hin[i], off, j->in[i], j->count * j0.in_elem[i]);
        off += j->count * j0.in_elem[i];
      }
    }
    const size_t ho = call.out(total * j0.out_elem);
    call.upload();
    void* dev_in[2] = {j0.nin > 0 ? call.ptr<char>(hin[0]) : nullptr, j0.nin > 1 ? call.ptr<char>(hin[1]) : nullptr};
    ck(engine_call(j0, dev_in, total, call.ptr<char>(ho)));
    size_t off = 0;
    for (const Job* j : g) {
      call.download(ho, j->out, j->count * j0.out_elem, off);
      off += j->count * j0.out_elem;
    }
    call.finish();
    for (Job* j : g) j->ran = true;
    Runtime& rt = Runtime::get();
    std::lock_guard<std::mutex> lk(rt.qmu);
    rt.stats.device_calls += 1;
    rt.stats.max_calls_per_device_call = std::max<uint64_t>(rt.stats.max_calls_per_device_call, g.size());
  } catch (...) {
    if (g.size() == 1) {
      g[0]->err = std::current_exception();
      g[0]->ran = true;
      return;
    }
    for (Job* j : g) run_group(std::vector<Job*>{j});
  }
}

// Queues `j` and returns once it ran; rethrows its error.
void submit(Job& j) {
  Runtime& rt = Runtime::get();
  std::unique_lock<std::mutex> lk(rt.qmu);
  rt.pending.push_back(&j);
  rt.stats.calls += 1;
  while (!j.done) {
    if (!rt.running && !rt.hold) {
      rt.running = true;
      std::vector<Job*> take;
      take.swap(rt.pending);
      lk.unlock();
      try {
        std::map<std::tuple<int, heu_key*, uint32_t>, std::vector<Job*>> groups;
        std::vector<std::tuple<int, heu_key*, uint32_t>> order;
        for (Job* t : take) {
          auto k = std::make_tuple(t->op, t->key, t->param);
          auto& v = groups[k];
          if (v.empty()) order.push_back(k);
          v.push_back(t);
        }
        for (auto& k : order) run_group(groups[k]);
      } catch (...) {
        // run_group keeps errors per call; anything else fails the calls not yet run
        for (Job* t : take)
          if (!t->ran) t->err = std::current_exception();
      }
      lk.lock();
      for (Job* t : take) t->done = true;
      rt.running = false;
      rt.qcv.notify_all();
    } else {
      rt.qcv.wait(lk);
    }
  }
  if (j.err) std::rethrow_exception(j.err);
}

Job make_job(int op, heu_key* key, uint32_t param, size_t count, void* out, size_t out_elem) {
  Job j;
  j.op = op;
  j.key = key;
  j.param = param;
  j.count = count;
  j.out = out;
  j.out_elem = out_elem;
  return j;
}

void add_in(Job& j, const void* p, size_t elem) {
  j.in[j.nin] = p;
  j.in_elem[j.nin] = elem;
  j.nin += 1;
}

}  // namespace

SpiStats GetSpiStats() {
  Runtime& rt = Runtime::get();
  std::lock_guard<std::mutex> lk(rt.qmu);
  return rt.stats;
}

void HoldSpiQueue(bool hold) {
  Runtime& rt = Runtime::get();
  std::lock_guard<std::mutex> lk(rt.qmu);
  rt.hold = hold;
  rt.qcv.notify_all();
}

size_t SpiQueueLength() {
  Runtime& rt = Runtime::get();
  std::lock_guard<std::mutex> lk(rt.qmu);
  return rt.pending.size();
}

struct DeviceKey {
  heu_key* key = nullptr;
  uint32_t key_bits = 0, n_words = 0, ct_words = 0, r_words = 0;
  bool has_sk = false;
  std::vector<uint32_t> n2;  // n^2 limbs, for the ciphertext range check

  ~DeviceKey() {
    if (key) {
      std::lock_guard<std::mutex> lk(Runtime::get().mu);
      heu_key_destroy(key);
    }
  }
};

namespace {

std::vector<uint32_t> words_of(const BigInt& v, size_t words) {
  std::vector<uint32_t> w(words + 1);  // one spare word keeps a positive top bit clear
  v.ToWords(w.data(), words + 1);
  w.resize(words);
  return w;
}

// The reference draws r with key_size/2 bits (encryptor.cc:26-27).  OpenSSL's
// RAND_bytes is thread-safe (a DRBG per thread), so chunks draw in parallel.
Staged<uint32_t> random_r(const DeviceKey& dk, size_t count) {
  const size_t rw = dk.r_words;
  Staged<uint32_t> r(count * rw);
  const unsigned top = (dk.key_bits / 2) % 32;
  par_for(count, kGrain, [&](size_t b, size_t e) {
    if (e > b && RAND_bytes(reinterpret_cast<unsigned char*>(r.data() + b * rw), (int)((e - b) * rw * 4)) != 1)
      throw_msg("RAND_bytes failed");
    if (top)
      for (size_t i = b; i < e; ++i) r[i * rw + rw - 1] &= (1u << top) - 1u;
  });
  return r;
}

std::shared_ptr<DeviceKey> make_device_key(const BigInt& n, const BigInt& hs, const BigInt* p, const BigInt* q,
                                           size_t key_bits) {
  auto dk = std::make_shared<DeviceKey>();
  dk->key_bits = (uint32_t)key_bits;
  dk->n_words = (uint32_t)(key_bits / 32);
  dk->ct_words = 2 * dk->n_words;
  dk->r_words = (uint32_t)((key_bits / 2 + 31) / 32);
  auto nw = words_of(n, dk->n_words);
  auto hw = words_of(hs, dk->ct_words);
  dk->n2 = words_of(n * n, dk->ct_words);
  std::vector<uint32_t> pw, qw;
  if (p && q) {
    pw = words_of(*p, dk->n_words / 2);
    qw = words_of(*q, dk->n_words / 2);
  }
  Runtime& rt = Runtime::get();
  {
    std::lock_guard<std::mutex> lk(rt.mu);
    ck(heu_key_create(rt.ctx, nw.data(), hw.data(), dk->n_words, p ? pw.data() : nullptr, q ? qw.data() : nullptr,
                      dk->n_words / 2, dk->key_bits / 2, 0, &dk->key));
  }
  dk->has_sk = p && q;
  return dk;
}

const DeviceKey& dev(const PublicKey& pk) {
  if (!pk.dev_) throw_msg("PublicKey is not initialized (call Init())");
  return *pk.dev_;
}

// Evaluator VALIDATE (evaluator.cc:20-22): 0 <= c < n^2, packed into rows.
Staged<uint32_t> pack(const DeviceKey& dk, ConstSpan<Ciphertext> cts) {
  const size_t W = dk.ct_words;
  Staged<uint32_t> flat(cts.size() * W);
  par_for(cts.size(), kGrain, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) {
      const auto& c = cts[i]->c_;
      if (c.size() != W) throw_msg("Evaluator: Invalid ciphertext (wrong limb count)");
      int cmp = 0;
      for (size_t j = W; j-- > 0 && cmp == 0;) cmp = c[j] < dk.n2[j] ? -1 : (c[j] > dk.n2[j] ? 1 : 0);
      if (cmp >= 0) throw_msg("Evaluator: Invalid ciphertext");
      std::memcpy(flat.data() + i * W, c.data(), W * 4);
    }
  });
  return flat;
}

Staged<uint32_t> pack(const DeviceKey& dk, Span<Ciphertext> cts) {
  return pack(dk, ConstSpan<Ciphertext>(cts.data(), cts.size()));
}

std::vector<Ciphertext> unpack(const DeviceKey& dk, const uint32_t* flat, size_t count) {
  std::vector<Ciphertext> out(count);
  const size_t W = dk.ct_words;
  par_for(count, kGrain, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) out[i].c_.assign(flat + i * W, flat + (i + 1) * W);
  });
  return out;
}

std::vector<Ciphertext> unpack(const DeviceKey& dk, const Staged<uint32_t>& flat, size_t count) {
  return unpack(dk, flat.data(), count);
}

// Plaintexts as int64 when all fit, else two's complement rows of n_words.
struct PtBatch {
  bool small = true;
  Staged<int64_t> i64;
  std::vector<uint32_t> rows;
  uint32_t words = 0;
};

PtBatch pack_pts(const DeviceKey& dk, ConstSpan<Plaintext> pts, const BigInt* bound, bool negate) {
  PtBatch b;
  const size_t n = pts.size();
  std::atomic<bool> small{true};
  par_for(n, kGrain, [&](size_t lo, size_t hi) {
    bool s = true;
    for (size_t i = lo; i < hi; ++i) {
      const Plaintext* p = pts[i];
      if (bound && p->CompareAbs(*bound) > 0)
        throw_msg("plaintext out of range, message=" + p->ToHexString() + ", max (abs)=" + bound->ToHexString());
      s = s && p->FitsInt64() && !(negate && p->ToInt64() == std::numeric_limits<int64_t>::min());
    }
    if (!s) small = false;
  });
  b.small = small;
  if (b.small) {
    b.i64 = Staged<int64_t>(n);
    par_for(n, kGrain, [&](size_t lo, size_t hi) {
      for (size_t i = lo; i < hi; ++i) b.i64[i] = negate ? -pts[i]->ToInt64() : pts[i]->ToInt64();
    });
    return b;
  }
  b.words = dk.n_words;
  b.rows.resize(n * b.words);
  par_for(n, kGrain, [&](size_t lo, size_t hi) {
    for (size_t i = lo; i < hi; ++i) (negate ? -*pts[i] : *pts[i]).ToWords(b.rows.data() + i * b.words, b.words);
  });
  return b;
}

template <class T>
std::vector<const T*> ptrs(const std::vector<T>& v) {
  std::vector<const T*> out(v.size());
  for (size_t i = 0; i < v.size(); ++i) out[i] = &v[i];
  return out;
}

template <class T>
ConstSpan<T> cspan(const std::vector<const T*>& v) {
  return ConstSpan<T>(v.data(), v.size());
}

void check_same(size_t a, size_t b) {
  if (a != b) throw_msg("operand spans differ in length: " + std::to_string(a) + " vs " + std::to_string(b));
}

// ct * (1 + n*p): int64 plaintexts through heu_add_i64; longer ones become
// (1 + n*p) mod n^2 = Enc(p, r = 0) first, then one ct+ct product.
std::vector<Ciphertext> add_ct_pt(const DeviceKey& dk, const BigInt& bound, ConstSpan<Ciphertext> a,
                                  ConstSpan<Plaintext> b, bool negate_b) {
  check_same(a.size(), b.size());
  const size_t cnt = a.size();
  if (cnt == 0) return {};
  auto ca = pack(dk, a);
  PtBatch pb = pack_pts(dk, b, &bound, negate_b);
  Staged<uint32_t> res(ca.size());
  if (pb.small) {
    Job j = make_job(kJobAddI64, dk.key, 0, cnt, res.data(), dk.ct_words * 4);
    add_in(j, ca.data(), dk.ct_words * 4);
    add_in(j, pb.i64.data(), 8);
    submit(j);
    return unpack(dk, res, cnt);
  }
  // plaintexts beyond int64
  std::vector<uint32_t> rz(cnt, 0u);  // r = 0: h_s^0 = 1
  DeviceCall call;
  size_t hA = call.in(ca.data(), ca.size() * 4), hO = call.out(res.size() * 4);
  size_t hP = call.in(pb.rows.data(), pb.rows.size() * 4), hR = call.in(rz.data(), cnt * 4);
  size_t hG = call.out(res.size() * 4);
  call.upload();
  ck(heu_encrypt(dk.key, call.ptr<uint32_t>(hP), pb.rows.data(), pb.words, call.ptr<uint32_t>(hR), 1, cnt,
                 call.ptr<uint32_t>(hG)));
  ck(heu_add(dk.key, call.ptr<uint32_t>(hA), call.ptr<uint32_t>(hG), cnt, call.ptr<uint32_t>(hO)));
  call.download(hO, res.data(), res.size() * 4);
  call.finish();
  return unpack(dk, res, cnt);
}

std::vector<Ciphertext> ct_unary(const DeviceKey& dk, ConstSpan<Ciphertext> a, int op) {
  const size_t cnt = a.size();
  if (cnt == 0) return {};
  auto ca = pack(dk, a);
  Staged<uint32_t> res(ca.size());
  Job j = make_job(op, dk.key, 0, cnt, res.data(), dk.ct_words * 4);
  add_in(j, ca.data(), dk.ct_words * 4);
  submit(j);
  return unpack(dk, res, cnt);
}

std::vector<Ciphertext> ct_binary(const DeviceKey& dk, ConstSpan<Ciphertext> a, ConstSpan<Ciphertext> b, int op) {
  check_same(a.size(), b.size());
  const size_t cnt = a.size();
  if (cnt == 0) return {};
  auto ca = pack(dk, a), cb = pack(dk, b);
  Staged<uint32_t> res(ca.size());
  Job j = make_job(op, dk.key, 0, cnt, res.data(), dk.ct_words * 4);
  add_in(j, ca.data(), dk.ct_words * 4);
  add_in(j, cb.data(), dk.ct_words * 4);
  submit(j);
  return unpack(dk, res, cnt);
}

template <class T>
void write_back(Span<T> dst, std::vector<T>&& src) {
  for (size_t i = 0; i < dst.size(); ++i) *dst[i] = std::move(src[i]);
}

std::string hex_words(const std::vector<uint32_t>& w) { return BigInt::FromWords(w.data(), w.size(), false).ToHexString(); }

// x R or x R^-1 mod m for R = 2^(64 ceil(bits(m) / 64)): the Montgomery form of
// yacl's 64-bit-word backends (MontgomerySpace::MapIntoMSpace / MapBackToZSpace).
std::vector<uint32_t> mont_words(const std::vector<uint32_t>& x, const BigInt& m, bool into) {
  BN_CTX* ctx = BN_CTX_new();
  BIGNUM *R = BN_new(), *v = BN_new();
  BN_set_bit(R, 64 * ((BN_num_bits(m.bn()) + 63) / 64));
  BN_nnmod(R, R, m.bn(), ctx);
  if (!into) BN_mod_inverse(R, R, m.bn(), ctx);
  BN_lebin2bn(reinterpret_cast<const unsigned char*>(x.data()), (int)(4 * x.size()), v);
  BN_mod_mul(v, v, R, m.bn(), ctx);
  std::vector<uint32_t> out(x.size());
  BN_bn2lebinpad(v, reinterpret_cast<unsigned char*>(out.data()), (int)(4 * out.size()));
  BN_free(R);
  BN_free(v);
  BN_CTX_free(ctx);
  return out;
}

// The 8-byte little-endian schema index HEU's variant wrapper appends
// (phe/base/serializable_types.cc:50-70); the engine's wire readers expect it.
constexpr uint8_t kSchemaIndex[8] = {HEU_SCHEMA_ZPAILLIER, 0, 0, 0, 0, 0, 0, 0};

// Runs a heu_wire_put_* writer into a vector and drops the schema index: these
// are the backend object's own bytes.
template <class F>
std::vector<uint8_t> wire_bytes(F&& put) {
  size_t len = 0;
  (void)put(nullptr, 0, &len);  // HEU_ERANGE, with the length
  std::vector<uint8_t> out(len);
  ck(put(out.data(), len, &len));
  if (len < sizeof(kSchemaIndex)) throw_msg("wire writer returned too few bytes");
  out.resize(len - sizeof(kSchemaIndex));
  return out;
}

std::vector<uint8_t> with_index(std::span<const uint8_t> in) {
  std::vector<uint8_t> v(in.begin(), in.end());
  v.insert(v.end(), kSchemaIndex, kSchemaIndex + sizeof(kSchemaIndex));
  return v;
}

// A non-negative value as `words` little-endian limbs (unsigned, unlike ToWords).
std::vector<uint32_t> big_words(const BigInt& v, size_t words) {
  if (v.IsNegative() || v.BitCount() > 32 * words) throw_msg("key value does not fit its limbs");
  std::vector<uint32_t> w(words);
  bn_check(BN_bn2lebinpad(v.bn(), reinterpret_cast<unsigned char*>(w.data()), (int)(4 * words)) >= 0, "bn2lebinpad");
  return w;
}

}  // namespace

// ------------------------------------------------------------------ keys
std::string Ciphertext::ToString() const { return hex_words(c_); }

std::vector<uint8_t> Ciphertext::Serialize(const PublicKey& pk) const {
  if (c_.empty()) throw_msg("Ciphertext::Serialize: empty ciphertext");
  const auto m = mont_words(c_, pk.n_square_, true);
  return wire_bytes([&](uint8_t* o, size_t cap, size_t* len) {
    return heu_wire_put_ciphertext(m.data(), (uint32_t)m.size(), o, cap, len);
  });
}

Ciphertext Ciphertext::Deserialize(const PublicKey& pk, std::span<const uint8_t> in) {
  const auto buf = with_index(in);
  std::vector<uint32_t> m(pk.ct_words());
  ck(heu_wire_get_ciphertext(buf.data(), buf.size(), m.data(), (uint32_t)m.size()));
  if (BigInt::FromWords(m.data(), m.size(), false).Compare(pk.n_square_) >= 0)
    throw_msg("Ciphertext::Deserialize: value is not below n^2");
  return Ciphertext(mont_words(m, pk.n_square_, false));
}

std::vector<uint8_t> PublicKey::Serialize() const {
  const size_t nw = (n_.BitCount() + 31) / 32;
  const auto n = big_words(n_, nw), h = big_words(h_s_, 2 * nw);
  return wire_bytes([&](uint8_t* o, size_t cap, size_t* len) {
    return heu_wire_put_public_key(n.data(), (uint32_t)nw, h.data(), o, cap, len);
  });
}

void PublicKey::Deserialize(std::span<const uint8_t> in) {
  const auto buf = with_index(in);
  uint32_t nw = 0;
  ck(heu_wire_get_public_key(buf.data(), buf.size(), nullptr, nullptr, 0, &nw));
  std::vector<uint32_t> n(nw), h(2 * (size_t)nw);
  ck(heu_wire_get_public_key(buf.data(), buf.size(), n.data(), h.data(), nw, &nw));
  *this = PublicKey();
  n_ = BigInt::FromWords(n.data(), n.size(), false);
  h_s_ = BigInt::FromWords(h.data(), h.size(), false);
  Init();
}

std::vector<uint8_t> SecretKey::Serialize() const {
  const size_t pw = (std::max(p_.BitCount(), q_.BitCount()) + 31) / 32;
  const auto p = big_words(p_, pw), q = big_words(q_, pw);
  return wire_bytes([&](uint8_t* o, size_t cap, size_t* len) {
    return heu_wire_put_secret_key(p.data(), q.data(), (uint32_t)pw, o, cap, len);
  });
}

void SecretKey::Deserialize(std::span<const uint8_t> in) {
  const auto buf = with_index(in);
  uint32_t pw = 0;
  ck(heu_wire_get_secret_key(buf.data(), buf.size(), nullptr, nullptr, 0, &pw));
  std::vector<uint32_t> p(pw), q(pw);
  ck(heu_wire_get_secret_key(buf.data(), buf.size(), p.data(), q.data(), pw, &pw));
  *this = SecretKey();
  p_ = BigInt::FromWords(p.data(), pw, false);
  q_ = BigInt::FromWords(q.data(), pw, false);
  lambda_ = ((p_ - BigInt(1)) * (q_ - BigInt(1))) >> 1;
  const BigInt n = p_ * q_;
  bn_check(BN_mod_inverse(mu_.bn(), lambda_.bn(), n.bn(), tls_ctx()) != nullptr, "mod_inverse");
}

void PublicKey::Init() {
  if (key_size_ == 0) key_size_ = n_.BitCount();
  n_square_ = n_ * n_;
  n_half_ = n_ >> 1;
  if (!dev_) dev_ = make_device_key(n_, h_s_, nullptr, nullptr, key_size_);
}

size_t PublicKey::ct_words() const { return dev(*this).ct_words; }
size_t PublicKey::n_words() const { return dev(*this).n_words; }

std::string PublicKey::ToString() const {
  return "paillier_hip PK: n=" + n_.ToHexString() + " (" + std::to_string(key_size_) + " bits), h_s=" +
         h_s_.ToHexString();
}

void SecretKey::Init(const PublicKey& pk) {
  if (pk.dev_ && pk.dev_->has_sk) {
    dev_ = pk.dev_;
    return;
  }
  dev_ = make_device_key(pk.n_, pk.h_s_, &p_, &q_, pk.key_size_);
}

std::string SecretKey::ToString() const {
  return "paillier_hip SK: p=" + p_.ToHexString() + ", q=" + q_.ToHexString();
}

namespace {

void fill_keys(size_t key_size, const uint64_t* seed, SecretKey* sk, PublicKey* pk) {
  if (key_size % 64 != 0) throw_msg("Key size must be a multiple of 64");
  const size_t nw = key_size / 32;
  std::vector<uint32_t> n(nw), hs(2 * nw), p(nw / 2), q(nw / 2);
  ck(heu_keygen((uint32_t)key_size, seed, n.data(), hs.data(), p.data(), q.data()));
  pk->n_ = BigInt::FromWords(n.data(), nw, false);
  pk->h_s_ = BigInt::FromWords(hs.data(), 2 * nw, false);
  pk->key_size_ = key_size;
  sk->p_ = BigInt::FromWords(p.data(), nw / 2, false);
  sk->q_ = BigInt::FromWords(q.data(), nw / 2, false);
  // key_generator.cc:50-51: lambda = (p-1)(q-1)/2, mu = lambda^-1 mod n
  BigInt one(1);
  sk->lambda_ = ((sk->p_ - one) * (sk->q_ - one)) >> 1;
  bn_check(BN_mod_inverse(sk->mu_.bn(), sk->lambda_.bn(), pk->n_.bn(), tls_ctx()) != nullptr, "mod_inverse");
  // one device key (constants, CRT halves, fixed-base table) serves both
  pk->n_square_ = pk->n_ * pk->n_;
  pk->n_half_ = pk->n_ >> 1;
  pk->dev_ = make_device_key(pk->n_, pk->h_s_, &sk->p_, &sk->q_, key_size);
  sk->dev_ = pk->dev_;
}

}  // namespace

void KeyGenerator::Generate(size_t key_size, SecretKey* sk, PublicKey* pk) { fill_keys(key_size, nullptr, sk, pk); }

void KeyGenerator::GenerateSeeded(size_t key_size, uint64_t seed, SecretKey* sk, PublicKey* pk) {
  fill_keys(key_size, &seed, sk, pk);
}

// ------------------------------------------------------------------ Encryptor
std::vector<Ciphertext> Encryptor::EncryptZero(int64_t size) const {
  const DeviceKey& dk = dev(pk_);
  if (size < 0) throw_msg("EncryptZero: negative size");
  const size_t cnt = (size_t)size;
  if (cnt == 0) return {};
  auto r = random_r(dk, cnt);
  Staged<uint32_t> res(cnt * dk.ct_words);
  Job j = make_job(kJobEncZero, dk.key, dk.r_words, cnt, res.data(), dk.ct_words * 4);
  add_in(j, r.data(), dk.r_words * 4);
  submit(j);
  return unpack(dk, res, cnt);
}

std::vector<Ciphertext> Encryptor::EncryptWithR(ConstSpan<Plaintext> pts, const uint32_t* r, size_t r_words) const {
  const DeviceKey& dk = dev(pk_);
  const size_t cnt = pts.size();
  if (cnt == 0) return {};
  if (r_words == 0 || !r) throw_msg("EncryptWithR: empty randomness");
  PtBatch pb = pack_pts(dk, pts, &pk_.n_half_, false);
  Staged<uint32_t> res(cnt * dk.ct_words);
  if (pb.small) {
    Job j = make_job(kJobEncI64, dk.key, (uint32_t)r_words, cnt, res.data(), dk.ct_words * 4);
    add_in(j, pb.i64.data(), 8);
    add_in(j, r, r_words * 4);
    submit(j);
    return unpack(dk, res, cnt);
  }
  // plaintexts beyond int64
  DeviceCall call;
  size_t hR = call.in(r, cnt * r_words * 4), hO = call.out(res.size() * 4);
  size_t hM = call.in(pb.rows.data(), pb.rows.size() * 4);
  call.upload();
  ck(heu_encrypt(dk.key, call.ptr<uint32_t>(hM), pb.rows.data(), pb.words, call.ptr<uint32_t>(hR),
                 (uint32_t)r_words, cnt, call.ptr<uint32_t>(hO)));
  call.download(hO, res.data(), res.size() * 4);
  call.finish();
  return unpack(dk, res, cnt);
}

std::vector<Ciphertext> Encryptor::Encrypt(ConstSpan<Plaintext> pts) const {
  const DeviceKey& dk = dev(pk_);
  // range check first (encryptor.cc:38-40), so no randomness is drawn for a bad batch
  par_for(pts.size(), kGrain, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i)
      if (pts[i]->CompareAbs(pk_.n_half_) > 0)
        throw_msg("message number out of range, message=" + pts[i]->ToHexString() + ", max (abs)=" +
                  pk_.n_half_.ToHexString());
  });
  auto r = random_r(dk, pts.size());
  return EncryptWithR(pts, r.data(), dk.r_words);
}

// encryptor.cc:50-54: the audit string records p, rn = h_s^r and c in hex, rn
// and c in the Montgomery form the reference prints (R for 64-bit words).
std::pair<std::vector<Ciphertext>, std::vector<std::string>> Encryptor::EncryptWithAudit(
    ConstSpan<Plaintext> pts) const {
  const DeviceKey& dk = dev(pk_);
  const size_t cnt = pts.size();
  auto r = random_r(dk, cnt);
  auto cts = EncryptWithR(pts, r.data(), dk.r_words);
  std::vector<std::string> audit(cnt);
  if (cnt == 0) return {cts, audit};
  Staged<uint32_t> rn(cnt * dk.ct_words);
  {
    Job j = make_job(kJobEncZero, dk.key, dk.r_words, cnt, rn.data(), dk.ct_words * 4);
    add_in(j, r.data(), dk.r_words * 4);
    submit(j);
  }
  for (size_t i = 0; i < cnt; ++i) {
    std::vector<uint32_t> one(rn.data() + i * dk.ct_words, rn.data() + (i + 1) * dk.ct_words);
    audit[i] = "p:" + pts[i]->ToHexString() + ",rn:" + hex_words(mont_words(one, pk_.n_square_, true)) +
               ",c:" + hex_words(mont_words(cts[i].c_, pk_.n_square_, true));
  }
  return {cts, audit};
}

Ciphertext Encryptor::EncryptZero() const { return EncryptZero(1)[0]; }

Ciphertext Encryptor::Encrypt(const Plaintext& m) const {
  const Plaintext* p = &m;
  return Encrypt(ConstSpan<Plaintext>(&p, 1))[0];
}

std::pair<Ciphertext, std::string> Encryptor::EncryptWithAudit(const Plaintext& m) const {
  const Plaintext* p = &m;
  auto r = EncryptWithAudit(ConstSpan<Plaintext>(&p, 1));
  return {r.first[0], r.second[0]};
}

// ------------------------------------------------------------------ Decryptor
Decryptor::Decryptor(const PublicKey& pk, const SecretKey& sk) : pk_(pk), sk_(sk) {
  if (!sk_.dev_ || !sk_.dev_->has_sk) throw_msg("Decryptor: secret key is not initialized (call Init(pk))");
  if (!(sk_.p_ * sk_.q_ == pk_.n_)) throw_msg("Decryptor: pk and sk are not paired");
}

void Decryptor::Decrypt(ConstSpan<Ciphertext> in_cts, Span<Plaintext> out_pts) const {
  check_same(in_cts.size(), out_pts.size());
  const DeviceKey& dk = *sk_.dev_;
  const size_t cnt = in_cts.size();
  if (cnt == 0) return;
  auto ca = pack(dk, in_cts);
  const uint32_t ow = dk.n_words;  // |m| <= n/2 always fits n_words in two's complement
  Staged<uint32_t> res(cnt * ow);
  Job j = make_job(kJobDecrypt, dk.key, ow, cnt, res.data(), ow * 4);
  add_in(j, ca.data(), dk.ct_words * 4);
  submit(j);
  par_for(cnt, kGrain, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) *out_pts[i] = BigInt::FromWords(res.data() + i * ow, ow, true);
  });
}

std::vector<Plaintext> Decryptor::Decrypt(ConstSpan<Ciphertext> cts) const {
  std::vector<Plaintext> out(cts.size());
  std::vector<Plaintext*> op(out.size());
  for (size_t i = 0; i < out.size(); ++i) op[i] = &out[i];
  Decrypt(cts, Span<Plaintext>(op.data(), op.size()));
  return out;
}

void Decryptor::Decrypt(const Ciphertext& ct, Plaintext* out) const {
  const Ciphertext* c = &ct;
  Decrypt(ConstSpan<Ciphertext>(&c, 1), Span<Plaintext>(&out, 1));
}

Plaintext Decryptor::Decrypt(const Ciphertext& ct) const {
  Plaintext p;
  Decrypt(ct, &p);
  return p;
}

// ------------------------------------------------------------------ Evaluator
void Evaluator::Randomize(Span<Ciphertext> ct) const {
  const DeviceKey& dk = dev(pk_);
  const size_t cnt = ct.size();
  if (cnt == 0) return;
  auto ca = pack(dk, ct);
  auto r = random_r(dk, cnt);
  Staged<uint32_t> res(ca.size());
  Job j = make_job(kJobRandomize, dk.key, dk.r_words, cnt, res.data(), dk.ct_words * 4);
  add_in(j, ca.data(), dk.ct_words * 4);
  add_in(j, r.data(), dk.r_words * 4);
  submit(j);
  write_back(ct, unpack(dk, res, cnt));
}

std::vector<Ciphertext> Evaluator::Add(ConstSpan<Ciphertext> a, ConstSpan<Ciphertext> b) const {
  const DeviceKey& dk = dev(pk_);
  return ct_binary(dk, a, b, kJobAdd);
}

std::vector<Ciphertext> Evaluator::Add(ConstSpan<Ciphertext> a, ConstSpan<Plaintext> b) const {
  return add_ct_pt(dev(pk_), pk_.n_half_, a, b, false);
}

std::vector<Ciphertext> Evaluator::Add(ConstSpan<Plaintext> a, ConstSpan<Ciphertext> b) const { return Add(b, a); }

std::vector<Plaintext> Evaluator::Add(ConstSpan<Plaintext> a, ConstSpan<Plaintext> b) const {
  check_same(a.size(), b.size());
  std::vector<Plaintext> out;
  out.reserve(a.size());
  for (size_t i = 0; i < a.size(); ++i) out.push_back(*a[i] + *b[i]);
  return out;
}

void Evaluator::AddInplace(Span<Ciphertext> a, ConstSpan<Ciphertext> b) const {
  write_back(a, Add(ConstSpan<Ciphertext>(a.data(), a.size()), b));
}

void Evaluator::AddInplace(Span<Ciphertext> a, ConstSpan<Plaintext> b) const {
  write_back(a, Add(ConstSpan<Ciphertext>(a.data(), a.size()), b));
}

void Evaluator::AddInplace(Span<Plaintext> a, ConstSpan<Plaintext> b) const {
  write_back(a, Add(ConstSpan<Plaintext>(a.data(), a.size()), b));
}

// evaluator.cc:66-80: Sub(a, b) = Add(a, Negate(b)), Sub(ct, p) = Add(ct, -p),
// Sub(p, ct) = Add(Negate(ct), p).
std::vector<Ciphertext> Evaluator::Sub(ConstSpan<Ciphertext> a, ConstSpan<Ciphertext> b) const {
  const DeviceKey& dk = dev(pk_);
  return ct_binary(dk, a, b, kJobSub);
}

std::vector<Ciphertext> Evaluator::Sub(ConstSpan<Ciphertext> a, ConstSpan<Plaintext> b) const {
  return add_ct_pt(dev(pk_), pk_.n_half_, a, b, true);
}

std::vector<Ciphertext> Evaluator::Sub(ConstSpan<Plaintext> a, ConstSpan<Ciphertext> b) const {
  auto nb = Negate(b);
  return Add(cspan(ptrs(nb)), a);
}

std::vector<Plaintext> Evaluator::Sub(ConstSpan<Plaintext> a, ConstSpan<Plaintext> b) const {
  check_same(a.size(), b.size());
  std::vector<Plaintext> out;
  out.reserve(a.size());
  for (size_t i = 0; i < a.size(); ++i) out.push_back(*a[i] - *b[i]);
  return out;
}

void Evaluator::SubInplace(Span<Ciphertext> a, ConstSpan<Ciphertext> b) const {
  write_back(a, Sub(ConstSpan<Ciphertext>(a.data(), a.size()), b));
}

void Evaluator::SubInplace(Span<Ciphertext> a, ConstSpan<Plaintext> p) const {
  write_back(a, Sub(ConstSpan<Ciphertext>(a.data(), a.size()), p));
}

void Evaluator::SubInplace(Span<Plaintext> a, ConstSpan<Plaintext> b) const {
  write_back(a, Sub(ConstSpan<Plaintext>(a.data(), a.size()), b));
}

// evaluator.cc:99-123 (no range check on the scalar: ciphertext overflow is allowed)
std::vector<Ciphertext> Evaluator::Mul(ConstSpan<Ciphertext> a, ConstSpan<Plaintext> b) const {
  const DeviceKey& dk = dev(pk_);
  check_same(a.size(), b.size());
  const size_t cnt = a.size();
  if (cnt == 0) return {};
  std::atomic<bool> small{true};
  std::atomic<size_t> bits{1};
  par_for(cnt, kGrain, [&](size_t lo, size_t hi) {
    bool s = true;
    size_t nb = 1;
    for (size_t i = lo; i < hi; ++i) {
      s = s && b[i]->FitsInt64();
      nb = std::max(nb, b[i]->BitCount());
    }
    if (!s) small = false;
    for (size_t cur = bits.load(); nb > cur && !bits.compare_exchange_weak(cur, nb);) {
    }
  });
  auto ca = pack(dk, a);
  Staged<uint32_t> res(ca.size());
  if (small) {
    Staged<int64_t> k64(cnt);
    par_for(cnt, kGrain, [&](size_t lo, size_t hi) {
      for (size_t i = lo; i < hi; ++i) k64[i] = b[i]->ToInt64();
    });
    Job j = make_job(kJobMulI64, dk.key, 0, cnt, res.data(), dk.ct_words * 4);
    add_in(j, ca.data(), dk.ct_words * 4);
    add_in(j, k64.data(), 8);
    submit(j);
    return unpack(dk, res, cnt);
  }
  // scalars beyond int64
  const uint32_t kw = (uint32_t)((bits + 31) / 32);
  std::vector<uint32_t> mag(cnt * kw, 0u);
  std::vector<uint8_t> neg(cnt);
  for (size_t i = 0; i < cnt; ++i) {
    bn_check(BN_bn2lebinpad(b[i]->bn(), reinterpret_cast<unsigned char*>(mag.data() + i * kw), (int)kw * 4) >= 0,
             "bn2lebinpad");
    neg[i] = b[i]->IsNegative() ? 1 : 0;
  }
  DeviceCall call;
  size_t hA = call.in(ca.data(), ca.size() * 4), hO = call.out(res.size() * 4);
  size_t hK = call.in(mag.data(), mag.size() * 4), hS = call.in(neg.data(), cnt);
  call.upload();
  ck(heu_mul(dk.key, call.ptr<uint32_t>(hA), call.ptr<uint32_t>(hK), call.ptr<uint8_t>(hS), kw, cnt,
             call.ptr<uint32_t>(hO)));
  call.download(hO, res.data(), res.size() * 4);
  call.finish();
  return unpack(dk, res, cnt);
}

std::vector<Ciphertext> Evaluator::Mul(ConstSpan<Plaintext> a, ConstSpan<Ciphertext> b) const { return Mul(b, a); }

std::vector<Plaintext> Evaluator::Mul(ConstSpan<Plaintext> a, ConstSpan<Plaintext> b) const {
  check_same(a.size(), b.size());
  std::vector<Plaintext> out;
  out.reserve(a.size());
  for (size_t i = 0; i < a.size(); ++i) out.push_back(*a[i] * *b[i]);
  return out;
}

void Evaluator::MulInplace(Span<Ciphertext> a, ConstSpan<Plaintext> b) const {
  write_back(a, Mul(ConstSpan<Ciphertext>(a.data(), a.size()), b));
}

void Evaluator::MulInplace(Span<Plaintext> a, ConstSpan<Plaintext> b) const {
  write_back(a, Mul(ConstSpan<Plaintext>(a.data(), a.size()), b));
}

std::vector<Ciphertext> Evaluator::Negate(ConstSpan<Ciphertext> a) const {
  const DeviceKey& dk = dev(pk_);
  return ct_unary(dk, a, kJobNegate);
}

void Evaluator::NegateInplace(Span<Ciphertext> a) const { write_back(a, Negate(ConstSpan<Ciphertext>(a.data(), a.size()))); }

Ciphertext Evaluator::ReduceSum(ConstSpan<Ciphertext> a) const {
  const DeviceKey& dk = dev(pk_);
  auto ca = pack(dk, a);
  std::vector<uint32_t> res(dk.ct_words);
  const int64_t offs[2] = {0, (int64_t)a.size()};
  DeviceCall call;
  size_t hA = call.in(ca.data(), ca.size() * 4), hO = call.out(res.size() * 4);
  call.upload();
  ck(heu_sum(dk.key, call.ptr<uint32_t>(hA), offs, 1, call.ptr<uint32_t>(hO)));
  call.download(hO, res.data(), res.size() * 4);
  call.finish();
  return Ciphertext(std::move(res));
}

// single-element forms
namespace {
template <class T>
ConstSpan<T> one(const T& v, const T*& slot) {
  slot = &v;
  return ConstSpan<T>(&slot, 1);
}
}  // namespace

void Evaluator::Randomize(Ciphertext* ct) const { Randomize(Span<Ciphertext>(&ct, 1)); }

Ciphertext Evaluator::Add(const Ciphertext& a, const Ciphertext& b) const {
  const Ciphertext *x, *y;
  return Add(one(a, x), one(b, y))[0];
}

Ciphertext Evaluator::Add(const Ciphertext& a, const Plaintext& b) const {
  const Ciphertext* x;
  const Plaintext* y;
  return Add(one(a, x), one(b, y))[0];
}

void Evaluator::AddInplace(Ciphertext* a, const Ciphertext& b) const { *a = Add(*a, b); }
void Evaluator::AddInplace(Ciphertext* a, const Plaintext& b) const { *a = Add(*a, b); }

Ciphertext Evaluator::Sub(const Ciphertext& a, const Ciphertext& b) const {
  const Ciphertext *x, *y;
  return Sub(one(a, x), one(b, y))[0];
}

Ciphertext Evaluator::Sub(const Ciphertext& a, const Plaintext& b) const {
  const Ciphertext* x;
  const Plaintext* y;
  return Sub(one(a, x), one(b, y))[0];
}

Ciphertext Evaluator::Sub(const Plaintext& a, const Ciphertext& b) const {
  const Plaintext* x;
  const Ciphertext* y;
  return Sub(one(a, x), one(b, y))[0];
}

void Evaluator::SubInplace(Ciphertext* a, const Ciphertext& b) const { *a = Sub(*a, b); }
void Evaluator::SubInplace(Ciphertext* a, const Plaintext& b) const { *a = Sub(*a, b); }

Ciphertext Evaluator::Mul(const Ciphertext& a, const Plaintext& b) const {
  const Ciphertext* x;
  const Plaintext* y;
  return Mul(one(a, x), one(b, y))[0];
}

void Evaluator::MulInplace(Ciphertext* a, const Plaintext& b) const { *a = Mul(*a, b); }

Ciphertext Evaluator::Negate(const Ciphertext& a) const {
  const Ciphertext* x;
  return Negate(one(a, x))[0];
}

void Evaluator::NegateInplace(Ciphertext* a) const { *a = Negate(*a); }

}  // namespace paillier_hip
}  // namespace heu_amd
