// span_bench.cc -- throughput of the vectorized-SPI classes (paillier_hip.h)
// with host spans, which is what a HEU caller gets: the conversions between
// BigInt objects and limbs, one host-to-device and one device-to-host copy and
// a sync per engine call, in steady state (results of each round are dropped
// before the next, so the allocator reuses their pages).  Beside it, the same engine entry points on
// device-resident data for the same key and the same randomness width, so the
// host's share shows.  The SPI draws r of key_bits / 2 bits, as the
// reference's GetRn does (paillier_zahlen/encryptor.cc:26-27); the
// device-resident column uses the same width.  Paillier-2048, seeded key.
//
// Usage: span_bench [count]   (default 65536).  Prints one JSON object.
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <random>
#include <string>
#include <vector>

#include "heu_paillier.h"
#include "paillier_hip.h"

using heu_amd::BigInt;
using heu_amd::ConstSpan;
using namespace heu_amd::paillier_hip;

namespace {

void ck(int rc, const char* what) {
  if (rc != HEU_OK) {
    std::fprintf(stderr, "%s failed: %d %s\n", what, rc, heu_last_error());
    std::exit(1);
  }
}

void hip_ck(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    std::exit(1);
  }
}

double secs(const std::function<void()>& f) {
  const auto t0 = std::chrono::steady_clock::now();
  f();
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

template <class T>
std::vector<const T*> ptrs(const std::vector<T>& v) {
  std::vector<const T*> o;
  o.reserve(v.size());
  for (auto& x : v) o.push_back(&x);
  return o;
}

std::vector<uint32_t> words(const BigInt& v, size_t n) {
  std::vector<uint32_t> w(n + 1);
  v.ToWords(w.data(), n + 1);  // one spare word: a positive value's top bit may be set
  w.resize(n);
  return w;
}

}  // namespace

int main(int argc, char** argv) {
  const size_t count = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 65536;
  SecretKey sk;
  PublicKey pk;
  KeyGenerator::GenerateSeeded(2048, 0, &sk, &pk);
  sk.Init(pk);
  Encryptor enc(pk);
  Decryptor dec(pk, sk);
  Evaluator ev(pk);
  std::mt19937_64 g(5);
  std::vector<BigInt> m, w;
  std::vector<int64_t> m64(count), w64(count);
  for (size_t i = 0; i < count; ++i) {
    m64[i] = (int64_t)g();
    w64[i] = (int64_t)(int32_t)g();
    m.emplace_back(m64[i]);
    w.emplace_back(w64[i]);
  }
  // SPI, host spans, in steady state: each round builds its results into
  // fresh vectors and drops them at its end, as a streaming caller does; round
  // 0 warms up (staging buffers, engine scratch, the allocator's pages) and the
  // mean of the kReps rounds after it is reported
  constexpr int kReps = 3;
  const auto mp = ptrs(m), wp = ptrs(w);
  const std::vector<Ciphertext> c = enc.Encrypt(ConstSpan<BigInt>(mp));
  const auto cp = ptrs(c);
  double te = 0, ta = 0, tm = 0, td = 0;
  bool ok = true;
  for (int rep = 0; rep <= kReps; ++rep) {
    const double wgt = rep ? 1.0 / kReps : 0.0;
    std::vector<Ciphertext> c2, s2, p2;
    std::vector<BigInt> d2;
    te += wgt * secs([&] { c2 = enc.Encrypt(ConstSpan<BigInt>(mp)); });
    ta += wgt * secs([&] { s2 = ev.Add(ConstSpan<Ciphertext>(cp), ConstSpan<Ciphertext>(cp)); });
    tm += wgt * secs([&] { p2 = ev.Mul(ConstSpan<Ciphertext>(cp), ConstSpan<BigInt>(wp)); });
    const auto c2p = ptrs(c2);
    td += wgt * secs([&] { d2 = dec.Decrypt(ConstSpan<Ciphertext>(c2p)); });
    ok = ok && d2.size() == count && s2.size() == count && p2.size() == count;
    for (size_t i = 0; ok && i < count; ++i) ok = d2[i] == m[i];
  }

  // the engine on device-resident data: same key, r of the same width
  const uint32_t nw = 64, cw = 128, pw = 32, rw = 32;
  heu_ctx* ctx = nullptr;
  heu_key* key = nullptr;
  ck(heu_ctx_create(0, &ctx), "heu_ctx_create");
  const auto n = words(pk.n_, nw), hs = words(pk.h_s_, cw), pp = words(sk.p_, pw), qq = words(sk.q_, pw);
  ck(heu_key_create(ctx, n.data(), hs.data(), nw, pp.data(), qq.data(), pw, 1024, 0, &key), "heu_key_create");
  std::vector<uint32_t> r(count * rw);
  for (auto& x : r) x = (uint32_t)g();
  int64_t *dm = nullptr, *dw = nullptr;
  uint32_t *dr = nullptr, *dc = nullptr, *ds = nullptr, *dp = nullptr, *dd = nullptr;
  hip_ck(hipMalloc(&dm, count * 8), "hipMalloc");
  hip_ck(hipMalloc(&dw, count * 8), "hipMalloc");
  hip_ck(hipMalloc(&dr, r.size() * 4), "hipMalloc");
  hip_ck(hipMalloc(&dc, count * cw * 4), "hipMalloc");
  hip_ck(hipMalloc(&ds, count * cw * 4), "hipMalloc");
  hip_ck(hipMalloc(&dp, count * cw * 4), "hipMalloc");
  hip_ck(hipMalloc(&dd, count * nw * 4), "hipMalloc");
  hip_ck(hipMemcpy(dm, m64.data(), count * 8, hipMemcpyHostToDevice), "hipMemcpy");
  hip_ck(hipMemcpy(dw, w64.data(), count * 8, hipMemcpyHostToDevice), "hipMemcpy");
  hip_ck(hipMemcpy(dr, r.data(), r.size() * 4, hipMemcpyHostToDevice), "hipMemcpy");
  auto run = [&](const std::function<int()>& f) {
    return secs([&] {
      ck(f(), "engine call");
      ck(heu_ctx_sync(ctx), "heu_ctx_sync");
    });
  };
  auto e_enc = [&] { return heu_encrypt_i64(key, dm, dr, rw, count, dc); };
  auto e_add = [&] { return heu_add(key, dc, dc, count, ds); };
  auto e_mul = [&] { return heu_mul_i64(key, dc, dw, count, dp); };
  auto e_dec = [&] { return heu_decrypt(key, dc, count, dd, nw, nullptr); };
  run(e_enc);  // warm-up
  run(e_add);
  run(e_mul);
  run(e_dec);
  double de = 0, da = 0, dmul = 0, ddec = 0;
  for (int rep = 0; rep < kReps; ++rep) {
    de += run(e_enc) / kReps;
    da += run(e_add) / kReps;
    dmul += run(e_mul) / kReps;
    ddec += run(e_dec) / kReps;
  }
  std::printf("{\"elements\": %zu, \"key_bits\": 2048, \"r_bits\": 1024, \"decrypt_ok\": %s, "
              "\"spi_host_spans_per_s\": {\"enc\": %.1f, \"add\": %.1f, \"ctpt\": %.1f, \"dec\": %.1f}, "
              "\"engine_device_resident_per_s\": {\"enc\": %.1f, \"add\": %.1f, \"ctpt\": %.1f, \"dec\": %.1f}}\n",
              count, ok ? "true" : "false", count / te, count / ta, count / tm, count / td, count / de, count / da,
              count / dmul, count / ddec);
  for (void* q : {(void*)dm, (void*)dw, (void*)dr, (void*)dc, (void*)ds, (void*)dp, (void*)dd}) (void)hipFree(q);
  heu_key_destroy(key);
  heu_ctx_destroy(ctx);
  return ok ? 0 : 1;
}
