// span_bench.cc -- throughput of the vectorized-SPI classes (paillier_hip.h)
// with host spans, which is what a HEU caller gets: the conversions between
// BigInt objects and limbs, one host-to-device and one device-to-host copy and
// a sync per engine call.  Beside it, the same engine entry points on
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
  // SPI, host spans (one warm-up round first: staging arena, engine scratch)
  std::vector<Ciphertext> c, s, p;
  std::vector<BigInt> d;
  c = enc.Encrypt(ConstSpan<BigInt>(ptrs(m)));
  d = dec.Decrypt(ConstSpan<Ciphertext>(ptrs(c)));
  const auto mp = ptrs(m), wp = ptrs(w);
  const double te = secs([&] { c = enc.Encrypt(ConstSpan<BigInt>(mp)); });
  const auto cp = ptrs(c);
  const double ta = secs([&] { s = ev.Add(ConstSpan<Ciphertext>(cp), ConstSpan<Ciphertext>(cp)); });
  const double tm = secs([&] { p = ev.Mul(ConstSpan<Ciphertext>(cp), ConstSpan<BigInt>(wp)); });
  const double td = secs([&] { d = dec.Decrypt(ConstSpan<Ciphertext>(cp)); });
  bool ok = d.size() == count;
  for (size_t i = 0; ok && i < count; ++i) ok = d[i] == m[i];

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
  run(e_dec);
  const double de = run(e_enc), da = run(e_add), dmul = run(e_mul), ddec = run(e_dec);
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
