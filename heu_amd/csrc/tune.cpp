// tune.cpp -- the autotuner (heu_autotune in include/heu_paillier.h).
//
// For one key it times, on synthetic inputs and on the device at hand, the
// candidates of three parameters and keeps the fastest of each:
//   - the fixed-base table window (encryption throughput; candidates whose
//     table fits HEU_TABLE_BYTES, default 4 GiB);
//   - the lanes per bignum of the CRT decryption (decryption throughput);
//   - the grid multiplier of the persistent work-list launches, i.e. how many
//     times the resident workgroups per CU a launch starts (occupancy), for the
//     key's n^2 shape (ct x scalar) and decryption shape.
// Every candidate runs once to warm up and then twice; the best time counts.
// A decryption shape with more lanes replaces the default one, and a grid
// multiplier above 1 replaces 1, only when it is at least 2% faster (the
// candidates are usually within timing noise of each other).
// The parameters never change results, only speed.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "engine_internal.hpp"

using namespace heu_engine;

namespace {

uint64_t mix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Best of two timed runs of `fn` after one warm-up, on the context's stream (ms).
template <class F>
int best_ms(heu_ctx* ctx, F&& fn, double* out) {
  hipEvent_t a = nullptr, b = nullptr;
  HIP_TRY(hipEventCreate(&a));
  HIP_TRY(hipEventCreate(&b));
  int rc = fn();
  if (!rc) rc = heu_ctx_sync(ctx);
  double best = 1e30;
  for (int rep = 0; rep < 2 && !rc; ++rep) {
    (void)hipEventRecord(a, ctx->stream);
    rc = fn();
    (void)hipEventRecord(b, ctx->stream);
    if (!rc) rc = heu_ctx_sync(ctx);
    float ms = 0;
    if (!rc && hipEventElapsedTime(&ms, a, b) == hipSuccess) best = std::min(best, (double)ms);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *out = best;
  return rc;
}

std::string fmt(const char* f, double v) {
  char buf[64];
  snprintf(buf, sizeof(buf), f, v);
  return buf;
}

}  // namespace

extern "C" int heu_autotune(heu_key* key, uint32_t what, char* report, size_t report_len) {
  if (!key || !key->ops) return fail(HEU_EINVAL, "null key");
  heu_ctx* ctx = key->ctx;
  HIP_TRY(hipSetDevice(ctx->device));
  const int bits = key->key_bits;
  const long n_enc = bits == 2048 ? 131072 : 65536;
  const long n_dec = bits == 2048 ? 65536 : 32768;
  const long n = std::max(n_enc, n_dec);
  const uint32_t rw = (key->r_bits + 31) / 32;
  // synthetic inputs
  std::vector<int64_t> m(n), k(n);
  std::vector<uint32_t> r((size_t)n * rw);
  uint64_t s = 0x5eed + (uint64_t)bits;
  for (long i = 0; i < n; ++i) {
    m[i] = (int64_t)mix(s);
    k[i] = (int64_t)(int32_t)mix(s);
    for (uint32_t w = 0; w < rw; ++w) r[(size_t)i * rw + w] = (uint32_t)mix(s);
    if (key->r_bits % 32) r[(size_t)i * rw + rw - 1] &= (1u << (key->r_bits % 32)) - 1u;
  }
  DevBuf dm, dk, dr, dct, dout;
  int rc;
  if ((rc = dm.alloc(n * 8)) || (rc = dk.alloc(n * 8)) || (rc = dr.alloc(r.size() * 4)) ||
      (rc = dct.alloc((size_t)n * key->ct_words * 4)) || (rc = dout.alloc((size_t)n * key->ct_words * 4)))
    return rc;
  HIP_TRY(hipMemcpy(dm.p, m.data(), n * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dk.p, k.data(), n * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dr.p, r.data(), r.size() * 4, hipMemcpyHostToDevice));
  if ((rc = heu_encrypt_i64(key, dm.as<int64_t>(), dr.as<uint32_t>(), rw, n, dct.as<uint32_t>()))) return rc;
  if ((rc = heu_ctx_sync(ctx))) return rc;
  std::string js = "{\"key_bits\": " + std::to_string(bits);

  if (what & HEU_TUNE_TABLE) {
    const size_t budget = table_budget_bytes();
    uint32_t best_w = (uint32_t)key->unit;
    double best_t = 1e30;
    js += ", \"table_window\": {\"elements\": " + std::to_string(n_enc) + ", \"ms\": {";
    bool first = true;
    for (uint32_t w : {8u, 10u, 12u, 13u, 14u, 15u, 16u}) {
      const size_t bytes = ((size_t)((key->r_bits + w - 1) / w) << w) * (size_t)key->nl_n * 4;
      if (bytes > budget) continue;
      if ((rc = heu_key_set_param(key, HEU_PARAM_TABLE_WINDOW, w))) return rc;
      double t;
      rc = best_ms(ctx, [&] {
        return heu_encrypt_i64(key, dm.as<int64_t>(), dr.as<uint32_t>(), rw, n_enc, dout.as<uint32_t>());
      }, &t);
      if (rc) return rc;
      js += std::string(first ? "" : ", ") + "\"" + std::to_string(w) + "\": " + fmt("%.3f", t);
      first = false;
      if (t < best_t) {
        best_t = t;
        best_w = w;
      }
    }
    if ((rc = heu_key_set_param(key, HEU_PARAM_TABLE_WINDOW, best_w))) return rc;
    js += "}, \"chosen\": " + std::to_string(best_w) + "}";
  }

  if ((what & HEU_TUNE_DEC_LANES) && key->has_sk) {
    int best_v = key->dec_shape;
    double best_t = 1e30;
    js += ", \"dec_lanes\": {\"elements\": " + std::to_string(n_dec) + ", \"ms\": {";
    for (int v = 0; v < heu_hip::kDecShapes; ++v) {
      const uint32_t lanes = (uint32_t)key->ops->dec[v].L;
      if ((rc = heu_key_set_param(key, HEU_PARAM_DEC_LANES, lanes))) return rc;
      double t;
      rc = best_ms(ctx, [&] { return heu_decrypt(key, dct.as<uint32_t>(), n_dec, dout.as<uint32_t>(), 2, nullptr); },
                   &t);
      if (rc) return rc;
      js += std::string(v ? ", " : "") + "\"" + std::to_string(lanes) + "\": " + fmt("%.3f", t);
      // the first shape (fewest lanes) stays unless another is clearly faster:
      // timings within 2% are noise, and a stable choice keeps runs comparable
      if (t < best_t * (v ? 0.98 : 1.0)) {
        best_t = t;
        best_v = v;
      }
    }
    if ((rc = heu_key_set_param(key, HEU_PARAM_DEC_LANES, (uint32_t)key->ops->dec[best_v].L))) return rc;
    js += "}, \"chosen\": " + std::to_string(key->ops->dec[best_v].L) + "}";
  }

  if (what & HEU_TUNE_GRID) {
    js += ", \"grid_multiplier\": {";
    for (int dec = 0; dec < 2; ++dec) {
      if (dec && !key->has_sk) continue;
      heu_task t{};
      t.key = key;
      t.a = dct.p;
      t.out = dout.p;
      if (dec) {
        t.op = HEU_OP_DECRYPT;
        t.aux = 2;
        t.count = (size_t)n_dec;
      } else {
        t.op = HEU_OP_MUL_I64;
        t.b = dk.p;
        t.count = (size_t)n_enc;
      }
      uint32_t best_m = 1;
      double best_t = 1e30;
      js += std::string(dec ? ", " : "") + (dec ? "\"decrypt\"" : "\"mul\"") + ": {\"ms\": {";
      for (uint32_t mult : {1u, 2u, 4u}) {
        if ((rc = heu_ctx_set_grid_multiplier(ctx, (uint32_t)bits, dec, mult))) return rc;
        heu_batch* b = nullptr;
        if ((rc = heu_batch_create(ctx, &t, 1, 0, &b))) return rc;
        double tm;
        rc = best_ms(ctx, [&] { return heu_batch_run(b); }, &tm);
        heu_batch_destroy(b);
        if (rc) return rc;
        js += std::string(mult == 1 ? "" : ", ") + "\"" + std::to_string(mult) + "\": " + fmt("%.3f", tm);
        // a wider grid only queues more workgroups: keep it only for a 2% gain
        if (tm < best_t * (mult == 1 ? 1.0 : 0.98)) {
          best_t = tm;
          best_m = mult;
        }
      }
      if ((rc = heu_ctx_set_grid_multiplier(ctx, (uint32_t)bits, dec, best_m))) return rc;
      js += "}, \"chosen\": " + std::to_string(best_m) + "}";
    }
    js += "}";
  }
  js += "}";
  if (report && report_len) {
    const size_t c = std::min(report_len - 1, js.size());
    std::memcpy(report, js.data(), c);
    report[c] = '\0';
  }
  return HEU_OK;
}
