// wire.cpp -- HEU's serialized form of ZPaillier objects (heu_wire_* in
// include/heu_paillier.h).  Host code only: ciphertexts arrive here in the
// Montgomery form heu_to_mont writes on the device.
//
// The layout follows the reference's writers:
//   - paillier_z::Ciphertext is MSGPACK_DEFINE(c_) (paillier_zahlen/
//     ciphertext.h:37): a msgpack array holding the BigInt c_;
//   - PublicKey packs [n, h_s] (public_key.h:64-76), SecretKey
//     [lambda, mu, p, q] (secret_key.h:64-80);
//   - phe's SerializableVariant appends the schema index as a little-endian
//     size_t (phe/base/serializable_types.cc:50-70), 4 for ZPaillier
//     (phe/base/schema.h:49);
//   - DenseMatrix packs [rows, cols, ndim, [items]] with every item's bytes as
//     a msgpack str, items in Eigen's column-major order (numpy/matrix.h:221-265).
// A BigInt is packed by yacl's msgpack adaptor, which the reference does not
// vendor; here it is a msgpack str holding one sign byte (0 non-negative,
// 1 negative) and the magnitude's bytes, little-endian, without high zero
// bytes.  Integers use msgpack's shortest encoding, as msgpack-c writes them.
#include <openssl/bn.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/heu_paillier.h"
#include "engine_internal.hpp"

using heu_engine::fail;

namespace {

// Writes into [p, p + cap); with p == nullptr it only counts.  `n` is the
// length so far whether or not it fit.
struct Out {
  uint8_t* p;
  size_t cap, n = 0;
  void byte(uint8_t b) {
    if (p && n < cap) p[n] = b;
    ++n;
  }
  void bytes(const uint8_t* b, size_t len) {
    if (p && n + len <= cap) std::memcpy(p + n, b, len);
    n += len;
  }
  void be(uint64_t v, int width) {
    for (int i = width - 1; i >= 0; --i) byte((uint8_t)(v >> (8 * i)));
  }
  void array(uint32_t count) {
    if (count < 16) byte((uint8_t)(0x90 | count));
    else if (count < 65536) { byte(0xdc); be(count, 2); }
    else { byte(0xdd); be(count, 4); }
  }
  void str_header(size_t len) {
    if (len < 32) byte((uint8_t)(0xa0 | len));
    else if (len < 256) { byte(0xd9); be(len, 1); }
    else if (len < 65536) { byte(0xda); be(len, 2); }
    else { byte(0xdb); be(len, 4); }
  }
  void integer(int64_t v) {
    if (v >= 0) {
      const uint64_t u = (uint64_t)v;
      if (u < 128) byte((uint8_t)u);
      else if (u < 256) { byte(0xcc); be(u, 1); }
      else if (u < 65536) { byte(0xcd); be(u, 2); }
      else if (u < (1ull << 32)) { byte(0xce); be(u, 4); }
      else { byte(0xcf); be(u, 8); }
    } else if (v >= -32) {
      byte((uint8_t)(0xe0 | (v & 0x1f)));
    } else if (v >= -128) {
      byte(0xd0); be((uint64_t)v, 1);
    } else if (v >= -32768) {
      byte(0xd1); be((uint64_t)v, 2);
    } else if (v >= INT32_MIN) {
      byte(0xd2); be((uint64_t)v, 4);
    } else {
      byte(0xd3); be((uint64_t)v, 8);
    }
  }
  // A non-negative BigInt given as `words` little-endian uint32 limbs.
  void bigint_words(const uint32_t* w, size_t words) {
    size_t nb = words * 4;
    const uint8_t* b = reinterpret_cast<const uint8_t*>(w);  // little-endian host
    while (nb > 0 && b[nb - 1] == 0) --nb;
    str_header(nb + 1);
    byte(0);
    bytes(b, nb);
  }
  void bigint(const BIGNUM* v) {
    std::vector<uint8_t> mag((size_t)BN_num_bytes(v));
    if (!mag.empty()) BN_bn2lebinpad(v, mag.data(), (int)mag.size());
    str_header(mag.size() + 1);
    byte(BN_is_negative(v) ? 1 : 0);
    bytes(mag.data(), mag.size());
  }
  void schema_index() {
    const uint64_t idx = HEU_SCHEMA_ZPAILLIER;
    for (int i = 0; i < 8; ++i) byte((uint8_t)(idx >> (8 * i)));
  }
};

int finish(const Out& o, size_t* len) {
  if (len) *len = o.n;
  if (!o.p || o.n > o.cap) return fail(HEU_ERANGE, "output buffer holds %zu bytes; %zu are needed", o.cap, o.n);
  return HEU_OK;
}

// Reads the msgpack subset written above (and bin / every integer form).
struct In {
  const uint8_t* p;
  size_t len, at = 0;
  bool ok = true;
  uint8_t byte() {
    if (at >= len) {
      ok = false;
      return 0;
    }
    return p[at++];
  }
  uint64_t be(int width) {
    uint64_t v = 0;
    for (int i = 0; i < width; ++i) v = (v << 8) | byte();
    return v;
  }
  bool array(uint32_t* count) {
    const uint8_t t = byte();
    if ((t & 0xf0) == 0x90) *count = t & 0x0f;
    else if (t == 0xdc) *count = (uint32_t)be(2);
    else if (t == 0xdd) *count = (uint32_t)be(4);
    else return ok = false;
    return ok;
  }
  // str or bin: its body in [*body, *body + *n)
  bool blob(const uint8_t** body, size_t* n) {
    const uint8_t t = byte();
    if ((t & 0xe0) == 0xa0) *n = t & 0x1f;
    else if (t == 0xd9 || t == 0xc4) *n = (size_t)be(1);
    else if (t == 0xda || t == 0xc5) *n = (size_t)be(2);
    else if (t == 0xdb || t == 0xc6) *n = (size_t)be(4);
    else return ok = false;
    if (!ok || *n > len - at) return ok = false;
    *body = p + at;
    at += *n;
    return true;
  }
  bool integer(int64_t* v) {
    const uint8_t t = byte();
    if (t < 0x80) *v = t;
    else if (t >= 0xe0) *v = (int8_t)t;
    else if (t == 0xcc) *v = (int64_t)be(1);
    else if (t == 0xcd) *v = (int64_t)be(2);
    else if (t == 0xce) *v = (int64_t)be(4);
    else if (t == 0xcf) {
      const uint64_t u = be(8);
      if (u >> 63) return ok = false;
      *v = (int64_t)u;
    } else if (t == 0xd0) *v = (int8_t)be(1);
    else if (t == 0xd1) *v = (int16_t)be(2);
    else if (t == 0xd2) *v = (int32_t)be(4);
    else if (t == 0xd3) *v = (int64_t)be(8);
    else return ok = false;
    return ok;
  }
  // A BigInt into BIGNUM r.
  bool bigint(BIGNUM* r) {
    const uint8_t* b;
    size_t n;
    if (!blob(&b, &n) || n < 1 || b[0] > 1 || n - 1 > (size_t)INT_MAX) return ok = false;
    if (!BN_lebin2bn(b + 1, (int)(n - 1), r)) return ok = false;
    BN_set_negative(r, b[0] == 1 && !BN_is_zero(r));
    return true;
  }
  // A non-negative BigInt into `words` limbs (fails if it does not fit).
  bool bigint_words(uint32_t* w, size_t words) {
    const uint8_t* b;
    size_t n;
    if (!blob(&b, &n) || n < 1 || b[0] != 0) return ok = false;
    // b[1..n) is the magnitude: bytes beyond `words` limbs must be zero
    if (n - 1 > words * 4) {
      for (size_t i = words * 4; i < n - 1; ++i)
        if (b[1 + i]) return ok = false;
    }
    std::memset(w, 0, words * 4);
    std::memcpy(w, b + 1, std::min(n - 1, words * 4));
    return true;
  }
  bool schema_index_at_end(size_t* payload_len) {
    if (len < 9) return false;
    uint64_t idx = 0;
    for (int i = 7; i >= 0; --i) idx = (idx << 8) | p[len - 8 + i];
    if (idx != HEU_SCHEMA_ZPAILLIER) return false;
    *payload_len = len - 8;
    return true;
  }
};

struct Bn {
  BIGNUM* b = BN_new();
  Bn() = default;
  Bn(const Bn&) = delete;
  ~Bn() { BN_free(b); }
  operator BIGNUM*() const { return b; }
};

void put_ciphertext(Out& o, const uint32_t* mont, uint32_t words) {
  o.array(1);  // MSGPACK_DEFINE(c_)
  o.bigint_words(mont, words);
  o.schema_index();
}

size_t ciphertext_len(const uint32_t* mont, uint32_t words) {
  Out o{nullptr, 0};
  put_ciphertext(o, mont, words);
  return o.n;
}

bool get_ciphertext(const uint8_t* b, size_t n, uint32_t* mont, uint32_t words) {
  In in{b, n};
  size_t plen;
  if (!in.schema_index_at_end(&plen)) return false;
  in.len = plen;
  uint32_t cnt = 0;
  return in.array(&cnt) && cnt == 1 && in.bigint_words(mont, words) && in.at == in.len;
}

int to_words(const BIGNUM* v, uint32_t* w, uint32_t words) {
  if (BN_is_negative(v) || BN_num_bytes(v) > (int)(4 * words))
    return fail(HEU_ERANGE, "a key value does not fit %u words", words);
  BN_bn2lebinpad(v, reinterpret_cast<unsigned char*>(w), (int)(4 * words));
  return HEU_OK;
}

uint32_t words_of(const BIGNUM* v) { return (uint32_t)((BN_num_bits(v) + 31) / 32); }

// rows x cols == count, without the product wrapping around
bool shape_holds(int64_t rows, int64_t cols, uint64_t count) {
  if (rows < 0 || cols < 0) return false;
  if (cols != 0 && (uint64_t)rows > UINT64_MAX / (uint64_t)cols) return false;
  return (uint64_t)rows * (uint64_t)cols == count;
}

// ---- interconnection form (numpy/matrix.cc:39-111) -------------------------
// A protobuf DataExchangeProtocol (package org.interconnection.v2.runtime).
// Its .proto is part of org_interconnection 0.0.1 (MODULE.bazel:45), which the
// reference does not vendor, so these numbers are this engine's reading of it,
// held here alone:
//   message DataExchangeProtocol { ScalarType scalar_type = 1;
//     string scalar_type_name = 2; oneof container { NdArray v_ndarray = 4; } }
//   message NdArray { repeated int64 shape = 1; repeated bytes items = 3; }
//   SCALAR_TYPE_OBJECT = 12
// Fields are written in field-number order, shape packed, as protobuf's own
// serializers do; the reader takes either shape encoding, any field order and
// skips fields it does not know.
constexpr uint32_t kIcScalarType = 1, kIcScalarTypeName = 2, kIcNdarray = 4;
constexpr uint32_t kIcShape = 1, kIcItems = 3;
constexpr uint64_t kIcScalarObject = 12;
constexpr const char kIcTypeName[] = "ciphertext";  // matrix.cc:29-30 (Typename<phe::Ciphertext>)

enum : uint32_t { kVarint = 0, kFixed64 = 1, kDelimited = 2, kFixed32 = 5 };

void pb_varint(Out& o, uint64_t v) {
  while (v >= 0x80) {
    o.byte((uint8_t)(v | 0x80));
    v >>= 7;
  }
  o.byte((uint8_t)v);
}

void pb_tag(Out& o, uint32_t field, uint32_t wire) { pb_varint(o, ((uint64_t)field << 3) | wire); }

size_t varint_len(uint64_t v) {
  size_t n = 1;
  while (v >= 0x80) {
    v >>= 7;
    ++n;
  }
  return n;
}

// Reads protobuf's wire format: varints, tags and length-delimited bodies.
struct Pb {
  const uint8_t* p;
  size_t len, at = 0;
  bool varint(uint64_t* v) {
    *v = 0;
    for (int shift = 0; shift < 70; shift += 7) {
      if (at >= len) return false;
      const uint8_t b = p[at++];
      if (shift == 63 && b > 1) return false;  // beyond 64 bits
      *v |= (uint64_t)(b & 0x7f) << shift;
      if (!(b & 0x80)) return true;
    }
    return false;
  }
  bool tag(uint32_t* field, uint32_t* wire) {
    uint64_t t;
    if (!varint(&t) || (t >> 3) == 0 || (t >> 3) > 0x1fffffff) return false;
    *field = (uint32_t)(t >> 3);
    *wire = (uint32_t)(t & 7);
    return true;
  }
  bool delimited(const uint8_t** body, size_t* n) {
    uint64_t l;
    if (!varint(&l) || l > len - at) return false;
    *body = p + at;
    *n = (size_t)l;
    at += (size_t)l;
    return true;
  }
  bool skip(uint32_t wire) {
    uint64_t v;
    const uint8_t* b;
    size_t n;
    switch (wire) {
      case kVarint: return varint(&v);
      case kFixed64: return len - at >= 8 ? (at += 8, true) : false;
      case kDelimited: return delimited(&b, &n);
      case kFixed32: return len - at >= 4 ? (at += 4, true) : false;
      default: return false;  // groups are not used by proto3 messages
    }
  }
};

struct IcMatrix {
  std::vector<int64_t> shape;
  std::vector<std::pair<const uint8_t*, size_t>> items;
};

// NdArray fields of one v_ndarray body, merged into m (a repeated message
// field's occurrences merge, so shapes and items append).
bool ic_ndarray(const uint8_t* b, size_t n, IcMatrix* m) {
  Pb r{b, n};
  while (r.at < r.len) {
    uint32_t f, w;
    if (!r.tag(&f, &w)) return false;
    if (f == kIcShape && w == kDelimited) {
      const uint8_t* body;
      size_t bn;
      if (!r.delimited(&body, &bn)) return false;
      Pb s{body, bn};
      while (s.at < s.len) {
        uint64_t v;
        if (!s.varint(&v)) return false;
        m->shape.push_back((int64_t)v);
      }
    } else if (f == kIcShape && w == kVarint) {
      uint64_t v;
      if (!r.varint(&v)) return false;
      m->shape.push_back((int64_t)v);
    } else if (f == kIcItems && w == kDelimited) {
      const uint8_t* body;
      size_t bn;
      if (!r.delimited(&body, &bn)) return false;
      m->items.emplace_back(body, bn);
    } else if (f == kIcShape || f == kIcItems || !r.skip(w)) {
      return false;  // a known field with the wrong wire type, or a bad field
    }
  }
  return true;
}

}  // namespace

extern "C" {

int heu_wire_put_ciphertext(const uint32_t* mont, uint32_t words, uint8_t* out, size_t cap, size_t* len) {
  if (!mont || words == 0 || !len) return fail(HEU_EINVAL, "null argument");
  Out o{out, cap};
  put_ciphertext(o, mont, words);
  return finish(o, len);
}

int heu_wire_get_ciphertext(const uint8_t* in, size_t len, uint32_t* mont, uint32_t words) {
  if (!in || !mont || words == 0) return fail(HEU_EINVAL, "null argument");
  if (!get_ciphertext(in, len, mont, words))
    return fail(HEU_EINVAL, "not a serialized ZPaillier ciphertext of at most %u words", words);
  return HEU_OK;
}

int heu_wire_put_matrix(const uint32_t* mont, int64_t rows, int64_t cols, int64_t ndim, uint32_t words,
                        uint8_t* out, size_t cap, size_t* len) {
  if (!len || words == 0 || rows < 0 || cols < 0 || ndim < 0 || ndim > 2) return fail(HEU_EINVAL, "bad argument");
  const int64_t count = rows * cols;
  if (count > 0 && !mont) return fail(HEU_EINVAL, "null argument");
  if (count > (int64_t)UINT32_MAX) return fail(HEU_EINVAL, "too many elements for a msgpack array");
  Out o{out, cap};
  o.array(4);
  o.integer(rows);
  o.integer(cols);
  o.integer(ndim);
  o.array((uint32_t)count);
  for (int64_t i = 0; i < count; ++i) {
    const uint32_t* v = mont + (size_t)i * words;
    o.str_header(ciphertext_len(v, words));
    put_ciphertext(o, v, words);
  }
  return finish(o, len);
}

int heu_wire_get_matrix(const uint8_t* in, size_t len, int64_t* rows, int64_t* cols, int64_t* ndim, uint32_t* mont,
                        uint32_t words, size_t cap_values) {
  if (!in) return fail(HEU_EINVAL, "null argument");
  In r{in, len};
  uint32_t four = 0, count = 0;
  int64_t nr = 0, nc = 0, nd = 0;
  if (!r.array(&four) || four != 4 || !r.integer(&nr) || !r.integer(&nc) || !r.integer(&nd) || !r.array(&count) ||
      nd < 0 || nd > 2 || !shape_holds(nr, nc, count))
    return fail(HEU_EINVAL, "not a serialized ciphertext matrix: [rows, cols, ndim, items] expected");
  if (rows) *rows = nr;
  if (cols) *cols = nc;
  if (ndim) *ndim = nd;
  if (!mont) return HEU_OK;
  if (words == 0) return fail(HEU_EINVAL, "words must be positive");
  if (cap_values < count) return fail(HEU_ERANGE, "the matrix holds %u values; room for %zu", count, cap_values);
  for (uint32_t i = 0; i < count; ++i) {
    const uint8_t* b;
    size_t n;
    if (!r.blob(&b, &n) || !get_ciphertext(b, n, mont + (size_t)i * words, words))
      return fail(HEU_EINVAL, "matrix item %u is not a serialized ZPaillier ciphertext of at most %u words", i, words);
  }
  if (r.at != len) return fail(HEU_EINVAL, "trailing bytes after the matrix");
  return HEU_OK;
}

int heu_wire_put_matrix_ic(const uint32_t* mont, int64_t rows, int64_t cols, int64_t ndim, uint32_t words,
                           uint8_t* out, size_t cap, size_t* len) {
  if (!len || words == 0 || rows < 0 || cols < 0 || ndim < 0 || ndim > 2) return fail(HEU_EINVAL, "bad argument");
  if ((ndim == 0 && (rows != 1 || cols != 1)) || (ndim == 1 && cols != 1))
    return fail(HEU_EINVAL, "a %lld-d matrix cannot be %lld x %lld", (long long)ndim, (long long)rows,
                (long long)cols);
  const int64_t count = rows * cols;
  if (count > 0 && !mont) return fail(HEU_EINVAL, "null argument");
  // the NdArray body's length comes first, so count it
  const int64_t shape[2] = {rows, cols};
  size_t shape_len = 0;
  for (int64_t i = 0; i < ndim; ++i) shape_len += varint_len((uint64_t)shape[i]);
  size_t nd_len = ndim ? 1 + varint_len(shape_len) + shape_len : 0;
  for (int64_t i = 0; i < count; ++i) {
    const size_t item = ciphertext_len(mont + (size_t)i * words, words);
    nd_len += 1 + varint_len(item) + item;
  }
  Out o{out, cap};
  pb_tag(o, kIcScalarType, kVarint);
  pb_varint(o, kIcScalarObject);
  pb_tag(o, kIcScalarTypeName, kDelimited);
  pb_varint(o, sizeof(kIcTypeName) - 1);
  o.bytes(reinterpret_cast<const uint8_t*>(kIcTypeName), sizeof(kIcTypeName) - 1);
  pb_tag(o, kIcNdarray, kDelimited);
  pb_varint(o, nd_len);
  const size_t body_at = o.n;
  if (ndim) {
    pb_tag(o, kIcShape, kDelimited);
    pb_varint(o, shape_len);
    for (int64_t i = 0; i < ndim; ++i) pb_varint(o, (uint64_t)shape[i]);
  }
  for (int64_t i = 0; i < count; ++i) {
    const uint32_t* v = mont + (size_t)i * words;
    pb_tag(o, kIcItems, kDelimited);
    pb_varint(o, ciphertext_len(v, words));
    put_ciphertext(o, v, words);
  }
  if (o.n - body_at != nd_len) return fail(HEU_EINVAL, "internal: NdArray length mismatch");
  return finish(o, len);
}

int heu_wire_get_matrix_ic(const uint8_t* in, size_t len, int64_t* rows, int64_t* cols, int64_t* ndim, uint32_t* mont,
                           uint32_t words, size_t cap_values) {
  if (!in && len) return fail(HEU_EINVAL, "null argument");
  Pb r{in, len};
  uint64_t scalar_type = 0;
  bool have_nd = false;
  IcMatrix m;
  while (r.at < r.len) {
    uint32_t f, w;
    if (!r.tag(&f, &w)) return fail(HEU_EINVAL, "not a protobuf DataExchangeProtocol");
    bool ok;
    if (f == kIcScalarType && w == kVarint) {
      ok = r.varint(&scalar_type);
    } else if (f == kIcNdarray && w == kDelimited) {
      const uint8_t* b;
      size_t n;
      ok = r.delimited(&b, &n) && ic_ndarray(b, n, &m);
      have_nd = true;
    } else {
      ok = f != kIcScalarType && f != kIcNdarray && r.skip(w);
    }
    if (!ok) return fail(HEU_EINVAL, "not a protobuf DataExchangeProtocol (field %u)", f);
  }
  if (scalar_type != kIcScalarObject)
    return fail(HEU_EINVAL, "Buffer format illegal, scalar_type=%llu", (unsigned long long)scalar_type);
  if (!have_nd) return fail(HEU_EINVAL, "unsupported container type: an NdArray is expected");
  const size_t nd = m.shape.size();
  const int64_t nr = nd > 0 ? m.shape[0] : 1, nc = nd > 1 ? m.shape[1] : 1;
  if (nd > 2 || nr < 0 || nc < 0) return fail(HEU_EINVAL, "a ciphertext matrix has at most 2 non-negative dims");
  if (!shape_holds(nr, nc, m.items.size()))
    return fail(HEU_EINVAL, "Pb: shape and len not match (%lld x %lld, %zu items)", (long long)nr, (long long)nc,
                m.items.size());
  if (rows) *rows = nr;
  if (cols) *cols = nc;
  if (ndim) *ndim = (int64_t)nd;
  if (!mont) return HEU_OK;
  if (words == 0) return fail(HEU_EINVAL, "words must be positive");
  if (cap_values < m.items.size())
    return fail(HEU_ERANGE, "the matrix holds %zu values; room for %zu", m.items.size(), cap_values);
  for (size_t i = 0; i < m.items.size(); ++i)
    if (!get_ciphertext(m.items[i].first, m.items[i].second, mont + i * words, words))
      return fail(HEU_EINVAL, "matrix item %zu is not a serialized ZPaillier ciphertext of at most %u words", i, words);
  return HEU_OK;
}

int heu_wire_put_public_key(const uint32_t* n, uint32_t n_words, const uint32_t* h_s, uint8_t* out, size_t cap,
                            size_t* len) {
  if (!n || !h_s || n_words == 0 || !len) return fail(HEU_EINVAL, "null argument");
  Out o{out, cap};
  o.array(2);
  o.bigint_words(n, n_words);
  o.bigint_words(h_s, 2 * (size_t)n_words);
  o.schema_index();
  return finish(o, len);
}

int heu_wire_get_public_key(const uint8_t* in, size_t len, uint32_t* n, uint32_t* h_s, uint32_t cap_words,
                            uint32_t* n_words) {
  if (!in || !n_words) return fail(HEU_EINVAL, "null argument");
  In r{in, len};
  size_t plen;
  if (!r.schema_index_at_end(&plen)) return fail(HEU_EINVAL, "not a serialized ZPaillier public key");
  r.len = plen;
  Bn bn_n, bn_h;
  uint32_t two = 0;
  if (!r.array(&two) || two != 2 || !r.bigint(bn_n) || !r.bigint(bn_h) || r.at != r.len || BN_is_negative(bn_n) ||
      BN_is_negative(bn_h))
    return fail(HEU_EINVAL, "not a serialized ZPaillier public key: [n, h_s] expected");
  *n_words = words_of(bn_n);
  if (!n || !h_s) return HEU_OK;
  if (cap_words < *n_words) return fail(HEU_ERANGE, "n has %u words; room for %u", *n_words, cap_words);
  int rc = to_words(bn_n, n, cap_words);
  if (!rc) rc = to_words(bn_h, h_s, 2 * cap_words);
  return rc;
}

int heu_wire_put_secret_key(const uint32_t* p, const uint32_t* q, uint32_t prime_words, uint8_t* out, size_t cap,
                            size_t* len) {
  if (!p || !q || prime_words == 0 || !len) return fail(HEU_EINVAL, "null argument");
  BN_CTX* ctx = BN_CTX_new();
  Bn bp, bq, n, lam, mu, t;
  BN_lebin2bn(reinterpret_cast<const unsigned char*>(p), (int)(4 * prime_words), bp);
  BN_lebin2bn(reinterpret_cast<const unsigned char*>(q), (int)(4 * prime_words), bq);
  BN_mul(n, bp, bq, ctx);
  // key_generator.cc:26-60: lambda = (p - 1)(q - 1) / 2, mu = lambda^-1 mod n
  BN_sub(lam, bp, BN_value_one());
  BN_sub(t, bq, BN_value_one());
  BN_mul(lam, lam, t, ctx);
  BN_rshift1(lam, lam);
  const bool inv = BN_mod_inverse(mu, lam, n, ctx) != nullptr;
  BN_CTX_free(ctx);
  if (!inv) return fail(HEU_EKEY, "lambda has no inverse mod n: p, q are not a Paillier key");
  Out o{out, cap};
  o.array(4);
  o.bigint(lam);
  o.bigint(mu);
  o.bigint(bp);
  o.bigint(bq);
  o.schema_index();
  return finish(o, len);
}

int heu_wire_get_secret_key(const uint8_t* in, size_t len, uint32_t* p, uint32_t* q, uint32_t cap_words,
                            uint32_t* prime_words) {
  if (!in || !prime_words) return fail(HEU_EINVAL, "null argument");
  In r{in, len};
  size_t plen;
  if (!r.schema_index_at_end(&plen)) return fail(HEU_EINVAL, "not a serialized ZPaillier secret key");
  r.len = plen;
  Bn lam, mu, bp, bq;
  uint32_t four = 0;
  if (!r.array(&four) || four != 4 || !r.bigint(lam) || !r.bigint(mu) || !r.bigint(bp) || !r.bigint(bq) ||
      r.at != r.len)
    return fail(HEU_EINVAL, "not a serialized ZPaillier secret key: [lambda, mu, p, q] expected");
  if (BN_is_negative(bp) || BN_is_negative(bq) || BN_cmp(bp, BN_value_one()) <= 0 || BN_cmp(bq, BN_value_one()) <= 0)
    return fail(HEU_EKEY, "p and q must be integers above 1");
  // the packed lambda and mu must be those of p and q
  BN_CTX* ctx = BN_CTX_new();
  Bn n, l2, t, m2;
  BN_mul(n, bp, bq, ctx);
  BN_sub(l2, bp, BN_value_one());
  BN_sub(t, bq, BN_value_one());
  BN_mul(l2, l2, t, ctx);
  BN_rshift1(l2, l2);
  const bool inv = BN_mod_inverse(m2, l2, n, ctx) != nullptr;
  BN_CTX_free(ctx);
  if (!inv || BN_cmp(lam, l2) != 0 || BN_cmp(mu, m2) != 0)
    return fail(HEU_EKEY, "lambda and mu do not belong to p and q");
  *prime_words = std::max(words_of(bp), words_of(bq));
  if (!p || !q) return HEU_OK;
  if (cap_words < *prime_words) return fail(HEU_ERANGE, "the primes have %u words; room for %u", *prime_words, cap_words);
  int rc = to_words(bp, p, cap_words);
  if (!rc) rc = to_words(bq, q, cap_words);
  return rc;
}

}  // extern "C"
