"""HEU's serialized form of ZPaillier objects (heu_wire_* of the C ABI).

Host functions over numpy limb arrays; ciphertext values are the Montgomery
form c R mod n^2 with R = 2^(64 ceil(bits(n^2) / 64)), which is what
paillier_z::Ciphertext::c_ holds under yacl's 64-bit-word backends
(Key.to_mont(ct, 64) on the device).  The layouts are the reference's:
  ciphertext    msgpack [c_] + schema index 4 as a little-endian size_t
                (paillier_zahlen/ciphertext.h:37, phe/base/serializable_types.cc:50-70)
  matrix        msgpack [rows, cols, ndim, [item str ...]], column-major
                (numpy/matrix.h:221-314)
  matrix, ic    protobuf DataExchangeProtocol {scalar_type OBJECT, "ciphertext",
                NdArray {shape, items}} (numpy/matrix.cc:39-111); the .proto
                is not in the reference, the field numbers are wire.cpp's
  public key    msgpack [n, h_s] + index (paillier_zahlen/public_key.h:64-90)
  secret key    msgpack [lambda, mu, p, q] + index (secret_key.h:64-94)
A BigInt inside is a msgpack str of one sign byte and the little-endian
magnitude; yacl's own BigInt packing is not in the reference, so that inner
layout is the engine's.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from ._native import check, lib

SCHEMA_ZPAILLIER = _native.HEU_SCHEMA_ZPAILLIER


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _write(fn, *args) -> bytes:
    """Calls a heu_wire_put_* writer twice: once for the length, once to write."""
    need = ctypes.c_size_t()
    rc = fn(*args, None, 0, ctypes.byref(need))
    if rc not in (_native.HEU_OK, _native.HEU_ERANGE):
        check(rc)
    buf = ctypes.create_string_buffer(max(1, need.value))
    got = ctypes.c_size_t()
    check(fn(*args, buf, need.value, ctypes.byref(got)))
    return buf.raw[: got.value]


def _words(v: int, words: int) -> np.ndarray:
    return np.frombuffer(int(v).to_bytes(4 * words, "little"), dtype=np.uint32).copy()


def _int(a: np.ndarray) -> int:
    return int.from_bytes(np.ascontiguousarray(a, dtype=np.uint32).tobytes(), "little")


def put_ciphertext(mont: np.ndarray) -> bytes:
    """One serialized phe::Ciphertext from its Montgomery-form limbs."""
    m = np.ascontiguousarray(mont, dtype=np.uint32).ravel()
    return _write(lib().heu_wire_put_ciphertext, _ptr(m), m.size)


def get_ciphertext(buf: bytes, words: int) -> np.ndarray:
    out = np.zeros(words, np.uint32)
    check(lib().heu_wire_get_ciphertext(buf, len(buf), _ptr(out), words))
    return out


def put_matrix(mont_colmajor: np.ndarray, rows: int, cols: int, ndim: int) -> bytes:
    """A serialized DenseMatrix<Ciphertext>; values of shape (rows * cols,
    words) in column-major order."""
    m = np.ascontiguousarray(mont_colmajor, dtype=np.uint32)
    if m.ndim != 2 or m.shape[0] != rows * cols or m.shape[1] == 0:
        raise ValueError(f"values must have shape ({rows * cols}, words)")
    return _write(lib().heu_wire_put_matrix, _ptr(m) if m.size else None, rows, cols, ndim, m.shape[1])


def matrix_shape(buf: bytes):
    r, c, d = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    check(lib().heu_wire_get_matrix(buf, len(buf), ctypes.byref(r), ctypes.byref(c), ctypes.byref(d), None, 0, 0))
    return r.value, c.value, d.value


def get_matrix(buf: bytes, words: int):
    """(rows, cols, ndim, values (rows * cols, words) in column-major order)."""
    rows, cols, ndim = matrix_shape(buf)
    out = np.zeros((rows * cols, words), np.uint32)
    check(lib().heu_wire_get_matrix(buf, len(buf), None, None, None, _ptr(out), words, rows * cols))
    return rows, cols, ndim, out


def put_matrix_ic(mont_colmajor: np.ndarray, rows: int, cols: int, ndim: int) -> bytes:
    """The interconnection form of a DenseMatrix<Ciphertext> (Serialize4Ic);
    values as for put_matrix."""
    m = np.ascontiguousarray(mont_colmajor, dtype=np.uint32)
    if m.ndim != 2 or m.shape[0] != rows * cols or m.shape[1] == 0:
        raise ValueError(f"values must have shape ({rows * cols}, words)")
    return _write(lib().heu_wire_put_matrix_ic, _ptr(m) if m.size else None, rows, cols, ndim, m.shape[1])


def matrix_shape_ic(buf: bytes):
    r, c, d = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    check(lib().heu_wire_get_matrix_ic(buf, len(buf), ctypes.byref(r), ctypes.byref(c), ctypes.byref(d), None, 0, 0))
    return r.value, c.value, d.value


def get_matrix_ic(buf: bytes, words: int):
    """LoadFromIc: (rows, cols, ndim, values (rows * cols, words) column-major)."""
    rows, cols, ndim = matrix_shape_ic(buf)
    out = np.zeros((rows * cols, words), np.uint32)
    check(lib().heu_wire_get_matrix_ic(buf, len(buf), None, None, None, _ptr(out), words, rows * cols))
    return rows, cols, ndim, out


def put_public_key(n: int, h_s: int) -> bytes:
    nw = (n.bit_length() + 31) // 32
    return _write(lib().heu_wire_put_public_key, _ptr(_words(n, nw)), nw, _ptr(_words(h_s, 2 * nw)))


def get_public_key(buf: bytes):
    """(n, h_s) as Python integers."""
    nw = ctypes.c_uint32()
    check(lib().heu_wire_get_public_key(buf, len(buf), None, None, 0, ctypes.byref(nw)))
    n, h = np.zeros(nw.value, np.uint32), np.zeros(2 * nw.value, np.uint32)
    check(lib().heu_wire_get_public_key(buf, len(buf), _ptr(n), _ptr(h), nw.value, ctypes.byref(nw)))
    return _int(n), _int(h)


def put_secret_key(p: int, q: int) -> bytes:
    pw = (max(p.bit_length(), q.bit_length()) + 31) // 32
    return _write(lib().heu_wire_put_secret_key, _ptr(_words(p, pw)), _ptr(_words(q, pw)), pw)


def get_secret_key(buf: bytes):
    """(p, q) as Python integers; lambda and mu are checked against them."""
    pw = ctypes.c_uint32()
    check(lib().heu_wire_get_secret_key(buf, len(buf), None, None, 0, ctypes.byref(pw)))
    p, q = np.zeros(pw.value, np.uint32), np.zeros(pw.value, np.uint32)
    check(lib().heu_wire_get_secret_key(buf, len(buf), _ptr(p), _ptr(q), pw.value, ctypes.byref(pw)))
    return _int(p), _int(q)
