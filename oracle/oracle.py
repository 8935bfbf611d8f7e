"""ctypes front end of the CPU oracle (paillier_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() to
check its result and by bench.py's cpu_baseline leg.  The product package
(heu_amd) never imports this module.

Big integers travel as numpy uint32 arrays of little-endian limbs; signed
plaintexts as two's complement over a fixed number of limbs.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_SRC = os.path.join(_HERE, "paillier_oracle.c")

ORC_ERRORS = {1: "invalid argument", 2: "out of range", 3: "not invertible", 4: "out of memory"}


class OracleError(RuntimeError):
    pass


def build(force: bool = False) -> str:
    """Compile paillier_oracle.c into oracle/liboracle.so (gcc, OpenSSL libcrypto)."""
    if force or not os.path.exists(_LIB_PATH) or os.path.getmtime(_LIB_PATH) < os.path.getmtime(_SRC):
        tmp = _LIB_PATH + ".tmp"
        subprocess.check_call(["gcc", "-O2", "-fPIC", "-shared", "-Wall", "-o", tmp, _SRC, "-lcrypto", "-lpthread"])
        os.replace(tmp, _LIB_PATH)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P, S, I, U64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint64
        L.orc_prng_word.restype = U64
        L.orc_prng_word.argtypes = [U64, U64, U64]
        L.orc_prng_fill.argtypes = [U64, U64, U64, P, S]
        L.orc_keygen.argtypes = [I, U64, P, P, P, P]
        L.orc_key_new.restype = P
        L.orc_key_new.argtypes = [P, P, I, P, P, I]
        L.orc_key_free.argtypes = [P]
        L.orc_key_build_table.argtypes = [P, I, I]
        L.orc_key_export.argtypes = [P, ctypes.c_char_p, P, I]
        L.orc_encrypt_i64.argtypes = [P, P, P, I, S, P, I, I]
        L.orc_encrypt_twos.argtypes = [P, P, I, P, I, S, P, I, I]
        L.orc_decrypt.argtypes = [P, P, S, P, I, I]
        L.orc_add.argtypes = [P, P, P, S, P, I]
        L.orc_add_pt_i64.argtypes = [P, P, P, S, P, I]
        L.orc_sub.argtypes = [P, P, P, S, P, I]
        L.orc_negate.argtypes = [P, P, S, P, I]
        L.orc_mul_i64.argtypes = [P, P, P, S, P, I]
        L.orc_mul_twos.argtypes = [P, P, P, I, S, P, I]
        L.orc_randomize.argtypes = [P, P, P, I, S, P, I, I]
        L.orc_dot_i64.argtypes = [P, P, P, P, S, P]
        L.orc_to_mont.argtypes = [P, P, S, I, I, P]
        _lib = L
    return _lib


def _check(rc):
    if rc:
        raise OracleError(ORC_ERRORS.get(rc, f"error {rc}"))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------- ints/limbs
def int_to_limbs(v: int, words: int) -> np.ndarray:
    if v < 0:
        raise ValueError("negative value")
    return np.frombuffer(v.to_bytes(words * 4, "little"), dtype=np.uint32).copy()


def limbs_to_int(a) -> int:
    return int.from_bytes(np.ascontiguousarray(a, dtype=np.uint32).tobytes(), "little")


def twos_to_int(a) -> int:
    a = np.ascontiguousarray(a, dtype=np.uint32)
    v = limbs_to_int(a)
    bits = 32 * a.shape[-1]
    return v - (1 << bits) if v >> (bits - 1) else v


def int_to_twos(v: int, words: int) -> np.ndarray:
    return int_to_limbs(v % (1 << (32 * words)), words)


def ints_to_rows(values, words: int) -> np.ndarray:
    out = np.zeros((len(values), words), dtype=np.uint32)
    for i, v in enumerate(values):
        out[i] = int_to_limbs(v, words)
    return out


def rows_to_ints(rows):
    return [limbs_to_int(r) for r in np.asarray(rows)]


# ---------------------------------------------------------------- PRNG
def prng_words(seed: int, stream: int, count: int, offset: int = 0) -> np.ndarray:
    out = np.empty(count, dtype=np.uint64)
    lib().orc_prng_fill(seed, stream, offset, _ptr(out), count)
    return out


# ---------------------------------------------------------------- keys
def keygen(key_bits: int, seed: int) -> dict:
    """Seeded restatement of paillier_zahlen/key_generator.cc:26-60."""
    pw, nw = key_bits // 64, key_bits // 32
    p, q = np.zeros(pw, np.uint32), np.zeros(pw, np.uint32)
    n, hs = np.zeros(nw, np.uint32), np.zeros(2 * nw, np.uint32)
    _check(lib().orc_keygen(key_bits, seed, _ptr(p), _ptr(q), _ptr(n), _ptr(hs)))
    return {"p": limbs_to_int(p), "q": limbs_to_int(q), "n": limbs_to_int(n), "hs": limbs_to_int(hs), "bits": key_bits}


class Key:
    """A paillier_z PublicKey (+ optional SecretKey) held by the oracle."""

    def __init__(self, n: int, hs: int, p: int | None = None, q: int | None = None):
        self.n, self.hs, self.p, self.q = n, hs, p, q
        self.n_words = (n.bit_length() + 31) // 32
        self.ct_words = 2 * self.n_words
        self.prime_words = self.n_words // 2 + (self.n_words % 2)
        nl = int_to_limbs(n, self.n_words)
        hl = int_to_limbs(hs, self.ct_words)
        if p is not None:
            pl = int_to_limbs(p, self.prime_words)
            ql = int_to_limbs(q, self.prime_words)
            self._h = lib().orc_key_new(_ptr(nl), _ptr(hl), self.n_words, _ptr(pl), _ptr(ql), self.prime_words)
        else:
            self._h = lib().orc_key_new(_ptr(nl), _ptr(hl), self.n_words, None, None, 0)
        self.table_unit = None

    @classmethod
    def from_keygen(cls, kd: dict) -> "Key":
        return cls(kd["n"], kd["hs"], kd["p"], kd["q"])

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.orc_key_free(h)
            self._h = None

    def build_table(self, unit_bits: int = 10, max_exp_bits: int = 256):
        _check(lib().orc_key_build_table(self._h, unit_bits, max_exp_bits))
        self.table_unit = unit_bits

    def export(self, name: str) -> int:
        words = self.ct_words
        buf = np.zeros(words, np.uint32)
        _check(lib().orc_key_export(self._h, name.encode(), _ptr(buf), words))
        return limbs_to_int(buf)

    # ---------------------------------------------------------- batch ops
    def _ct(self, count):
        return np.zeros((count, self.ct_words), np.uint32)

    def encrypt_i64(self, m, r, use_table=False, threads=1):
        m = np.ascontiguousarray(m, dtype=np.int64)
        r = np.ascontiguousarray(r, dtype=np.uint32)
        out = self._ct(len(m))
        _check(lib().orc_encrypt_i64(self._h, _ptr(m), _ptr(r), r.shape[1], len(m), _ptr(out), int(use_table), threads))
        return out

    def encrypt_twos(self, m, r, use_table=False, threads=1):
        m = np.ascontiguousarray(m, dtype=np.uint32)
        r = np.ascontiguousarray(r, dtype=np.uint32)
        out = self._ct(len(m))
        _check(lib().orc_encrypt_twos(self._h, _ptr(m), m.shape[1], _ptr(r), r.shape[1], len(m), _ptr(out),
                                      int(use_table), threads))
        return out

    def decrypt(self, ct, out_words=None, threads=1):
        ct = np.ascontiguousarray(ct, dtype=np.uint32)
        ow = out_words or self.n_words
        out = np.zeros((len(ct), ow), np.uint32)
        _check(lib().orc_decrypt(self._h, _ptr(ct), len(ct), _ptr(out), ow, threads))
        return out

    def add(self, a, b, threads=1):
        a, b = np.ascontiguousarray(a, np.uint32), np.ascontiguousarray(b, np.uint32)
        out = self._ct(len(a))
        _check(lib().orc_add(self._h, _ptr(a), _ptr(b), len(a), _ptr(out), threads))
        return out

    def add_pt_i64(self, a, p, threads=1):
        a, p = np.ascontiguousarray(a, np.uint32), np.ascontiguousarray(p, np.int64)
        out = self._ct(len(a))
        _check(lib().orc_add_pt_i64(self._h, _ptr(a), _ptr(p), len(a), _ptr(out), threads))
        return out

    def sub(self, a, b, threads=1):
        a, b = np.ascontiguousarray(a, np.uint32), np.ascontiguousarray(b, np.uint32)
        out = self._ct(len(a))
        _check(lib().orc_sub(self._h, _ptr(a), _ptr(b), len(a), _ptr(out), threads))
        return out

    def negate(self, a, threads=1):
        a = np.ascontiguousarray(a, np.uint32)
        out = self._ct(len(a))
        _check(lib().orc_negate(self._h, _ptr(a), len(a), _ptr(out), threads))
        return out

    def mul_i64(self, a, s, threads=1):
        a, s = np.ascontiguousarray(a, np.uint32), np.ascontiguousarray(s, np.int64)
        out = self._ct(len(a))
        _check(lib().orc_mul_i64(self._h, _ptr(a), _ptr(s), len(a), _ptr(out), threads))
        return out

    def mul(self, a, scalars, threads=1):
        """Mul by Python ints of any size (two's complement rows)."""
        a = np.ascontiguousarray(a, np.uint32)
        words = max(2, max((abs(int(v)).bit_length() + 32) // 32 for v in scalars))
        s = np.stack([int_to_twos(int(v), words) for v in scalars]) if len(scalars) else np.zeros((0, words), np.uint32)
        out = self._ct(len(a))
        _check(lib().orc_mul_twos(self._h, _ptr(a), _ptr(s), words, len(a), _ptr(out), threads))
        return out

    def randomize(self, a, r, use_table=False, threads=1):
        a, r = np.ascontiguousarray(a, np.uint32), np.ascontiguousarray(r, np.uint32)
        out = self._ct(len(a))
        _check(lib().orc_randomize(self._h, _ptr(a), _ptr(r), r.shape[1], len(a), _ptr(out), int(use_table), threads))
        return out

    def dot_i64(self, cts, weights, offsets):
        cts = np.ascontiguousarray(cts, np.uint32)
        offsets = np.ascontiguousarray(offsets, np.int64)
        w = None if weights is None else np.ascontiguousarray(weights, np.int64)
        segs = len(offsets) - 1
        out = self._ct(segs)
        _check(lib().orc_dot_i64(self._h, _ptr(cts), _ptr(w), _ptr(offsets), segs, _ptr(out)))
        return out

    def to_mont(self, cts, word_bits=64, inverse=False):
        cts = np.ascontiguousarray(cts, np.uint32)
        out = self._ct(len(cts))
        _check(lib().orc_to_mont(self._h, _ptr(cts), len(cts), word_bits, 0 if inverse else 1, _ptr(out)))
        return out
