"""Python front end of the engine: contexts, resident keys and batch ops on
device tensors.

Device memory and the stream come from PyTorch (plumbing only); every
operation runs in libheu_paillier.so.  Ciphertexts are int32 tensors of shape
(N, ct_words) holding little-endian uint32 limbs of the Z-space value
c = (1 + n*m) * h_s^r mod n^2 (paillier_zahlen/encryptor.cc:42-49).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native
from ._native import check, lib

# ---------------------------------------------------------------------------
# limbs <-> Python ints
# ---------------------------------------------------------------------------


def int_to_words(v: int, words: int) -> np.ndarray:
    return np.frombuffer(int(v).to_bytes(words * 4, "little"), dtype=np.uint32).copy()


def words_to_int(a) -> int:
    return int.from_bytes(np.ascontiguousarray(a, dtype=np.uint32).tobytes(), "little")


def twos_to_int(a) -> int:
    a = np.ascontiguousarray(a, dtype=np.uint32)
    v = words_to_int(a)
    bits = 32 * a.shape[-1]
    return v - (1 << bits) if v >> (bits - 1) else v


def ints_to_words(values, words: int, signed: bool = False) -> np.ndarray:
    mod = 1 << (32 * words)
    out = np.empty((len(values), words), dtype=np.uint32)
    for i, v in enumerate(values):
        v = int(v)
        if signed:
            v %= mod
        elif v < 0 or v >= mod:
            raise ValueError("value does not fit")
        out[i] = np.frombuffer(v.to_bytes(words * 4, "little"), dtype=np.uint32)
    return out


def words_to_ints(rows, signed: bool = False):
    rows = np.ascontiguousarray(rows, dtype=np.uint32)
    f = twos_to_int if signed else words_to_int
    return [f(r) for r in rows]


# ---------------------------------------------------------------------------
# Seeded counter PRNG (same definition as the engine's key generator):
#   word(seed, stream, i) = mix(mix(seed*G + mix(stream + C)) + (i+1)*G)
# ---------------------------------------------------------------------------
_G = np.uint64(0x9E3779B97F4A7C15)
_C = np.uint64(0x632BE59BD9B4E019)


def _mix64(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def prng_words(seed: int, stream: int, count: int, offset: int = 0) -> np.ndarray:
    """`count` 64-bit words of stream `stream` under `seed`, from index `offset`."""
    with np.errstate(over="ignore"):
        key = _mix64(np.uint64(seed) * _G + _mix64(np.uint64(stream) + _C))
        idx = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        return _mix64(key + idx * _G)


def seeded_r(seed: int, stream: int, count: int, r_bits: int = 256, offset: int = 0) -> np.ndarray:
    """Randomness r uniform in [1, 2^r_bits) as (count, ceil(r_bits/32)) uint32 limbs
    (r = 0 is replaced by 1)."""
    w64 = (r_bits + 63) // 64
    raw = prng_words(seed, stream, count * w64, offset * w64).reshape(count, w64)
    if r_bits % 64:
        raw[:, -1] &= np.uint64((1 << (r_bits % 64)) - 1)
    zero = ~raw.any(axis=1)
    raw[zero, 0] = 1
    words = raw.view(np.uint32).reshape(count, 2 * w64)
    return np.ascontiguousarray(words[:, : (r_bits + 31) // 32])


# ---------------------------------------------------------------------------
# Context and keys
# ---------------------------------------------------------------------------


def _ptr(t):
    if t is None:
        return None
    if isinstance(t, torch.Tensor):
        return ctypes.c_void_p(t.data_ptr())
    if isinstance(t, np.ndarray):
        return t.ctypes.data_as(ctypes.c_void_p)
    raise TypeError(type(t))


class Context:
    """A device plus the stream work is queued on (torch's current stream)."""

    def __init__(self, device: int = 0):
        self.device = torch.device("cuda", device)
        h = ctypes.c_void_p()
        check(lib().heu_ctx_create(device, ctypes.byref(h)))
        self._h = h
        stream = torch.cuda.current_stream(self.device)
        check(lib().heu_ctx_set_stream(self._h, ctypes.c_void_p(stream.cuda_stream)))

    @property
    def handle(self):
        return self._h

    def sync(self) -> None:
        check(lib().heu_ctx_sync(self._h))

    def profile(self, enable: bool = True) -> None:
        """Per-kernel event timing on/off (clears the records)."""
        check(lib().heu_ctx_profile(self._h, int(enable)))

    def profile_read(self, kernel: str):
        """(total milliseconds, launches) of `kernel` since profiling was enabled."""
        ms, n = ctypes.c_double(), ctypes.c_uint64()
        check(lib().heu_ctx_profile_read(self._h, kernel.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and _native._lib is not None:
            lib().heu_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def keygen(key_bits: int, seed: int | None = None) -> dict:
    """paillier_zahlen/key_generator.cc:26-60 (seeded when seed is not None)."""
    n = np.zeros(key_bits // 32, np.uint32)
    hs = np.zeros(key_bits // 16, np.uint32)
    p = np.zeros(key_bits // 64, np.uint32)
    q = np.zeros(key_bits // 64, np.uint32)
    sp = ctypes.byref(ctypes.c_uint64(seed)) if seed is not None else None
    check(lib().heu_keygen(key_bits, sp, _ptr(n), _ptr(hs), _ptr(p), _ptr(q)))
    return {"bits": key_bits, "n": words_to_int(n), "hs": words_to_int(hs), "p": words_to_int(p),
            "q": words_to_int(q)}


class Key:
    """A Paillier key resident on the device (public part, optional secret part)."""

    def __init__(self, ctx: Context, n: int, hs: int, p: int | None = None, q: int | None = None,
                 r_bits: int = 256, window_bits: int = 0):
        self.ctx = ctx
        self.n, self.hs, self.p, self.q = n, hs, p, q
        n_words = (n.bit_length() + 31) // 32
        nw = int_to_words(n, n_words)
        hw = int_to_words(hs, 2 * n_words)
        h = ctypes.c_void_p()
        if p is not None:
            pw_words = (max(p.bit_length(), q.bit_length()) + 31) // 32
            pw, qw = int_to_words(p, pw_words), int_to_words(q, pw_words)
            check(lib().heu_key_create(ctx.handle, _ptr(nw), _ptr(hw), n_words, _ptr(pw), _ptr(qw), pw_words,
                                       r_bits, window_bits, ctypes.byref(h)))
        else:
            check(lib().heu_key_create(ctx.handle, _ptr(nw), _ptr(hw), n_words, None, None, 0, r_bits,
                                       window_bits, ctypes.byref(h)))
        self._h = h
        kb, nwd, cw, sk = (ctypes.c_uint32() for _ in range(4))
        check(lib().heu_key_info(h, ctypes.byref(kb), ctypes.byref(nwd), ctypes.byref(cw), ctypes.byref(sk)))
        self.key_bits, self.n_words, self.ct_words, self.has_sk = kb.value, nwd.value, cw.value, bool(sk.value)
        self.r_bits = r_bits
        self.window_bits = window_bits

    @classmethod
    def from_keygen(cls, ctx: Context, kd: dict, **kw) -> "Key":
        return cls(ctx, kd["n"], kd["hs"], kd.get("p"), kd.get("q"), **kw)

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and _native._lib is not None:
            lib().heu_key_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def device_bytes(self) -> int:
        return lib().heu_key_device_bytes(self._h)

    # ------------------------------------------------------------- helpers
    def _dev(self, a, dtype) -> torch.Tensor:
        if isinstance(a, torch.Tensor):
            t = a.to(self.ctx.device)
        else:
            a = np.ascontiguousarray(a)
            if not a.flags.writeable:  # a broadcast view: torch shares only writable memory
                a = a.copy()
            if a.dtype == np.uint32:
                a = a.view(np.int32)
            elif a.dtype == np.uint8:
                pass
            t = torch.from_numpy(a).to(self.ctx.device)
        if t.dtype != dtype:
            t = t.to(dtype)
        return t.contiguous()

    def _ct_out(self, count: int) -> torch.Tensor:
        return torch.empty((count, self.ct_words), dtype=torch.int32, device=self.ctx.device)

    def _rows(self, t: torch.Tensor) -> int:
        if t.dim() != 2 or t.shape[1] != self.ct_words:
            raise ValueError(f"ciphertexts must have shape (N, {self.ct_words})")
        return t.shape[0]

    def _check_rows(self, n: int, cts=(), rows=(), rands=()) -> None:
        """Every further operand holds exactly n rows, so each device read stays
        inside its array: ciphertext arrays (n, ct_words), per-element arrays
        (n,), randomness (n, r_words)."""
        for t in cts:
            if self._rows(t) != n:
                raise ValueError(f"ciphertext operands must have {n} rows, got {t.shape[0]}")
        for t in rows:
            if t.dim() != 1 or t.shape[0] != n:
                raise ValueError(f"per-element operands must have shape ({n},), got {tuple(t.shape)}")
        for t in rands:
            if t.dim() != 2 or t.shape[0] != n or t.shape[1] == 0:
                raise ValueError(f"r must have shape ({n}, r_words), got {tuple(t.shape)}")

    @staticmethod
    def _segments(offsets, positions: int) -> np.ndarray:
        """Host offsets of a segmented reduction, checked so that every device
        read stays inside the `positions` elements it addresses."""
        offs = np.ascontiguousarray(offsets, dtype=np.int64)
        if offs.ndim != 1 or offs.size < 1:
            raise ValueError("offsets must be a 1-D array of segments + 1 entries")
        if offs.size > 1 and (offs[0] < 0 or offs[-1] > positions or bool(np.any(np.diff(offs) < 0))):
            raise ValueError(f"offsets must be non-decreasing within [0, {positions}]")
        return offs

    @staticmethod
    def _check_per_position(t: torch.Tensor, offs: np.ndarray, what: str) -> None:
        if offs.size > 1 and t.shape[0] < offs[-1]:
            raise ValueError(f"{what} has {t.shape[0]} entries; the segments address {int(offs[-1])}")

    @staticmethod
    def _check_indices(idx: torch.Tensor, rows: int) -> None:
        if idx.numel() and (int(idx.min()) < 0 or int(idx.max()) >= rows):
            raise ValueError(f"indices must lie in [0, {rows})")

    # ------------------------------------------------------------- tuning
    @property
    def table_window(self) -> int:
        v = ctypes.c_uint32()
        check(lib().heu_key_get_param(self._h, _native.HEU_PARAM_TABLE_WINDOW, ctypes.byref(v)))
        return v.value

    @table_window.setter
    def table_window(self, bits: int) -> None:
        check(lib().heu_key_set_param(self._h, _native.HEU_PARAM_TABLE_WINDOW, bits))

    @property
    def dec_lanes(self) -> int:
        v = ctypes.c_uint32()
        check(lib().heu_key_get_param(self._h, _native.HEU_PARAM_DEC_LANES, ctypes.byref(v)))
        return v.value

    @dec_lanes.setter
    def dec_lanes(self, lanes: int) -> None:
        check(lib().heu_key_set_param(self._h, _native.HEU_PARAM_DEC_LANES, lanes))

    def autotune(self, what: int = _native.HEU_TUNE_ALL) -> dict:
        """Times the table window, decryption lanes per bignum and work-list
        grid multiplier on this device, keeps the fastest; returns the report."""
        import json

        buf = ctypes.create_string_buffer(8192)
        check(lib().heu_autotune(self._h, what, buf, len(buf)))
        return json.loads(buf.value.decode())

    # ------------------------------------------------------------- ops
    def encrypt_i64(self, m, r) -> torch.Tensor:
        m = self._dev(m, torch.int64)
        r = self._dev(r, torch.int32)
        self._check_rows(m.shape[0], rows=(m,), rands=(r,))
        out = self._ct_out(m.shape[0])
        check(lib().heu_encrypt_i64(self._h, _ptr(m), _ptr(r), r.shape[1], m.shape[0], _ptr(out)))
        return out

    def encrypt(self, values, r) -> torch.Tensor:
        """Arbitrary Python ints (|m| <= n/2)."""
        mw = self.n_words
        mh = ints_to_words(values, mw, signed=True)
        md = self._dev(mh, torch.int32)
        r = self._dev(r, torch.int32)
        self._check_rows(len(values), rands=(r,))
        out = self._ct_out(len(values))
        check(lib().heu_encrypt(self._h, _ptr(md), _ptr(mh), mw, _ptr(r), r.shape[1], len(values), _ptr(out)))
        return out

    def randomize(self, ct, r) -> torch.Tensor:
        r = self._dev(r, torch.int32)
        if ct is None:
            n = r.shape[0]
            out = self._ct_out(n)
            check(lib().heu_randomize(self._h, None, _ptr(r), r.shape[1], n, _ptr(out)))
            return out
        ct = self._dev(ct, torch.int32)
        n = self._rows(ct)
        self._check_rows(n, rands=(r,))
        out = self._ct_out(n)
        check(lib().heu_randomize(self._h, _ptr(ct), _ptr(r), r.shape[1], n, _ptr(out)))
        return out

    def decrypt(self, ct, out_words: int | None = None, overflow: bool = False):
        ct = self._dev(ct, torch.int32)
        n = self._rows(ct)
        ow = out_words or self.n_words
        out = torch.empty((n, ow), dtype=torch.int32, device=self.ctx.device)
        ovf = torch.empty((n,), dtype=torch.uint8, device=self.ctx.device) if overflow else None
        check(lib().heu_decrypt(self._h, _ptr(ct), n, _ptr(out), ow, _ptr(ovf)))
        return (out, ovf) if overflow else out

    def add(self, a, b) -> torch.Tensor:
        a, b = self._dev(a, torch.int32), self._dev(b, torch.int32)
        n = self._rows(a)
        self._check_rows(n, cts=(b,))
        out = self._ct_out(n)
        check(lib().heu_add(self._h, _ptr(a), _ptr(b), n, _ptr(out)))
        return out

    def add_i64(self, a, p) -> torch.Tensor:
        a, p = self._dev(a, torch.int32), self._dev(p, torch.int64)
        n = self._rows(a)
        self._check_rows(n, rows=(p,))
        out = self._ct_out(n)
        check(lib().heu_add_i64(self._h, _ptr(a), _ptr(p), n, _ptr(out)))
        return out

    def sub(self, a, b) -> torch.Tensor:
        a, b = self._dev(a, torch.int32), self._dev(b, torch.int32)
        n = self._rows(a)
        self._check_rows(n, cts=(b,))
        out = self._ct_out(n)
        check(lib().heu_sub(self._h, _ptr(a), _ptr(b), n, _ptr(out)))
        return out

    def negate(self, a) -> torch.Tensor:
        a = self._dev(a, torch.int32)
        n = self._rows(a)
        out = self._ct_out(n)
        check(lib().heu_negate(self._h, _ptr(a), n, _ptr(out)))
        return out

    def mul_i64(self, a, k) -> torch.Tensor:
        a, k = self._dev(a, torch.int32), self._dev(k, torch.int64)
        n = self._rows(a)
        self._check_rows(n, rows=(k,))
        out = self._ct_out(n)
        check(lib().heu_mul_i64(self._h, _ptr(a), _ptr(k), n, _ptr(out)))
        return out

    def mul(self, a, scalars) -> torch.Tensor:
        """ct x scalar for Python ints of any size (evaluator.cc:99-123)."""
        a = self._dev(a, torch.int32)
        n = self._rows(a)
        if len(scalars) != n:
            raise ValueError(f"need {n} scalars, got {len(scalars)}")
        mags = [abs(int(v)) for v in scalars]
        kw = max(1, max((m.bit_length() + 31) // 32 for m in mags) if mags else 1)
        kmag = self._dev(ints_to_words(mags, kw), torch.int32)
        kneg = self._dev(np.array([int(v) < 0 for v in scalars], dtype=np.uint8), torch.uint8)
        out = self._ct_out(n)
        check(lib().heu_mul(self._h, _ptr(a), _ptr(kmag), _ptr(kneg), kw, n, _ptr(out)))
        return out

    def sum(self, ct, offsets) -> torch.Tensor:
        ct = self._dev(ct, torch.int32)
        offs = self._segments(offsets, self._rows(ct))
        segs = len(offs) - 1
        out = self._ct_out(segs)
        check(lib().heu_sum(self._h, _ptr(ct), _ptr(offs), segs, _ptr(out)))
        return out

    def sum_noalloc(self, ct, offsets, out: torch.Tensor) -> torch.Tensor:
        """sum() into a caller-provided output, with no device allocation by the
        engine (heu_sum_noalloc: scratch in a fixed per-GPU array)."""
        ct = self._dev(ct, torch.int32)
        offs = self._segments(offsets, self._rows(ct))
        segs = len(offs) - 1
        if not out.is_contiguous() or tuple(out.shape) != (segs, self.ct_words) or out.dtype != torch.int32:
            raise ValueError(f"out must be a contiguous int32 tensor of shape ({segs}, {self.ct_words})")
        check(lib().heu_sum_noalloc(self._h, _ptr(ct), _ptr(offs), segs, _ptr(out)))
        return out

    def dot_i64(self, ct, w, offsets) -> torch.Tensor:
        ct, w = self._dev(ct, torch.int32), self._dev(w, torch.int64)
        offs = self._segments(offsets, self._rows(ct))
        self._check_per_position(w, offs, "w")
        segs = len(offs) - 1
        out = self._ct_out(segs)
        check(lib().heu_dot_i64(self._h, _ptr(ct), _ptr(w), _ptr(offs), segs, _ptr(out)))
        return out

    def sum_gather(self, ct, idx, offsets) -> torch.Tensor:
        """out[s] = prod ct[idx[j]] over positions j of segment s."""
        ct, idx = self._dev(ct, torch.int32), self._dev(idx, torch.int64)
        offs = self._segments(offsets, idx.shape[0])
        self._check_indices(idx[: int(offs[-1])], self._rows(ct))
        segs = len(offs) - 1
        out = self._ct_out(segs)
        check(lib().heu_sum_gather(self._h, _ptr(ct), _ptr(idx), _ptr(offs), segs, _ptr(out)))
        return out

    def dot_gather(self, ct, idx, w, offsets) -> torch.Tensor:
        """out[s] = prod ct[idx[j]]^w[j] over positions j of segment s."""
        ct, idx, w = self._dev(ct, torch.int32), self._dev(idx, torch.int64), self._dev(w, torch.int64)
        offs = self._segments(offsets, idx.shape[0])
        self._check_indices(idx[: int(offs[-1])], self._rows(ct))
        self._check_per_position(w, offs, "w")
        segs = len(offs) - 1
        out = self._ct_out(segs)
        check(lib().heu_dot_gather(self._h, _ptr(ct), _ptr(idx), _ptr(w), _ptr(offs), segs, _ptr(out)))
        return out

    def prefix_mul(self, ct, offsets) -> torch.Tensor:
        """Inclusive prefix products within each segment."""
        ct = self._dev(ct, torch.int32)
        offs = self._segments(offsets, self._rows(ct))
        out = self._ct_out(self._rows(ct))
        check(lib().heu_prefix_mul(self._h, _ptr(ct), _ptr(offs), len(offs) - 1, _ptr(out)))
        return out

    # ------------------------------------------------------------- engine form
    @property
    def mform_words(self) -> int:
        w = ctypes.c_uint32()
        check(lib().heu_key_mform_words(self._h, ctypes.byref(w)))
        return w.value

    def _mform_rows(self, t: torch.Tensor) -> int:
        if t.dim() != 2 or t.shape[1] != self.mform_words:
            raise ValueError(f"engine-form values must have shape (N, {self.mform_words})")
        return t.shape[0]

    def to_mform(self, ct) -> torch.Tensor:
        """Z-space ciphertexts -> engine form (R29 Montgomery limbs, lazy)."""
        ct = self._dev(ct, torch.int32)
        n = self._rows(ct)
        out = torch.empty((n, self.mform_words), dtype=torch.int32, device=self.ctx.device)
        check(lib().heu_to_mform(self._h, _ptr(ct), n, _ptr(out)))
        return out

    def from_mform(self, mf) -> torch.Tensor:
        mf = self._dev(mf, torch.int32)
        n = self._mform_rows(mf)
        out = self._ct_out(n)
        check(lib().heu_from_mform(self._h, _ptr(mf), n, _ptr(out)))
        return out

    def add_mform(self, a, b) -> torch.Tensor:
        a, b = self._dev(a, torch.int32), self._dev(b, torch.int32)
        n = self._mform_rows(a)
        if self._mform_rows(b) != n:
            raise ValueError("engine-form operands must have the same number of rows")
        out = torch.empty_like(a)
        check(lib().heu_add_mform(self._h, _ptr(a), _ptr(b), n, _ptr(out)))
        return out

    def sum_mform(self, mf, offsets) -> torch.Tensor:
        mf = self._dev(mf, torch.int32)
        offs = self._segments(offsets, self._mform_rows(mf))
        segs = len(offs) - 1
        out = torch.empty((segs, self.mform_words), dtype=torch.int32, device=self.ctx.device)
        check(lib().heu_sum_mform(self._h, _ptr(mf), _ptr(offs), segs, _ptr(out)))
        return out

    def to_mont(self, ct, word_bits: int = 64) -> torch.Tensor:
        ct = self._dev(ct, torch.int32)
        n = self._rows(ct)
        out = self._ct_out(n)
        check(lib().heu_to_mont(self._h, _ptr(ct), n, word_bits, _ptr(out)))
        return out

    def from_mont(self, ct, word_bits: int = 64) -> torch.Tensor:
        ct = self._dev(ct, torch.int32)
        n = self._rows(ct)
        out = self._ct_out(n)
        check(lib().heu_from_mont(self._h, _ptr(ct), n, word_bits, _ptr(out)))
        return out


class _TaskOps:
    """Task-recording methods shared by Batch and Server: each converts its
    operands to device tensors, allocates the output and hands one heu_task to
    `_task`, returning what `_task` returns."""

    def _record(self, key, op, a=None, b=None, count=0, out=None, aux=0, offsets=None, segments=0):
        keep = [x for x in (a, b, out, offsets) if x is not None]
        addr = lambda x: None if x is None else (x.data_ptr() if isinstance(x, torch.Tensor) else x.ctypes.data)
        t = _native.HeuTask(key._h.value, op, aux, addr(a), addr(b), count, addr(offsets), segments, addr(out))
        return self._task(t, out, keep)

    def encrypt_i64(self, key: "Key", m, r, out: torch.Tensor | None = None) -> torch.Tensor:
        m, r = key._dev(m, torch.int64), key._dev(r, torch.int32)
        n = m.shape[0]
        key._check_rows(n, rows=(m,), rands=(r,))
        if out is None:
            out = key._ct_out(n)
        elif not out.is_contiguous() or tuple(out.shape) != (n, key.ct_words) or out.dtype != torch.int32:
            raise ValueError(f"out must be a contiguous int32 tensor of shape ({n}, {key.ct_words})")
        return self._record(key, _native.HEU_OP_ENCRYPT_I64, m, r, n, out, aux=r.shape[1])

    def randomize(self, key: "Key", ct, r) -> torch.Tensor:
        r = key._dev(r, torch.int32)
        ct = None if ct is None else key._dev(ct, torch.int32)
        n = r.shape[0]
        key._check_rows(n, cts=() if ct is None else (ct,), rands=(r,))
        return self._record(key, _native.HEU_OP_RANDOMIZE, ct, r, n, key._ct_out(n), aux=r.shape[1])

    def add(self, key: "Key", a, b) -> torch.Tensor:
        a, b = key._dev(a, torch.int32), key._dev(b, torch.int32)
        n = key._rows(a)
        key._check_rows(n, cts=(b,))
        return self._record(key, _native.HEU_OP_ADD, a, b, n, key._ct_out(n))

    def add_i64(self, key: "Key", a, p) -> torch.Tensor:
        a, p = key._dev(a, torch.int32), key._dev(p, torch.int64)
        n = key._rows(a)
        key._check_rows(n, rows=(p,))
        return self._record(key, _native.HEU_OP_ADD_I64, a, p, n, key._ct_out(n))

    def mul_i64(self, key: "Key", a, k) -> torch.Tensor:
        a, k = key._dev(a, torch.int32), key._dev(k, torch.int64)
        n = key._rows(a)
        key._check_rows(n, rows=(k,))
        return self._record(key, _native.HEU_OP_MUL_I64, a, k, n, key._ct_out(n))

    def decrypt(self, key: "Key", ct, out_words: int | None = None) -> torch.Tensor:
        ct = key._dev(ct, torch.int32)
        n = key._rows(ct)
        ow = out_words or key.n_words
        out = torch.empty((n, ow), dtype=torch.int32, device=key.ctx.device)
        return self._record(key, _native.HEU_OP_DECRYPT, ct, None, n, out, aux=ow)

    def sum(self, key: "Key", ct, offsets) -> torch.Tensor:
        ct = key._dev(ct, torch.int32)
        offs = key._segments(offsets, key._rows(ct))
        segs = len(offs) - 1
        return self._record(key, _native.HEU_OP_SUM, ct, None, 0, key._ct_out(segs), offsets=offs, segments=segs)

    def dot_i64(self, key: "Key", ct, w, offsets) -> torch.Tensor:
        ct, w = key._dev(ct, torch.int32), key._dev(w, torch.int64)
        offs = key._segments(offsets, key._rows(ct))
        key._check_per_position(w, offs, "w")
        segs = len(offs) - 1
        return self._record(key, _native.HEU_OP_DOT_I64, ct, w, 0, key._ct_out(segs), offsets=offs, segments=segs)


class Batch(_TaskOps):
    """Many tasks (vectors of any length, on any mix of 2048/3072-bit keys) run
    as one dispatch: persistent work-list launches plus follow-up steps,
    optionally captured into a hipGraph and replayed (heu_batch_* C ABI).

    Each method records a task and returns its (preallocated) output tensor;
    `run()` plans the batch on first use and enqueues one run.  Inputs are kept
    alive by the batch; the same buffers are read at every run.  Tasks between
    barriers must not read each other's outputs."""

    def __init__(self, ctx: Context, graph: bool = False):
        self.ctx = ctx
        self.graph = graph
        self._tasks = []
        self._keep = []
        self._h = None

    def _task(self, t, out, keep):
        if self._h is not None:
            raise RuntimeError("batch already planned")
        self._keep += keep
        self._tasks.append(t)
        return out

    def barrier(self) -> None:
        """Tasks recorded after this see the outputs of every task before it."""
        if self._h is not None:
            raise RuntimeError("batch already planned")
        self._tasks.append(_native.HeuTask(None, _native.HEU_OP_BARRIER, 0, None, None, 0, None, 0, None))

    def plan(self) -> None:
        if self._h is not None:
            return
        arr = (_native.HeuTask * len(self._tasks))(*self._tasks)
        h = ctypes.c_void_p()
        flags = _native.HEU_BATCH_GRAPH if self.graph else 0
        check(lib().heu_batch_create(self.ctx._h, arr, len(self._tasks), flags, ctypes.byref(h)))
        self._h = h

    def run(self) -> None:
        self.plan()
        check(lib().heu_batch_run(self._h))

    def stats(self) -> dict:
        self.plan()
        la, un, by = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_size_t()
        check(lib().heu_batch_stats(self._h, ctypes.byref(la), ctypes.byref(un), ctypes.byref(by)))
        return {"launches": la.value, "units": un.value, "device_bytes": by.value}

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and _native._lib is not None:
            lib().heu_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server(_TaskOps):
    """Request server with continuous batching (heu_server_* C ABI).

    Every op method submits one request and returns (ticket, output tensor);
    `wait(ticket)` blocks until the output is written and returns the request's
    latency in ms (submit -> completion).  Inputs must be on the device (they
    are, once an op method returns: host operands are copied synchronously), and
    the server keeps every operand alive until the request is waited."""

    def __init__(self, ctx: Context, workspace_bytes: int = 0, max_inflight: int = 0, max_batch_tasks: int = 0,
                 batch_rounds: int = 0, parts_per_round: int = 0, urgent_ms: int = 0, max_workspace_bytes: int = 0):
        self.ctx = ctx
        self._keep = {}
        cfg = _native.HeuServerConfig(workspace_bytes, max_inflight, max_batch_tasks, batch_rounds, parts_per_round,
                                      urgent_ms, max_workspace_bytes)
        h = ctypes.c_void_p()
        check(lib().heu_server_create(ctx._h, ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h

    def _task(self, t, out, keep):
        tk = ctypes.c_uint64()
        check(lib().heu_server_submit(self._h, ctypes.byref(t), ctypes.byref(tk)))
        self._keep[tk.value] = keep
        return tk.value, out

    def wait(self, ticket: int, timeout_ms: float = -1.0) -> float:
        lat = ctypes.c_double()
        rc = lib().heu_server_wait(self._h, ticket, timeout_ms, ctypes.byref(lat))
        if rc == _native.HEU_OK or "timeout" not in (lib().heu_last_error() or b"").decode():
            self._keep.pop(ticket, None)
        check(rc)
        return lat.value

    def stats(self) -> dict:
        st = _native.HeuServerStats()
        check(lib().heu_server_stats(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in st._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and _native._lib is not None:
            lib().heu_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
