"""Arrays of Paillier ciphertexts on the device: the HeMatrix surface of
HEU's `heu.numpy` (heu/pylib/numpy_binding/bind_numpy.cc:321-510,
heu/library/numpy/evaluator.h:24-92) for the Paillier path, over the engine.

  kit = hnp.setup(2048)                        # keys resident on the GPU
  x = kit.encryptor().encrypt(np_int64_array)  # CiphertextArray (1-D or 2-D)
  y = kit.evaluator().add(x, other)            # numpy broadcasting, ct/pt mixes
  s = kit.evaluator().select_sum(x, (slice(None), 0))
  h = kit.evaluator().feature_wise_bucket_sum(x, subgroup_map, order_map, 16)
  z = kit.decryptor().decrypt(y)               # numpy int64 (object if wider)

Plaintext arrays are numpy integer arrays (the reference wraps them in
PlaintextArray).  Ciphertext arrays keep C order on the device; every
reduction, selection and bucket sum runs as one gathered engine call
(`heu_sum_gather` / `heu_dot_gather`), so no gathered copy is made.

Ciphertext arrays index as HeMatrix does (pylib/numpy_binding/py_slicer.cc),
which differs from numpy: a 2-tuple of row and column keys selects the block
of their cross product (`x[[1, 2], [0, 3]]` is 2 x 2), and any other key
selects rows over every column (`x[()]` selects nothing, `x[(5,)]` is a
one-row matrix).  Keys are ints, slices and sequences of ints; 1-D boolean
masks are also taken.  `select_sum` and `batch_select_sum` read keys the same
way for plaintext and ciphertext arrays; plaintext arrays themselves index as
numpy does.
"""
from __future__ import annotations

import os

import numpy as np
import torch

from . import engine as E
from . import wire


class MatrixSerializeFormat:
    """numpy/matrix.h MatrixSerializeFormat: Best is the msgpack form,
    Interconnection the org_interconnection protobuf (numpy/matrix.cc:39-111;
    its .proto is not in the reference tree, so its field numbers are the
    engine's stated ones)."""

    Best = "best"
    Interconnection = "interconnection"


# yacl's BigInt backends (OpenSSL, GMP) use 64-bit words: R = 2^(64 ceil(bits / 64))
MONT_WORD_BITS = 64


class CiphertextArray:
    """A 0-, 1- or 2-D array of ciphertexts of one key, `data` being a device
    tensor of shape (size, ct_words), C order."""

    def __init__(self, key: E.Key, data: torch.Tensor, shape):
        self.key = key
        self.data = data
        self.shape = tuple(int(s) for s in shape)
        if data.shape[0] != int(np.prod(self.shape, dtype=np.int64)):
            raise ValueError(f"{data.shape[0]} ciphertexts do not fill shape {self.shape}")

    @property
    def ndim(self) -> int:
        return len(self.shape)

    @property
    def size(self) -> int:
        return int(np.prod(self.shape, dtype=np.int64))

    @property
    def rows(self) -> int:
        """Rows of the array as the reference's matrix (bind_numpy.cc:62-63)."""
        return self.shape[0] if self.ndim else 1

    @property
    def cols(self) -> int:
        """Columns of the array as the reference's matrix (a vector has one)."""
        return self.shape[1] if self.ndim == 2 else 1

    def _flat_index(self, key) -> np.ndarray:
        """Positions in `data` of the elements a HeMatrix key selects, in the
        C order of the selected block."""
        return _select_positions(self.shape, key)

    def __getitem__(self, key) -> "CiphertextArray":
        """HeMatrix indexing (pylib/numpy_binding/py_slicer.cc:58-86): a
        2-tuple selects the block of its row and column indices, any other
        key selects rows; an int index drops its axis."""
        block, shape, transposed = _heu_block(self.shape, key)
        sel = block.T if transposed else block
        idx = torch.from_numpy(np.ascontiguousarray(sel.ravel())).to(self.data.device)
        return CiphertextArray(self.key, self.data.index_select(0, idx), shape)

    def __setitem__(self, key, value) -> None:
        """HeMatrix assignment (py_slicer.cc:88-120, numpy/matrix.h:125-135):
        the block the key selects (ints do not drop axes here) takes a
        ciphertext array of the block's shape; a 1-D array fills a one-row
        block.  Where an index repeats, the last write wins."""
        if not isinstance(value, CiphertextArray):
            raise TypeError(f"Unsupported value type [{type(value).__name__}] for __setitem__")
        if value.key is not self.key:
            raise ValueError("the assigned ciphertexts belong to another key")
        block = _block_of(self.shape, key)[0]
        pr, pc = value.rows, value.cols
        if value.ndim == 1 and block.shape[0] == 1 and block.shape[1] > 1:
            pr, pc = pc, pr
        if (pr, pc) != block.shape:
            raise ValueError(f"cannot assign a {pr} x {pc} array to a {block.shape[0]} x {block.shape[1]} block")
        target = block.ravel()
        # last write wins, as in the element-by-element assignment
        _, first_from_end = np.unique(target[::-1], return_index=True)
        src = (target.size - 1 - first_from_end).astype(np.int64)
        dev = self.data.device
        self.data[torch.from_numpy(target[src]).to(dev)] = value.data[torch.from_numpy(src).to(dev)]

    def transpose(self) -> "CiphertextArray":
        if self.ndim < 2:
            return self
        return self._take(np.arange(self.size, dtype=np.int64).reshape(self.shape).T)

    @property
    def T(self) -> "CiphertextArray":
        return self.transpose()

    def _take(self, grid: np.ndarray) -> "CiphertextArray":
        idx = torch.from_numpy(np.ascontiguousarray(grid.ravel())).to(self.data.device)
        return CiphertextArray(self.key, self.data.index_select(0, idx), grid.shape)

    def reshape(self, *shape) -> "CiphertextArray":
        shape = shape[0] if len(shape) == 1 and isinstance(shape[0], (tuple, list)) else shape
        shape = np.empty(self.size, dtype=np.int8).reshape(shape).shape
        return CiphertextArray(self.key, self.data, shape)

    def __len__(self) -> int:
        return self.shape[0] if self.ndim else 1

    def __repr__(self) -> str:
        return f"CiphertextArray(shape={self.shape}, key_bits={self.key.key_bits})"

    # ------------------------------------------------------------ wire format
    def _colmajor(self) -> np.ndarray:
        """Positions of `data` in Eigen's column-major order of (rows, cols)."""
        return self._colmajor_of(self.rows, self.cols)

    def serialize(self, format: str = MatrixSerializeFormat.Best) -> bytes:
        """DenseMatrix<Ciphertext>::Serialize (numpy/matrix.h:221-265), or
        Serialize4Ic (numpy/matrix.cc:39-76) with the Interconnection format:
        every ciphertext in the Montgomery form the reference stores,
        column-major."""
        put = {MatrixSerializeFormat.Best: wire.put_matrix, MatrixSerializeFormat.Interconnection: wire.put_matrix_ic}
        if format not in put:
            raise ValueError(f"unknown matrix serialize format {format!r}")
        key = self.key
        if self.size == 0:
            mont = np.zeros((0, key.ct_words), np.uint32)
        else:
            mont = key.to_mont(self.data, MONT_WORD_BITS).cpu().numpy().view(np.uint32)
            mont = mont[self._colmajor()]
        return put[format](mont, self.rows, self.cols, self.ndim)

    @staticmethod
    def load_from(key, buf: bytes, format: str = MatrixSerializeFormat.Best) -> "CiphertextArray":
        """DenseMatrix<Ciphertext>::LoadFrom (numpy/matrix.h:267-314), or
        LoadFromIc (numpy/matrix.cc:78-111), onto the device of `key` (an
        engine Key or a HeKit)."""
        get = {MatrixSerializeFormat.Best: wire.get_matrix, MatrixSerializeFormat.Interconnection: wire.get_matrix_ic}
        if format not in get:
            raise ValueError(f"unknown matrix serialize format {format!r}")
        key = key.key if isinstance(key, HeKit) else key
        rows, cols, ndim, mont = get[format](bytes(buf), key.ct_words)
        shape = () if ndim == 0 else ((rows,) if ndim == 1 else (rows, cols))
        if int(np.prod(shape, dtype=np.int64)) != rows * cols:
            raise ValueError(f"a {ndim}-d matrix cannot hold {rows} x {cols} elements")
        if rows * cols == 0:
            return CiphertextArray(key, key._ct_out(0), shape)
        z = key.from_mont(mont, MONT_WORD_BITS)
        # column-major item k holds C-order element (k % rows) * cols + k // rows
        inv = np.empty(rows * cols, dtype=np.int64)
        inv[CiphertextArray._colmajor_of(rows, cols)] = np.arange(rows * cols, dtype=np.int64)
        return CiphertextArray(key, z.index_select(0, torch.from_numpy(inv).to(z.device)), shape)

    @staticmethod
    def _colmajor_of(rows: int, cols: int) -> np.ndarray:
        k = np.arange(rows * cols, dtype=np.int64)
        return (k % rows) * cols + k // rows


# ---------------------------------------------------------------- HeMatrix keys
def _checked_indices(a: np.ndarray, n: int) -> np.ndarray:
    """slice_tool.cc:22-33: indices below n; a negative one counts from the end."""
    a = a.astype(np.int64, copy=True)
    hi = np.nonzero(a >= n)[0]
    if hi.size:
        raise IndexError(f"index {int(a[hi[0]])} is out of bounds [0, {n})")
    a[a < 0] += n
    lo = np.nonzero(a < 0)[0]
    if lo.size:
        raise IndexError(f"index {int(a[lo[0]]) - n} is out of bounds [{-n}, {n})")
    return a


def _axis_index(k, n: int):
    """slice_tool::Parse (slice_tool.cc:35-87) against an axis of length n:
    a slice, a sequence of ints (list, tuple, 1-D integer array) or an int
    -> (indices, whether the axis is dropped).  A 1-D boolean mask array of
    length n is also taken; the reference has no masks."""
    if isinstance(k, slice):
        return np.arange(*k.indices(n), dtype=np.int64), False
    if isinstance(k, (int, np.integer)):
        return _checked_indices(np.array([int(k)]), n), True
    if isinstance(k, np.ndarray):
        if k.dtype == np.bool_:
            if k.ndim != 1 or k.shape[0] != n:
                raise IndexError(f"a boolean index of shape {k.shape} does not match an axis of length {n}")
            return np.nonzero(k)[0].astype(np.int64), False
        if k.ndim != 1 or not np.issubdtype(k.dtype, np.integer):
            raise TypeError(f"an index array must be 1-D of integers, got {k.ndim}-D {k.dtype}")
        return _checked_indices(k, n), False
    if isinstance(k, (list, tuple, range)):
        if len(k) == 0:
            return np.zeros(0, np.int64), False
        for item in k:
            if not isinstance(item, (int, np.integer)):
                raise TypeError(f"indices array element must be integers, got {item!r} with type {type(item).__name__}")
        return _checked_indices(np.array([int(i) for i in k]), n), False
    raise TypeError(f"Unsupported indices type {type(k).__name__}")


def _heu_key(key, ndim: int, rows: int, cols: int):
    """py_slicer.cc:58-86 and extension_functions.cc:29-57: a 2-tuple indexes
    rows and columns, and selects the block of their cross product; any other
    key, a shorter tuple included, indexes rows over every column.
    -> (row indices, column indices, drop rows, drop columns)."""
    if isinstance(key, tuple):
        if len(key) > ndim:
            raise IndexError(f"too many indices for array, array is {ndim}-dimensional, but {len(key)} were indexed")
        if len(key) == 2:
            r, sr = _axis_index(key[0], rows)
            c, sc = _axis_index(key[1], cols)
            return r, c, sr, sc
    r, sr = _axis_index(key, rows)
    return r, np.arange(cols, dtype=np.int64), sr, False


def _dims(shape):
    """An array as the reference's rows x cols matrix (a vector is one
    column, a scalar 1 x 1)."""
    return (shape[0] if len(shape) else 1), (shape[1] if len(shape) == 2 else 1)


def _block_of(shape, key):
    rows, cols = _dims(shape)
    r, c, sr, sc = _heu_key(key, len(shape), rows, cols)
    return r[:, None] * cols + c[None, :], sr, sc


def _select_positions(shape, key) -> np.ndarray:
    """C-order positions of the block a key selects (no axis is dropped)."""
    return _block_of(shape, key)[0].ravel()


def _heu_block(shape, key):
    """The block a key selects and the shape of the result, after
    DenseMatrix::GetItem's squeeze rules (numpy/matrix.h:80-123).
    -> (block of C-order positions, result shape, whether the block is read
    transposed)."""
    ndim = len(shape)
    block, sq_row, sq_col = _block_of(shape, key)
    if ndim == 1 and sq_col:
        raise IndexError("axis doesn't exist, you cannot squeeze shape[1] of a vector")
    if ndim == 0 and (sq_row or sq_col):
        raise IndexError("axis doesn't exist, the array is 0-d")
    nr, nc = block.shape
    min_dim = 1 if (nr > 1 or nc > 1) else 0
    new_dim, transposed = ndim, False
    if (sq_row or sq_col) and ndim != min_dim:
        if sq_col and nc <= 1:  # a column vector, or a scalar
            new_dim = ndim - 1 - (1 if sq_row and nr <= 1 else 0)
        elif sq_row and nr <= 1:  # a dropped row: the one-row block becomes a vector
            new_dim, transposed = ndim - 1, True
        else:
            raise IndexError(f"cannot drop an axis of a {nr} x {nc} block")
    if new_dim == 2:
        out_shape = (nr, nc)
    elif new_dim == 1:
        out_shape = (block.size,)
    elif block.size == 1:
        out_shape = ()
    else:
        raise IndexError("a 0-d array has one element; this key selects none")
    return block, out_shape, transposed


def _is_ct(x) -> bool:
    return isinstance(x, CiphertextArray)


_I64_MIN, _I64_MAX = -(2**63), 2**63 - 1


def _pt(x) -> np.ndarray:
    """A plaintext array: int64 when every value fits, else an object array of
    Python ints (the reference's plaintexts are BigInts)."""
    a = np.asarray(x)
    if a.dtype == object:
        vals = a.ravel().tolist()
        if not all(isinstance(v, (int, np.integer)) and not isinstance(v, (bool, np.bool_)) for v in vals):
            raise TypeError("plaintext arrays must hold integers")
        vals = [int(v) for v in vals]
        if all(_I64_MIN <= v <= _I64_MAX for v in vals):
            return np.array(vals, dtype=np.int64).reshape(a.shape)
        return np.array(vals, dtype=object).reshape(a.shape)
    if not np.issubdtype(a.dtype, np.integer):
        raise TypeError("plaintext arrays must hold integers")
    if a.dtype == np.uint64 and a.size and int(a.max()) > _I64_MAX:
        return np.array([int(v) for v in a.ravel()], dtype=object).reshape(a.shape)
    return a.astype(np.int64)


def _ints(a: np.ndarray) -> list:
    return [int(v) for v in a.ravel()]


def _negated(p: np.ndarray) -> np.ndarray:
    """-p, as Python ints where int64 cannot hold it."""
    if p.dtype == object or (p.size and int(p.min()) == _I64_MIN):
        return np.array([-v for v in _ints(p)], dtype=object)
    return -p


def _add_pt(key: E.Key, cts: torch.Tensor, p: np.ndarray) -> torch.Tensor:
    """Add(ct, p) = ct (1 + n p) mod n^2 (evaluator.cc:46-64): int64 plaintexts
    in one kernel, wider ones as the encryption of p with r = 0, then ct+ct."""
    if p.dtype == object:
        return key.add(cts, key.encrypt(_ints(p), np.zeros((p.size, 1), np.uint32)))
    return key.add_i64(cts, p)


class Encryptor:
    """Encrypts numpy integer arrays (numpy/encryptor.cc:19-66)."""

    def __init__(self, kit: "HeKit"):
        self.kit = kit

    def _r(self, count: int, seed=None) -> np.ndarray:
        key = self.kit.key
        rw = (key.r_bits + 31) // 32
        if seed is not None:
            return E.seeded_r(seed, 1, count, key.r_bits)
        raw = np.frombuffer(os.urandom(count * rw * 4), dtype=np.uint32).reshape(count, rw).copy()
        if key.r_bits % 32:
            raw[:, -1] &= np.uint32((1 << (key.r_bits % 32)) - 1)
        return raw

    def encrypt(self, pts, seed=None) -> CiphertextArray:
        """Integers of any size with |m| <= n/2 (encryptor.cc:35-58); int64
        arrays take the int64 kernel."""
        a = _pt(pts)
        flat = np.ascontiguousarray(a.ravel())
        key = self.kit.key
        if flat.size == 0:
            return CiphertextArray(key, key._ct_out(0), a.shape)
        r = self._r(flat.size, seed)
        ct = key.encrypt(_ints(flat), r) if flat.dtype == object else key.encrypt_i64(flat, r)
        return CiphertextArray(key, ct, a.shape)

    def encrypt_with_audit(self, pts, seed=None):
        """(ciphertexts, audit strings) as numpy/encryptor.cc:107-121 and
        paillier_zahlen/encryptor.cc:50-54 build them: "p:{m},rn:{h_s^r},c:{c}"
        per element, hex, with rn and c in the Montgomery form the reference
        stores (R for 64-bit words)."""
        a = _pt(pts)
        flat = np.ascontiguousarray(a.ravel())
        key = self.kit.key
        if flat.size == 0:
            return CiphertextArray(key, key._ct_out(0), a.shape), np.zeros(a.shape, dtype=object)
        r = self._r(flat.size, seed)
        ct = key.encrypt(_ints(flat), r) if flat.dtype == object else key.encrypt_i64(flat, r)
        rn = key.randomize(None, r)  # EncryptZero with the same r: h_s^r
        c_m = key.to_mont(ct, MONT_WORD_BITS).cpu().numpy().view(np.uint32)
        rn_m = key.to_mont(rn, MONT_WORD_BITS).cpu().numpy().view(np.uint32)
        audit = np.array([f"p:{_hex(int(m))},rn:{_hex(E.words_to_int(x))},c:{_hex(E.words_to_int(y))}"
                          for m, x, y in zip(flat, rn_m, c_m)], dtype=object).reshape(a.shape)
        return CiphertextArray(key, ct, a.shape), audit


def _hex(v: int) -> str:
    """A BigInt's hex string: upper case, no leading zeros, '-' for negatives."""
    return ("-" if v < 0 else "") + format(abs(v), "X")


class Decryptor:
    """Decrypts to numpy arrays: int64 when every value fits, else Python ints
    (numpy/decryptor.cc:19-99)."""

    def __init__(self, kit: "HeKit"):
        self.kit = kit

    def decrypt(self, ct: CiphertextArray):
        key = self.kit.key
        if ct.size == 0:
            return np.zeros(ct.shape, dtype=np.int64)
        out, ovf = key.decrypt(ct.data, out_words=2, overflow=True)
        vals = out.cpu().numpy().view(np.int64).reshape(ct.shape)
        if not bool(ovf.any()):
            return vals
        wide = key.decrypt(ct.data, out_words=key.n_words).cpu().numpy().view(np.uint32)
        objs = np.array([E.twos_to_int(row) for row in wide], dtype=object)
        return objs.reshape(ct.shape)

    def decrypt_in_range(self, ct: CiphertextArray, range_bits: int = 128):
        """Decrypt, and throw unless every plaintext lies in (-2^range_bits,
        2^range_bits), i.e. has at most range_bits bits (numpy/decryptor.h:37,
        numpy/decryptor.cc:19-45; the check blocks plaintext-overflow attacks)."""
        key = self.kit.key
        if ct.size == 0:
            return np.zeros(ct.shape, dtype=np.int64)
        # room for range_bits of magnitude plus the sign, at most the plaintext width
        ow = min(key.n_words, range_bits // 32 + 2)
        out, ovf = key.decrypt(ct.data, out_words=ow, overflow=True)
        rows = out.cpu().numpy().view(np.uint32)
        bad = np.nonzero(ovf.cpu().numpy())[0]
        vals = [E.twos_to_int(r) for r in rows] if ow > 2 else list(rows.view(np.int64).ravel())
        if bad.size == 0:
            over = [i for i, v in enumerate(vals) if abs(int(v)).bit_length() > range_bits]
        else:
            over = [int(bad[0])]
        if over:
            i = over[0]
            found = abs(int(vals[i])).bit_length() if bad.size == 0 else f"> {32 * ow - 1}"
            raise RuntimeError("Dangerous!!! HE ciphertext range check failed, there may be a malicious party "
                               "stealing your data, please stop computing immediately. "
                               f"found pt.BitCount()={found}, expected {range_bits}")
        if all(-(2**63) <= int(v) < 2**63 for v in vals):
            return np.array([int(v) for v in vals], dtype=np.int64).reshape(ct.shape)
        return np.array([int(v) for v in vals], dtype=object).reshape(ct.shape)


class Evaluator:
    """Elementwise ops with numpy broadcasting, matrix products and sums
    (numpy/evaluator.h:24-92, pylib extension_functions.cc)."""

    def __init__(self, kit: "HeKit"):
        self.kit = kit

    # ------------------------------------------------------------ helpers
    def _bcast(self, x, y):
        """Broadcasts two operands; ciphertext operands become flat index
        arrays into their data, plaintext operands flat int64 arrays."""
        sx = x.shape if _is_ct(x) else np.shape(x)
        sy = y.shape if _is_ct(y) else np.shape(y)
        shape = np.broadcast_shapes(sx, sy)

        def expand(v, s):
            if _is_ct(v):
                grid = np.broadcast_to(np.arange(v.size, dtype=np.int64).reshape(s), shape)
                return np.ascontiguousarray(grid.ravel())
            return np.ascontiguousarray(np.broadcast_to(_pt(v), shape).ravel())

        return shape, expand(x, sx), expand(y, sy)

    def _rows_of(self, ct: CiphertextArray, idx: np.ndarray) -> torch.Tensor:
        if idx.size == ct.size and np.array_equal(idx, np.arange(ct.size)):
            return ct.data
        return ct.data.index_select(0, torch.from_numpy(idx).to(ct.data.device))

    def _key(self, *xs) -> E.Key:
        keys = {id(x.key): x.key for x in xs if _is_ct(x)}
        if len(keys) != 1:
            raise ValueError("ciphertext operands must belong to one key")
        return next(iter(keys.values()))

    # ------------------------------------------------------------ elementwise
    def add(self, x, y):
        if not _is_ct(x) and not _is_ct(y):
            return np.add(_pt(x), _pt(y))
        key = self._key(x, y)
        shape, ix, iy = self._bcast(x, y)
        if ix.size == 0:
            return CiphertextArray(key, key._ct_out(0), shape)
        if _is_ct(x) and _is_ct(y):
            out = key.add(self._rows_of(x, ix), self._rows_of(y, iy))
        elif _is_ct(x):
            out = _add_pt(key, self._rows_of(x, ix), iy)
        else:
            out = _add_pt(key, self._rows_of(y, iy), ix)
        return CiphertextArray(key, out, shape)

    def sub(self, x, y):
        if not _is_ct(x) and not _is_ct(y):
            return np.subtract(_pt(x), _pt(y))
        key = self._key(x, y)
        shape, ix, iy = self._bcast(x, y)
        if ix.size == 0:
            return CiphertextArray(key, key._ct_out(0), shape)
        if _is_ct(x) and _is_ct(y):  # evaluator.cc:66-68
            out = key.sub(self._rows_of(x, ix), self._rows_of(y, iy))
        elif _is_ct(x):  # Add(ct, -p), evaluator.cc:70-72
            out = _add_pt(key, self._rows_of(x, ix), _negated(iy))
        else:  # Add(Negate(ct), p), evaluator.cc:78-80
            out = _add_pt(key, key.negate(self._rows_of(y, iy)), ix)
        return CiphertextArray(key, out, shape)

    def mul(self, x, y):
        if not _is_ct(x) and not _is_ct(y):
            return np.multiply(_pt(x), _pt(y))
        if _is_ct(x) and _is_ct(y):
            raise TypeError("ciphertext x ciphertext is not defined for Paillier")
        key = self._key(x, y)
        shape, ix, iy = self._bcast(x, y)
        if ix.size == 0:
            return CiphertextArray(key, key._ct_out(0), shape)
        ct, idx, k = (x, ix, iy) if _is_ct(x) else (y, iy, ix)
        rows = self._rows_of(ct, idx)
        out = key.mul(rows, _ints(k)) if k.dtype == object else key.mul_i64(rows, k)
        return CiphertextArray(key, out, shape)

    # ------------------------------------------------------------ products
    def matmul(self, x, y):
        """x @ y for (ct, pt) or (pt, ct) operands, 1-D or 2-D (numpy/evaluator.cc:209-292):
        out[i, j] = sum_l x[i, l] * y[l, j], one gathered dot per output."""
        if not _is_ct(x) and not _is_ct(y):
            return np.matmul(_pt(x), _pt(y))
        if _is_ct(x) and _is_ct(y):
            raise TypeError("ciphertext @ ciphertext is not defined for Paillier")
        key = self._key(x, y)
        xs = x.shape if _is_ct(x) else np.shape(x)
        ys = y.shape if _is_ct(y) else np.shape(y)
        x2 = xs if len(xs) == 2 else (1, xs[0])
        y2 = ys if len(ys) == 2 else (ys[0], 1)
        n, k = x2
        k2, m = y2
        if k != k2:
            raise ValueError(f"matmul: shapes {xs} and {ys} not aligned")
        out_shape = tuple(([n] if len(xs) == 2 else []) + ([m] if len(ys) == 2 else []))
        i, j, l = np.meshgrid(np.arange(n), np.arange(m), np.arange(k), indexing="ij")
        if _is_ct(x):
            idx = (i * k + l).ravel()
            w = _pt(y).reshape(k, m)[l, j].ravel()
            ct = x
        else:
            idx = (l * m + j).ravel()
            w = _pt(x).reshape(n, k)[i, l].ravel()
            ct = y
        offs = np.arange(0, n * m * k + 1, k, dtype=np.int64)
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        if w.dtype == object:  # weights beyond int64: the powers, then one segmented sum
            terms = key.mul(self._rows_of(ct, idx), _ints(w)) if idx.size else key._ct_out(0)
            out = key.sum(terms, offs)
        else:
            out = key.dot_gather(ct.data, idx, np.ascontiguousarray(w), offs)
        return CiphertextArray(key, out, out_shape)

    # ------------------------------------------------------------ reductions
    def sum(self, x):
        """Sum of every element (numpy/evaluator.cc:355-373); a 0-D array."""
        if not _is_ct(x):
            return np.sum(_pt(x))
        out = x.key.sum(x.data, np.array([0, x.size], np.int64)) if x.size else None
        if out is None:
            out = x.key.sum(x.key._ct_out(0), np.array([0, 0], np.int64))
        return CiphertextArray(x.key, out, ())

    def select_sum(self, x, key):
        """The sum of the elements a HeMatrix key selects, without a copy
        (extension_functions.cc:29-57): a 2-tuple selects the block of its
        row and column indices, any other key (a shorter tuple included)
        selects rows.  An empty selection sums to 0 (an encryption of 0)."""
        if not _is_ct(x):
            a = _pt(x)
            return np.sum(a.ravel()[_select_positions(a.shape, key)])
        return self.batch_select_sum(x, [key]).reshape(())

    def batch_select_sum(self, x, keys):
        """[select_sum(x, k) for k in keys] as one 1-D array (extension_functions.cc:59-87)."""
        if not _is_ct(x):
            a = _pt(x)
            return np.array([np.sum(a.ravel()[_select_positions(a.shape, k)]) for k in keys], dtype=np.int64)
        sels = [x._flat_index(k) for k in keys]
        offs = np.concatenate([[0], np.cumsum([len(s) for s in sels])]).astype(np.int64)
        idx = np.concatenate(sels).astype(np.int64) if sels else np.zeros(0, np.int64)
        if idx.size == 0:
            idx = np.zeros(1, np.int64)  # a valid pointer; every segment is empty
        out = x.key.sum_gather(x.data, idx, offs)
        return CiphertextArray(x.key, out, (len(keys),))

    def feature_wise_bucket_sum(self, x, subgroup_map, order_map, bucket_num: int, cumsum: bool = False):
        """Per feature f, bucket b and column c: the sum of x[i, c] over the rows
        i of the subgroup with order_map[i, f] == b (cumulated over b with
        cumsum); shape (bucket_num * features, x.cols)
        (numpy/evaluator.cc:378-466, extension_functions.cc:89-107)."""
        return self.batch_feature_wise_bucket_sum(x, [subgroup_map], order_map, bucket_num, cumsum)[0]

    def batch_feature_wise_bucket_sum(self, x, subgroup_maps, order_map, bucket_num: int, cumsum: bool = False):
        if not _is_ct(x):
            return [self._bucket_sum_plain(_pt(x), np.asarray(sm), np.asarray(order_map), bucket_num, cumsum)
                    for sm in subgroup_maps]
        om = np.asarray(order_map)
        if om.ndim != 2 or om.shape[0] != x.rows:
            raise ValueError("order map and x should have the same number of rows")
        cols, feats, B = x.cols, om.shape[1], int(bucket_num)
        if om.size and (om.min() < 0 or om.max() >= B):
            raise ValueError("order map entries must be in [0, bucket_num)")
        idx_parts, lens = [], []
        for sm in subgroup_maps:
            sub = np.nonzero(np.asarray(sm).ravel() > 0)[0]
            for f in range(feats):
                b_of = om[sub, f].astype(np.int64)
                order = np.argsort(b_of, kind="stable")
                grouped = sub[order]
                counts = np.bincount(b_of, minlength=B)
                starts = np.concatenate([[0], np.cumsum(counts)])
                for c in range(cols):  # segments in (group, f, c, b) order
                    elem = grouped * cols + c
                    for b in range(B):
                        idx_parts.append(elem[starts[b]:starts[b + 1]])
                        lens.append(int(counts[b]))
        offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
        idx = np.concatenate(idx_parts).astype(np.int64) if idx_parts else np.zeros(0, np.int64)
        if idx.size == 0:
            idx = np.zeros(1, np.int64)
        sums = x.key.sum_gather(x.data, idx, offs)  # (G * F * C * B, W)
        if cumsum:
            seg = np.arange(0, sums.shape[0] + 1, B, dtype=np.int64)
            sums = x.key.prefix_mul(sums, seg)
        G = len(subgroup_maps)
        # (g, f, c, b) -> (g, f, b, c): rows f * B + b, columns c
        perm = np.arange(G * feats * cols * B).reshape(G, feats, cols, B).transpose(0, 1, 3, 2).ravel()
        sums = sums.index_select(0, torch.from_numpy(perm).to(sums.device))
        per = feats * B * cols
        return [CiphertextArray(x.key, sums[g * per:(g + 1) * per], (feats * B, cols)) for g in range(G)]

    @staticmethod
    def _bucket_sum_plain(x, sm, om, B, cumsum):
        x2 = x.reshape(x.shape[0], -1)
        sub = np.nonzero(sm.ravel() > 0)[0]
        out = np.zeros((B * om.shape[1], x2.shape[1]), dtype=x2.dtype)
        for f in range(om.shape[1]):
            for i in sub:
                out[f * B + om[i, f]] += x2[i]
            if cumsum:
                out[f * B:(f + 1) * B] = np.cumsum(out[f * B:(f + 1) * B], axis=0)
        return out


class PublicKey:
    """phe.PublicKey of the ZPaillier schema: n and h_s."""

    def __init__(self, n: int, h_s: int):
        self.n, self.h_s = int(n), int(h_s)

    def serialize(self) -> bytes:
        """msgpack [n, h_s] + schema index (paillier_zahlen/public_key.h:64-76)."""
        return wire.put_public_key(self.n, self.h_s)

    @staticmethod
    def load_from(buf: bytes) -> "PublicKey":
        return PublicKey(*wire.get_public_key(bytes(buf)))

    def plaintext_bound(self) -> int:
        """n >> 1: the largest |m| an encryption takes
        (paillier_zahlen/public_key.h:51-54, public_key.cc:30, encryptor.cc:38-40)."""
        return self.n >> 1

    def __eq__(self, other) -> bool:
        return isinstance(other, PublicKey) and (self.n, self.h_s) == (other.n, other.h_s)


class SecretKey:
    """phe.SecretKey of the ZPaillier schema: p and q (lambda and mu derive
    from them)."""

    def __init__(self, p: int, q: int):
        self.p, self.q = int(p), int(q)

    @property
    def lambda_(self) -> int:
        return (self.p - 1) * (self.q - 1) // 2

    @property
    def mu(self) -> int:
        return pow(self.lambda_, -1, self.p * self.q)

    def serialize(self) -> bytes:
        """msgpack [lambda, mu, p, q] + schema index (paillier_zahlen/secret_key.h:64-80)."""
        return wire.put_secret_key(self.p, self.q)

    @staticmethod
    def load_from(buf: bytes) -> "SecretKey":
        return SecretKey(*wire.get_secret_key(bytes(buf)))

    def __eq__(self, other) -> bool:
        return isinstance(other, SecretKey) and (self.p, self.q) == (other.p, other.q)


class HeKit:
    """Keys resident on the device plus the three actors (numpy HeKit,
    bind_numpy.cc:242-293)."""

    def __init__(self, key: E.Key):
        self.key = key
        self._enc, self._dec, self._ev = Encryptor(self), Decryptor(self), Evaluator(self)

    def encryptor(self) -> Encryptor:
        return self._enc

    def decryptor(self) -> Decryptor:
        return self._dec

    def evaluator(self) -> Evaluator:
        return self._ev

    def public_key(self) -> PublicKey:
        return PublicKey(self.key.n, self.key.hs)

    def secret_key(self) -> SecretKey | None:
        return SecretKey(self.key.p, self.key.q) if self.key.p is not None else None

    def array(self, ndarray) -> np.ndarray:
        """A plaintext array: numpy int64, or Python ints where a value needs
        more than 64 bits."""
        return _pt(ndarray)


# the ZPaillier aliases of phe/base/schema.cc:47-48, plus the names a HEU
# build registers this engine's schema under
_SCHEMA_ALIASES = ("z-paillier", "zpaillier", "paillier", "paillier_z", "paillier_zahlen",
                   "hpaillier", "paillier_hip", "paillier_mi355x")


def parse_schema(name: str) -> str:
    """ParseSchemaType (phe/base/schema.cc:77-87), case-insensitive, for the one
    schema this engine runs; any other schema raises ValueError."""
    if name.lower() not in _SCHEMA_ALIASES:
        raise ValueError(f"Unknown schema type {name}: this engine runs Paillier (z-paillier) only")
    return "z-paillier"


def setup(key_bits=2048, seed: int | None = None, ctx: E.Context | None = None,
          r_bits: int | None = None, window_bits: int = 0, secret_key: SecretKey | None = None) -> HeKit:
    """Generates a key (seeded or from the OS RNG) and makes it resident; or,
    given a PublicKey (and optionally its SecretKey) as the first argument,
    makes that key resident (hnp.setup(pk[, sk]) of bind_numpy.cc).
    r_bits defaults to key_bits / 2 like the reference's GetRn (encryptor.cc:27).
    The reference's setup(schema_string[, key_size]) form is taken too:
    setup("z-paillier", 2048), key size 2048 when none is given."""
    if isinstance(key_bits, str):
        parse_schema(key_bits)
        key_bits, seed = (2048 if seed is None else int(seed)), None
    ctx = ctx or E.Context(0)
    if isinstance(key_bits, PublicKey):
        pk = key_bits
        if isinstance(seed, SecretKey):  # setup(pk, sk), the reference's order
            if secret_key is not None:
                raise TypeError("setup: two secret keys given")
            secret_key, seed = seed, None
        bits = (pk.n.bit_length() + 31) // 32 * 32
        p, q = (secret_key.p, secret_key.q) if secret_key is not None else (None, None)
        if p is not None and p * q != pk.n:
            raise ValueError("the secret key does not belong to the public key")
        key = E.Key(ctx, pk.n, pk.h_s, p, q, r_bits=r_bits or bits // 2, window_bits=window_bits)
        return HeKit(key)
    kd = E.keygen(key_bits, seed)
    key = E.Key.from_keygen(ctx, kd, r_bits=r_bits or key_bits // 2, window_bits=window_bits)
    return HeKit(key)
