"""heu_sum_noalloc through the C ABI: the same bytes as heu_sum, with no device
allocation by the engine.

Two calls in a row on one context (a 4096 x 64 column-major matrix summed
along axis 0, then a smaller ragged set of segments with an empty one), a
3072-bit key, and a call that needs more scratch than the fixed array holds,
which must fail with HEU_ERANGE before any launch and leave the context
usable.  The device's free memory is the same before and after the
no-allocation calls.
"""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from heu_amd import engine

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return engine


@pytest.fixture(scope="module")
def ctx(eng):
    return eng.Context(0)


def host(t):
    return t.cpu().numpy().view(np.uint32)


def sum_noalloc_abi(eng, key, ct, offs, out):
    """One heu_sum_noalloc call, pointers straight through the C ABI."""
    lib = eng._native.lib()
    offs = np.ascontiguousarray(offs, dtype=np.int64)
    return lib.heu_sum_noalloc(key._h, ctypes.c_void_p(ct.data_ptr()), offs.ctypes.data_as(ctypes.c_void_p),
                               len(offs) - 1, ctypes.c_void_p(out.data_ptr()))


def cts(eng, key, n, seed):
    rng = np.random.default_rng(seed)
    m = rng.integers(-(2**40), 2**40, size=n, dtype=np.int64)
    return m, key.encrypt_i64(m, eng.seeded_r(seed, 3, n, 256))


@pytest.mark.parametrize("bits", [2048, 3072])
def test_noalloc_sum_equals_sum(eng, ctx, bits):
    key = eng.Key.from_keygen(ctx, eng.keygen(bits, 60 + bits), r_bits=256, window_bits=10)
    rows, cols = 4096, 64
    m1, c1 = cts(eng, key, rows * cols, 1)
    offs1 = np.arange(0, rows * cols + 1, rows, dtype=np.int64)
    m2, c2 = cts(eng, key, 1000, 2)
    offs2 = np.array([0, 1, 1, 37, 500, 999, 1000], dtype=np.int64)
    # the ordinary path first: it loads every kernel and grows the context's pool
    want1, want2 = key.sum(c1, offs1), key.sum(c2, offs2)
    ctx.sync()
    torch.cuda.synchronize()
    out1 = torch.full((cols, key.ct_words), -1, dtype=torch.int32, device=ctx.device)
    out2 = torch.full((len(offs2) - 1, key.ct_words), -1, dtype=torch.int32, device=ctx.device)
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    print(f"[{bits}] heu_sum_noalloc: {rows}x{cols} axis-0 sum, then {len(offs2) - 1} ragged segments", flush=True)
    assert sum_noalloc_abi(eng, key, c1, offs1, out1) == 0, eng._native.lib().heu_last_error()
    assert sum_noalloc_abi(eng, key, c2, offs2, out2) == 0, eng._native.lib().heu_last_error()
    ctx.sync()
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0, "the no-allocation calls changed the device's free memory"
    assert np.array_equal(host(out1), host(want1))
    assert np.array_equal(host(out2), host(want2))
    # the sums decrypt to the plaintext sums
    assert np.array_equal(host(key.decrypt(out1, out_words=2)).view(np.int64).ravel(),
                          m1.reshape(cols, rows).sum(axis=1))
    assert np.array_equal(host(key.decrypt(out2, out_words=2)).view(np.int64).ravel(),
                          np.array([m2[a:b].sum() for a, b in zip(offs2[:-1], offs2[1:])]))
    key.close()


def test_noalloc_rejects_oversized_call(eng, ctx):
    key = eng.Key.from_keygen(ctx, eng.keygen(2048, 77), r_bits=256, window_bits=10)
    n = 200000  # one segment per ciphertext: partials and totals far beyond the fixed array
    _, c = cts(eng, key, n, 4)
    offs = np.arange(n + 1, dtype=np.int64)
    out = torch.empty((n, key.ct_words), dtype=torch.int32, device=ctx.device)
    ctx.sync()
    print("heu_sum_noalloc: oversized call", flush=True)
    rc = sum_noalloc_abi(eng, key, c, offs, out)
    assert rc == eng._native.HEU_ERANGE
    assert b"fixed buffer" in eng._native.lib().heu_last_error()
    # the context still works, and a fitting call after it is exact
    small = torch.empty((1, key.ct_words), dtype=torch.int32, device=ctx.device)
    assert sum_noalloc_abi(eng, key, c[:100], np.array([0, 100]), small) == 0
    assert np.array_equal(host(small), host(key.sum(c[:100], np.array([0, 100]))))
    ctx.sync()
    key.close()
