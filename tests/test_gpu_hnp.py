"""GPU tests of the numpy-style HeMatrix layer (heu_amd/hnp.py): broadcasting
elementwise ops, matrix products, sums, selections and feature-wise bucket
sums, checked against numpy on the decrypted values, plus bit-exact
ciphertext closed forms of the gathered sum and dot."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kit():
    from heu_amd import engine, hnp

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return hnp.setup(2048, seed=80, ctx=engine.Context(0), r_bits=256, window_bits=10)


def test_encrypt_decrypt_broadcast_ops(kit):
    enc, dec, ev = kit.encryptor(), kit.decryptor(), kit.evaluator()
    rng = np.random.default_rng(1)
    x = rng.integers(-(2**40), 2**40, size=(6, 5), dtype=np.int64)
    p = rng.integers(-(2**20), 2**20, size=(5,), dtype=np.int64)
    col = rng.integers(-1000, 1000, size=(6, 1), dtype=np.int64)
    X = enc.encrypt(x, seed=3)
    print("hnp: encrypt/decrypt", flush=True)
    assert np.array_equal(dec.decrypt(X), x)
    assert np.array_equal(dec.decrypt(ev.add(X, X[0:1])), x + x[0:1])
    assert np.array_equal(dec.decrypt(ev.add(X, p)), x + p)
    assert np.array_equal(dec.decrypt(ev.add(p, X)), p + x)
    assert np.array_equal(dec.decrypt(ev.sub(X, p)), x - p)
    assert np.array_equal(dec.decrypt(ev.sub(p, X)), p - x)
    assert np.array_equal(dec.decrypt(ev.sub(X, X[:, 2:3])), x - x[:, 2:3])
    assert np.array_equal(dec.decrypt(ev.mul(X, col)), x * col)
    assert np.array_equal(dec.decrypt(ev.mul(col, X)), col * x)
    assert np.array_equal(dec.decrypt(X.T), x.T)
    assert np.array_equal(dec.decrypt(X[1:4, ::2]), x[1:4, ::2])
    big = np.array([np.iinfo(np.int64).min, 5], np.int64)
    two = enc.encrypt(np.array([3, 4], np.int64))
    got = dec.decrypt(ev.sub(two, big))
    assert list(got) == [3 - int(big[0]), -1]


def test_matmul(kit):
    enc, dec, ev = kit.encryptor(), kit.decryptor(), kit.evaluator()
    rng = np.random.default_rng(2)
    x = rng.integers(-(2**30), 2**30, size=(7, 5), dtype=np.int64)
    y = rng.integers(-(2**20), 2**20, size=(5, 3), dtype=np.int64)
    z = rng.integers(-(2**20), 2**20, size=(4, 7), dtype=np.int64)
    X = enc.encrypt(x)
    print("hnp: matmul", flush=True)
    assert np.array_equal(dec.decrypt(ev.matmul(X, y)), x @ y)
    assert np.array_equal(dec.decrypt(ev.matmul(z, X)), z @ x)
    v = rng.integers(-100, 100, size=(5,), dtype=np.int64)
    assert np.array_equal(dec.decrypt(ev.matmul(X, v)), x @ v)
    # bit-exact closed form of one output: prod ct^w mod n^2
    n2 = kit.key.n ** 2
    cts = [int.from_bytes(r.tobytes(), "little") for r in X.data.cpu().numpy().view(np.uint32)]
    out = ev.matmul(X, y)
    got = int.from_bytes(out.data[0].cpu().numpy().view(np.uint32).tobytes(), "little")
    want = 1
    for l in range(5):
        want = want * pow(cts[l], int(y[l, 0]), n2) % n2
    assert got == want


def test_sums_and_selections(kit):
    enc, dec, ev = kit.encryptor(), kit.decryptor(), kit.evaluator()
    rng = np.random.default_rng(3)
    x = rng.integers(-(2**40), 2**40, size=(9, 4), dtype=np.int64)
    X = enc.encrypt(x)
    print("hnp: sums", flush=True)
    assert dec.decrypt(ev.sum(X)) == x.sum()
    # HeMatrix keys: a 2-tuple of index arrays selects the block of their cross
    # product (pylib/numpy_binding/py_slicer.cc), unlike numpy's pairing
    keys = [(slice(None), 2), [1, 3, 5], x[:, 0] > 0, np.s_[2:6, 1:3], (np.array([0, 8]), np.array([3, 1])),
            np.s_[0:0], ()]
    want = [x[:, 2].sum(), x[[1, 3, 5]].sum(), x[x[:, 0] > 0].sum(), x[2:6, 1:3].sum(),
            x[0, 3] + x[0, 1] + x[8, 3] + x[8, 1], 0, 0]
    for k, w in zip(keys, want):
        assert dec.decrypt(ev.select_sum(X, k)) == w, k
    got = dec.decrypt(ev.batch_select_sum(X, keys))
    assert list(got) == want
    # bit-exact: a selection sum is the product of the selected ciphertexts mod n^2
    n2 = kit.key.n ** 2
    cts = [int.from_bytes(r.tobytes(), "little") for r in X.data.cpu().numpy().view(np.uint32)]
    sel = X._flat_index((slice(None), 2)).ravel()
    want = 1
    for i in sel:
        want = want * cts[i] % n2
    got_ct = ev.select_sum(X, (slice(None), 2)).data[0].cpu().numpy().view(np.uint32)
    assert int.from_bytes(got_ct.tobytes(), "little") == want


@pytest.mark.parametrize("cumsum", [False, True])
def test_feature_wise_bucket_sum(kit, cumsum):
    enc, dec, ev = kit.encryptor(), kit.decryptor(), kit.evaluator()
    rng = np.random.default_rng(4)
    rows, cols, feats, B = 50, 2, 3, 6
    x = rng.integers(-(2**30), 2**30, size=(rows, cols), dtype=np.int64)
    om = rng.integers(0, B, size=(rows, feats), dtype=np.int8)
    sm = (rng.random(rows) < 0.7).astype(np.int8)
    X = enc.encrypt(x)
    print(f"hnp: bucket sum cumsum={cumsum}", flush=True)
    got = dec.decrypt(ev.feature_wise_bucket_sum(X, sm, om, B, cumsum))
    want = ev.feature_wise_bucket_sum(x, sm, om, B, cumsum)
    assert got.shape == (feats * B, cols)
    assert np.array_equal(got, want)
    sms = [sm, 1 - sm, np.zeros(rows, np.int8)]
    outs = ev.batch_feature_wise_bucket_sum(X, sms, om, B, cumsum)
    for s, o in zip(sms, outs):
        assert np.array_equal(dec.decrypt(o), ev.feature_wise_bucket_sum(x, s, om, B, cumsum))
