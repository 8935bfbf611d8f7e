"""GPU restatement of the reference's Python HeMatrix tests
(heu/pylib/numpy_test.py) against heu_amd.hnp: test_evaluate,
test_encrypt_with_audit, test_matmul, test_evaluate_parallel,
test_ciphertext_slice_2d / _1d, test_batch_select_sum (at its size: a
1,000,000 x 2 matrix and 50,000 selections of 150 rows),
test_feature_wise_bucket_sum and test_batch_feature_wise_bucket_sum; and of
heu/pylib/phe_test.py's Paillier cases (test_enc_dec with the largest
plaintext, test_add_inplace, key and ciphertext serialization) on 0-d arrays.
The inputs and expected values are the reference's; its random inputs are drawn
here from seeded generators.  Plaintext matrices are numpy arrays and phe
scalars 0-d arrays.  Where the reference checks a plaintext-only path (matmul,
batch_select_sum), the same inputs also run encrypted and decrypt to the same
values."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kit():
    from heu_amd import engine, hnp

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return hnp.setup(2048, seed=92, ctx=engine.Context(0), r_bits=256, window_bits=10)


def test_evaluate(kit):
    """numpy_test.py:141-237."""
    enc, dec, ev = kit.encryptor(), kit.decryptor(), kit.evaluator()
    pt1, pt2 = np.array([[1, 2], [3, 4]]), np.array([[4, 5], [6, 7]])
    ct1, ct2 = enc.encrypt(pt1), enc.encrypt(pt2)
    print("pylib cases: test_evaluate", flush=True)
    assert (ct2.rows, ct2.cols, ct2.shape) == (2, 2, (2, 2))
    ans = np.array([[5, 7], [9, 11]])
    assert np.array_equal(ev.add(pt1, pt2), ans)
    for got in (ev.add(ct1, pt2), ev.add(pt1, ct2), ev.add(ct1, ct2)):
        assert np.array_equal(dec.decrypt(got), ans)
    ans = np.array([[-3, -3], [-3, -3]])
    assert np.array_equal(ev.sub(pt1, pt2), ans)
    for got in (ev.sub(ct1, pt2), ev.sub(pt1, ct2), ev.sub(ct1, ct2)):
        assert np.array_equal(dec.decrypt(got), ans)
    ans = np.array([[4, 10], [18, 28]])
    assert np.array_equal(ev.mul(pt1, pt2), ans)
    for got in (ev.mul(ct1, pt2), ev.mul(pt1, ct2)):
        assert np.array_equal(dec.decrypt(got), ans)
    ans = np.array([[16, 19], [36, 43]])
    assert np.array_equal(ev.matmul(pt1, pt2), ans)
    for got in (ev.matmul(ct1, pt2), ev.matmul(pt1, ct2)):
        assert np.array_equal(dec.decrypt(got), ans)
    assert ev.sum(pt1) == 10 and int(dec.decrypt(ev.sum(ct1))) == 10
    assert ev.sum(pt2) == 22 and int(dec.decrypt(ev.sum(ct2))) == 22
    # select sums: (rows, cols) keys select blocks, other keys rows
    assert ev.select_sum(pt1, ([0, 1], 0)) == 4
    assert int(dec.decrypt(ev.select_sum(ct1, [0]))) == 3
    assert ev.select_sum(pt2, ()) == 0
    assert int(dec.decrypt(ev.select_sum(ct2, (1, [0, 1])))) == 13
    assert list(ev.batch_select_sum(pt1, [(0, 0), (1, 1)])) == [1, 4]
    assert list(dec.decrypt(ev.batch_select_sum(ct1, [0]))) == [3]
    assert list(ev.batch_select_sum(pt2, [(), ()])) == [0, 0]
    assert list(dec.decrypt(ev.batch_select_sum(ct2, [(1, [0, 1]), ([0, 1], 0)]))) == [13, 10]


def test_encrypt_with_audit(kit):
    """numpy_test.py:130-139: the audit strings come back in the plaintexts'
    shape."""
    pt1 = np.array([[1], [3]])
    ct1, audit = kit.encryptor().encrypt_with_audit(pt1)
    print("pylib cases: test_encrypt_with_audit", flush=True)
    assert np.array_equal(kit.decryptor().decrypt(ct1), pt1)
    assert audit.ndim == 2 and audit.shape == (2, 1)
    assert len(audit[0, 0]) > 1


def test_matmul(kit):
    """numpy_test.py:239-251: a 64-vector times a 64 x 256 matrix, the
    transposed product and a vector dot product."""
    enc, dec, ev = kit.encryptor(), kit.decryptor(), kit.evaluator()
    rng = np.random.default_rng(239)
    nparr1 = rng.integers(-10000, 10000, (64,))
    nparr2 = rng.integers(-10000, 10000, (64, 256))
    print("pylib cases: test_matmul", flush=True)
    assert np.array_equal(ev.matmul(nparr1, nparr2), nparr1 @ nparr2)
    c1 = enc.encrypt(nparr1)
    assert np.array_equal(dec.decrypt(ev.matmul(c1, nparr2)), nparr1 @ nparr2)
    assert np.array_equal(dec.decrypt(ev.matmul(nparr1, enc.encrypt(nparr2))), nparr1 @ nparr2)
    assert np.array_equal(dec.decrypt(ev.matmul(nparr2.T, c1)), nparr2.T @ nparr1)
    assert np.array_equal(dec.decrypt(ev.matmul(enc.encrypt(nparr2).T, nparr1)), nparr2.T @ nparr1)
    nparr3 = rng.integers(-10000, 10000, (64,))
    assert int(dec.decrypt(ev.matmul(c1, nparr3))) == int(nparr1 @ nparr3)


def test_evaluate_parallel(kit):
    """numpy_test.py:253-297: 100 x 100 matrices, every operand order."""
    enc, dec, ev = kit.encryptor(), kit.decryptor(), kit.evaluator()
    rng = np.random.default_rng(253)
    nparr1 = rng.integers(-10000, 10000, (100, 100))
    nparr2 = rng.integers(-10000, 10000, (100, 100))
    harr2 = enc.encrypt(nparr2)
    print("pylib cases: test_evaluate_parallel", flush=True)
    assert np.array_equal(ev.matmul(nparr1, nparr1), nparr1 @ nparr1)
    assert np.array_equal(dec.decrypt(ev.add(harr2, nparr1)), nparr2 + nparr1)
    assert np.array_equal(dec.decrypt(ev.sub(harr2, nparr1)), nparr2 - nparr1)
    assert np.array_equal(dec.decrypt(ev.mul(harr2, nparr1)), nparr2 * nparr1)
    assert np.array_equal(dec.decrypt(ev.matmul(harr2, nparr1)), nparr2 @ nparr1)
    assert int(dec.decrypt(ev.sum(harr2))) == int(nparr2.sum())
    assert np.array_equal(dec.decrypt(ev.add(nparr1, harr2)), nparr1 + nparr2)
    assert np.array_equal(dec.decrypt(ev.sub(nparr1, harr2)), nparr1 - nparr2)
    assert np.array_equal(dec.decrypt(ev.mul(nparr1, harr2)), nparr1 * nparr2)
    assert np.array_equal(dec.decrypt(ev.matmul(nparr1, harr2)), nparr1 @ nparr2)
    assert np.array_equal(dec.decrypt(ev.add(harr2, harr2)), nparr2 + nparr2)
    assert np.array_equal(dec.decrypt(ev.sub(harr2, harr2)), nparr2 - nparr2)


def test_ciphertext_slice_2d(kit):
    """numpy_test.py:414-438: one element, a block, a row into a column and a
    row into a row, assigned on the device."""
    enc, dec = kit.encryptor(), kit.decryptor()
    nparr = np.arange(49).reshape((7, 7))
    carr = enc.encrypt(nparr)
    print("pylib cases: test_ciphertext_slice_2d", flush=True)
    carr[6, 6] = enc.encrypt(np.array(666))
    carr[6, 6] = enc.encrypt(np.array(1000))
    nparr[6, 6] = 1000
    carr[[5, 6], 6:4:-1] = carr[0:2, [3, 2]]
    nparr[[5, 6], 6:4:-1] = nparr[0:2, [3, 2]]
    assert np.array_equal(dec.decrypt(carr), nparr)
    carr[:, -1] = carr[0, :]
    nparr[:, -1] = nparr[0, :]
    assert np.array_equal(dec.decrypt(carr), nparr)
    carr[-2, :] = carr[0, :]
    nparr[-2, :] = nparr[0, :]
    assert np.array_equal(dec.decrypt(carr), nparr)


def test_ciphertext_slice_1d(kit):
    """numpy_test.py:440-451."""
    enc, dec = kit.encryptor(), kit.decryptor()
    nparr = np.arange(100)
    carr = enc.encrypt(nparr)
    print("pylib cases: test_ciphertext_slice_1d", flush=True)
    carr[-100] = enc.encrypt(np.array([1000]))
    nparr[-100] = 1000
    carr[1:3] = carr[98:]
    nparr[1:3] = nparr[98:]
    assert np.array_equal(dec.decrypt(carr), nparr)


def test_batch_select_sum(kit):
    """numpy_test.py:469-484 at its size: 50,000 selections of 150 random rows
    of a 1,000,000 x 2 matrix of values in [-100, 100), encrypted."""
    enc, dec, ev = kit.encryptor(), kit.decryptor(), kit.evaluator()
    rng = np.random.default_rng(469)
    sample_size = 100 * 10000
    m1 = rng.integers(-100, 100, (sample_size, 2))
    select_indices = [rng.integers(0, sample_size, 150).tolist() for _ in range(50000)]
    print("pylib cases: test_batch_select_sum (1,000,000 x 2, 50,000 selections)", flush=True)
    c1 = enc.encrypt(m1)
    got = dec.decrypt(ev.batch_select_sum(c1, select_indices))
    want = np.array([m1[a].sum() for a in select_indices])
    assert got.shape == (50000,) and np.array_equal(got, want)
    assert np.array_equal(ev.batch_select_sum(m1, select_indices[:1000]), want[:1000])


def test_feature_wise_bucket_sum(kit):
    """numpy_test.py:486-516: 450 rows, 10 features, the first 50 rows in the
    subgroup and in bucket 0."""
    enc, dec, ev = kit.encryptor(), kit.decryptor(), kit.evaluator()
    rng = np.random.default_rng(486)
    sample_size, feature_size = 450, 10
    m1 = rng.integers(-100, 100, (sample_size, 2))
    first_group = np.zeros(sample_size, dtype=np.int8)
    first_group[:50] = 1
    order_map = np.zeros((sample_size, feature_size), dtype=np.int8)
    order_map[50:, :] = 1
    print("pylib cases: test_feature_wise_bucket_sum", flush=True)
    first = dec.decrypt(ev.feature_wise_bucket_sum(enc.encrypt(m1), first_group, order_map, 2, False))
    assert first.shape == (2 * feature_size, 2)
    assert first[0, 0] == m1[:50, 0].sum() and first[0, 1] == m1[:50, 1].sum()
    assert first[1, 0] == 0 and first[1, 1] == 0


def test_batch_feature_wise_bucket_sum(kit):
    """numpy_test.py:518-559: two subgroups (rows 0-49, 50-99), two features,
    rows 50-99 in bucket 1."""
    enc, dec, ev = kit.encryptor(), kit.decryptor(), kit.evaluator()
    rng = np.random.default_rng(518)
    sample_size = 100
    m1 = rng.integers(-100, 100, (sample_size, 2))
    first_group = np.zeros(sample_size, dtype=np.int8)
    second_group = np.zeros(sample_size, dtype=np.int8)
    first_group[:50] = 1
    second_group[50:] = 1
    order_map = np.zeros((sample_size, 2), dtype=np.int8)
    order_map[50:, :] = 1
    print("pylib cases: test_batch_feature_wise_bucket_sum", flush=True)
    sums = ev.batch_feature_wise_bucket_sum(enc.encrypt(m1), [first_group, second_group], order_map, 2, False)
    first, second = dec.decrypt(sums[0]), dec.decrypt(sums[1])
    assert first[0, 0] == m1[:50, 0].sum() and first[0, 1] == m1[:50, 1].sum()
    assert first[1, 0] == 0 and first[1, 1] == 0
    assert second[0, 0] == 0 and second[0, 1] == 0
    assert second[1, 0] == m1[50:, 0].sum() and second[1, 1] == m1[50:, 1].sum()


def test_phe_enc_dec(kit):
    """phe_test.py:110-135 on 0-d arrays: ct+ct, ct-ct, ct+pt, ct-pt, and the
    largest plaintext, n/2 - 1 (far past int128), minus 1."""
    enc, dec, ev = kit.encryptor(), kit.decryptor(), kit.evaluator()
    ct1, ct2 = enc.encrypt(np.array(123)), enc.encrypt(np.array(456))
    print("pylib cases: phe test_enc_dec", flush=True)
    assert int(dec.decrypt(ev.add(ct1, ct2))) == 123 + 456
    assert int(dec.decrypt(ev.sub(ct1, ct2))) == 123 - 456
    assert int(dec.decrypt(ev.sub(0, ct1))) == -123
    assert int(dec.decrypt(ev.add(ct1, 100))) == 123 + 100
    assert int(dec.decrypt(ev.sub(ct1, 100))) == 123 - 100
    max_p = kit.public_key().plaintext_bound() - 1
    assert max_p > 2**128
    max_ct = ev.sub(enc.encrypt(np.array(max_p, dtype=object)), 1)
    assert int(dec.decrypt(max_ct)) == max_p - 1
    bound = kit.public_key().plaintext_bound()
    assert int(dec.decrypt(enc.encrypt(np.array(-bound, dtype=object)))) == -bound
    from heu_amd._native import HeuError

    with pytest.raises(HeuError):
        enc.encrypt(np.array(bound + 1, dtype=object))


def test_phe_add_inplace(kit):
    """phe_test.py:137-144: a running ciphertext sum of 0..99, plus 100."""
    enc, dec, ev = kit.encryptor(), kit.decryptor(), kit.evaluator()
    print("pylib cases: phe test_add_inplace", flush=True)
    total = enc.encrypt(np.array(0))
    for i in range(100):
        total = ev.add(total, enc.encrypt(np.array(i)))
    total = ev.add(total, 100)
    assert int(dec.decrypt(total)) == 5050


def test_phe_key_serialize(kit):
    """phe_test.py:256-296: keys and ciphertexts leave through their wire forms
    and a second kit built from the keys decrypts the first kit's ciphertexts."""
    from heu_amd import engine, hnp

    print("pylib cases: phe key serialize", flush=True)
    pk2 = hnp.PublicKey.load_from(kit.public_key().serialize())
    sk2 = hnp.SecretKey.load_from(kit.secret_key().serialize())
    assert pk2 == kit.public_key() and sk2 == kit.secret_key()
    kit2 = hnp.setup(pk2, sk2, ctx=engine.Context(0))
    ct1 = kit.encryptor().encrypt(np.array(123))
    ct2 = hnp.CiphertextArray.load_from(kit2, ct1.serialize())
    assert int(kit2.decryptor().decrypt(ct2)) == 123
    d = kit2.evaluator().sub(ct2, kit2.encryptor().encrypt(np.array(456)))
    assert int(kit.decryptor().decrypt(hnp.CiphertextArray.load_from(kit, d.serialize()))) == 123 - 456


def test_plaintexts_beyond_int64(kit):
    """Plaintexts and scalars past int64 (BigInts in the reference) through
    encrypt, ct+pt, ct-pt, ct x pt and matmul."""
    enc, dec, ev = kit.encryptor(), kit.decryptor(), kit.evaluator()
    rng = np.random.default_rng(7)
    big = np.array([int(v) << 90 for v in rng.integers(-(2**20), 2**20, 12)], dtype=object).reshape(3, 4)
    small = rng.integers(-1000, 1000, (3, 4))
    print("pylib cases: plaintexts beyond int64", flush=True)
    cb = enc.encrypt(big)
    assert dec.decrypt(cb).tolist() == big.tolist()
    cs = enc.encrypt(small)
    assert dec.decrypt(ev.add(cs, big)).tolist() == (small.astype(object) + big).tolist()
    assert dec.decrypt(ev.sub(cs, big)).tolist() == (small.astype(object) - big).tolist()
    assert dec.decrypt(ev.sub(big, cs)).tolist() == (big - small.astype(object)).tolist()
    scal = np.array([[1 << 100, -(1 << 70), 3, -1]], dtype=object)
    assert dec.decrypt(ev.mul(cs, scal)).tolist() == (small.astype(object) * scal).tolist()
    w = np.array([1 << 80, -(1 << 66), 7, 0], dtype=object)
    assert dec.decrypt(ev.matmul(cs, w)).tolist() == (small.astype(object) @ w).tolist()
    with_min = np.array([np.iinfo(np.int64).min, 5], np.int64)
    two = enc.encrypt(np.array([3, 4]))
    assert dec.decrypt(ev.sub(two, with_min)).tolist() == [3 - int(with_min[0]), -1]
