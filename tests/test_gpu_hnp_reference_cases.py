"""GPU restatement of the reference's HeMatrix test cases against heu_amd.hnp:
heu/library/numpy/test/numpy_test.cc (EvalWorks, EvalDimWorks, SumWorks,
SelectSumWorks, BinSumWorks, RangeCheckWorks, CtSerializeWorks),
matmul_test.cc (MatmulWorks over its nine shapes) and ic_test.cc
(CtSerializeIcWorks).  The inputs are the reference's own: GenMatrix fills a
rows x cols matrix row by row with consecutive integers from a start value
(numpy/test/test_tools.h:42-54), and the expected values are the ones those
tests assert.  The reference runs them on other schemas (OU, IcPaillier,
ZPaillier); here every ciphertext is this engine's Paillier-2048."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

INT64_MAX = np.iinfo(np.int64).max


def gen_matrix(rows, cols, start=0):
    """GenMatrix (test_tools.h:42-54): row-major consecutive integers."""
    return np.arange(start, start + rows * cols, dtype=np.int64).reshape(rows, cols)


def gen_vector(n, start=0):
    """GenVector (test_tools.h:56-68)."""
    return np.arange(start, start + n, dtype=np.int64)


def same_cts(a, b):
    return a.shape == b.shape and torch.equal(a.data.cpu(), b.data.cpu())


@pytest.fixture(scope="module")
def kit():
    from heu_amd import engine, hnp

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return hnp.setup(2048, seed=91, ctx=engine.Context(0), r_bits=256, window_bits=10)


def test_eval_works(kit):
    """numpy_test.cc:48-93: add, sub and mul with every ct/pt operand order."""
    enc, dec, ev = kit.encryptor(), kit.decryptor(), kit.evaluator()
    pts1, pts2 = gen_matrix(30, 10), gen_matrix(30, 10)
    cts1, cts2 = enc.encrypt(pts1), enc.encrypt(pts2)
    print("ref cases: EvalWorks", flush=True)
    pts_sum = ev.add(pts1, pts2)
    assert np.array_equal(pts_sum, pts1 + pts2)
    for got in (ev.add(cts1, cts2), ev.add(cts1, pts2), ev.add(pts1, cts2)):
        assert np.array_equal(dec.decrypt(got), pts_sum)
    pts_minus = ev.sub(pts1, pts2)
    assert np.array_equal(pts_minus, pts1 - pts2)
    for got in (ev.sub(cts1, cts2), ev.sub(cts1, pts2), ev.sub(pts1, cts2)):
        assert np.array_equal(dec.decrypt(got), pts_minus)
    pts_mul = ev.mul(pts1, pts2)
    assert np.array_equal(pts_mul, pts1 * pts2)
    for got in (ev.mul(cts1, pts2), ev.mul(pts1, cts2)):
        assert np.array_equal(dec.decrypt(got), pts_mul)


def test_eval_dim_works(kit):
    """numpy_test.cc:95-109: a sum keeps its operands' dimension count."""
    enc, ev = kit.encryptor(), kit.evaluator()
    print("ref cases: EvalDimWorks", flush=True)
    for a, nd in ((gen_matrix(2, 2), 2), (gen_matrix(1, 1), 2), (gen_vector(10), 1)):
        assert np.ndim(ev.add(a, a)) == nd
        assert ev.add(enc.encrypt(a), enc.encrypt(a)).ndim == nd
        assert ev.add(enc.encrypt(a), a).ndim == nd


def test_sum_works(kit):
    """numpy_test.cc:111-120."""
    enc, dec, ev = kit.encryptor(), kit.decryptor(), kit.evaluator()
    m = gen_matrix(30, 30)
    print("ref cases: SumWorks", flush=True)
    assert ev.sum(m) == 899 * 900 // 2
    assert int(dec.decrypt(ev.sum(enc.encrypt(m)))) == 899 * 900 // 2


def test_select_sum_works(kit):
    """numpy_test.cc:122-165: SelectSum(m, rows, cols) sums the submatrix
    m(rows, cols) (the pylib key (rows, cols)), with the scql case read back
    from its serialized form."""
    from heu_amd import hnp

    enc, dec, ev = kit.encryptor(), kit.decryptor(), kit.evaluator()
    m = gen_matrix(30, 30)
    print("ref cases: SelectSumWorks", flush=True)
    assert ev.select_sum(m, ([], [])) == 0
    assert ev.select_sum(m, ([0], [0])) == 0
    m2 = enc.encrypt(m)
    assert int(dec.decrypt(ev.select_sum(m2, ([0, 1], [0, 1])))) == 0 + 1 + 30 + 31
    assert int(dec.decrypt(ev.select_sum(m2, ([], [])))) == 0
    plain_in = np.array([[0], [0], [1000000], [1000000], [0]], np.int64)
    encrypted_in = enc.encrypt(plain_in)
    encrypted_in2 = hnp.CiphertextArray.load_from(kit, encrypted_in.serialize())
    assert encrypted_in2.rows == encrypted_in.rows
    assert encrypted_in2.cols == encrypted_in.cols
    assert encrypted_in2.ndim == encrypted_in.ndim
    assert int(dec.decrypt(ev.select_sum(encrypted_in2, ([0, 1], [0])))) == 0
    assert int(dec.decrypt(ev.select_sum(encrypted_in2, ([0, 1, 2], [0])))) == 1000000


def test_bin_sum_works(kit):
    """numpy_test.cc:167-183: FeatureWiseBucketSum of rows 2..4 of a 5 x 2
    GenMatrix under the test's order map; element (2, 0) is 4."""
    enc, dec, ev = kit.encryptor(), kit.decryptor(), kit.evaluator()
    m = gen_matrix(5, 2)
    sub = [2, 3, 4]
    order_map = np.array([[1, 3, 2, 2, 3, 3, 0, 0, 0, 0],
                          [2, 3, 0, 1, 0, 1, 2, 3, 0, 1],
                          [2, 1, 1, 1, 0, 2, 0, 1, 1, 1],
                          [3, 2, 2, 0, 0, 1, 1, 1, 2, 3],
                          [0, 0, 2, 2, 3, 3, 0, 3, 2, 2]], np.int8)
    every_row = np.ones(len(sub), np.int8)
    print("ref cases: BinSumWorks", flush=True)
    plain = ev.feature_wise_bucket_sum(m[sub], every_row, order_map[sub], 5)
    assert plain[2, 0] == 4
    got = dec.decrypt(ev.feature_wise_bucket_sum(enc.encrypt(m[sub]), every_row, order_map[sub], 5))
    assert got[2, 0] == 4
    assert np.array_equal(got, plain)


def test_range_check_works(kit):
    """numpy_test.cc:185-209: DecryptInRange throws once one plaintext needs
    more than range_bits bits; Decrypt still succeeds."""
    enc, dec, ev = kit.encryptor(), kit.decryptor(), kit.evaluator()
    pmatrix = gen_matrix(25, 25)
    cmatrix = enc.encrypt(pmatrix)
    print("ref cases: RangeCheckWorks", flush=True)
    assert np.array_equal(dec.decrypt_in_range(cmatrix, 64), pmatrix)
    scale = np.ones((25, 25), np.int64)
    scale[3, 3] = INT64_MAX
    cmatrix = ev.mul(cmatrix, scale)
    with pytest.raises(RuntimeError):
        dec.decrypt_in_range(cmatrix, 64)
    got = dec.decrypt(cmatrix)
    assert int(got[3, 3]) == int(pmatrix[3, 3]) * int(INT64_MAX)
    assert int(got[0, 1]) == 1
    scale = np.ones((25, 25), np.int64)
    scale[23, 23] = INT64_MAX
    cmatrix = ev.mul(cmatrix, scale)
    with pytest.raises(RuntimeError):
        dec.decrypt_in_range(cmatrix, 64)
    assert int(dec.decrypt(cmatrix)[23, 23]) == int(pmatrix[23, 23]) * int(INT64_MAX)
    with pytest.raises(RuntimeError):
        dec.decrypt_in_range(cmatrix, 1)
    dec.decrypt(cmatrix)


@pytest.mark.parametrize("n,k,m", [(1, 1, 1), (1, 1, 20), (10, 1, 1), (1, 100, 1), (10, 1, 20), (20, 30, 1),
                                   (1, 100, 3), (2, 3, 4), (10, 30, 20)])
def test_matmul_works(kit, n, k, m):
    """matmul_test.cc:35-73: in1 = GenMatrix(n, k, 10), in2 = GenMatrix(k, m, 5);
    ct @ pt and pt @ ct decrypt to in1 @ in2."""
    enc, dec, ev = kit.encryptor(), kit.decryptor(), kit.evaluator()
    in1, in2 = gen_matrix(n, k, 10), gen_matrix(k, m, 5)
    ans = in1 @ in2
    print(f"ref cases: MatmulWorks ({n}, {k}, {m})", flush=True)
    assert np.array_equal(ev.matmul(in1, in2), ans)
    assert np.array_equal(dec.decrypt(ev.matmul(enc.encrypt(in1), in2)), ans)
    assert np.array_equal(dec.decrypt(ev.matmul(in1, enc.encrypt(in2))), ans)


def test_ct_serialize_works(kit):
    """numpy_test.cc:38-46: a 10 x 30 ciphertext matrix survives Serialize /
    LoadFrom bit for bit."""
    from heu_amd import hnp

    cts1 = kit.encryptor().encrypt(gen_matrix(10, 30))
    print("ref cases: CtSerializeWorks", flush=True)
    cts2 = hnp.CiphertextArray.load_from(kit, cts1.serialize())
    assert same_cts(cts1, cts2)


def test_ct_serialize_ic_works(kit):
    """ic_test.cc:21-46: a 100 x 300 ciphertext matrix in the interconnection
    format loads back bit for bit and decrypts to the plaintexts."""
    from heu_amd import hnp

    fmt = hnp.MatrixSerializeFormat.Interconnection
    pts1 = gen_matrix(100, 300)
    cts1 = kit.encryptor().encrypt(pts1)
    print("ref cases: CtSerializeIcWorks", flush=True)
    buf = cts1.serialize(fmt)
    assert len(buf) > 0
    cts2 = hnp.CiphertextArray.load_from(kit, buf, fmt)
    assert same_cts(cts1, cts2)
    assert np.array_equal(kit.decryptor().decrypt(cts2), pts1)
