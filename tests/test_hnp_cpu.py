"""CPU checks of the HeMatrix layer's plaintext paths: the feature-wise bucket
sum restated from numpy/evaluator.cc:402-466 (a row loop per feature, then an
optional cumulative sum over buckets) against an independent numpy form, and
plaintext broadcasting / selections against numpy itself."""
import numpy as np
import pytest

pytest.importorskip("torch")


def test_plain_paths():
    from heu_amd.hnp import Evaluator

    ev = Evaluator(None)
    rng = np.random.default_rng(0)
    x = rng.integers(-1000, 1000, size=(40, 3))
    om = rng.integers(0, 5, size=(40, 4))
    sm = rng.integers(0, 2, size=40)
    for cumsum in (False, True):
        got = ev.feature_wise_bucket_sum(x, sm, om, 5, cumsum)
        want = np.zeros((20, 3), np.int64)
        for f in range(4):
            for b in range(5):
                want[f * 5 + b] = x[(sm > 0) & (om[:, f] == b)].sum(axis=0)
            if cumsum:
                want[f * 5:(f + 1) * 5] = np.cumsum(want[f * 5:(f + 1) * 5], axis=0)
        assert np.array_equal(got, want)
    assert np.array_equal(ev.add(x, x[0]), x + x[0])
    assert np.array_equal(ev.matmul(x.T, x), x.T @ x)
    assert ev.select_sum(x, (slice(None), 1)) == x[:, 1].sum()
    assert list(ev.batch_select_sum(x, [0, [1, 2]])) == [x[0].sum(), x[[1, 2]].sum()]


class _Key:
    """Stands in for an engine key: indexing only compares key identity."""


def _ids(shape, key=None):
    """A ciphertext array whose 'ciphertext' k is the one-word value k, on the
    CPU: indexing and assignment move rows of `data` without engine calls."""
    import torch
    from heu_amd.hnp import CiphertextArray

    size = int(np.prod(shape, dtype=np.int64))
    return CiphertextArray(key or _Key(), torch.arange(size, dtype=torch.int32).reshape(size, 1), shape)


def _vals(arr):
    return arr.data[:, 0].numpy().reshape(arr.shape)


def test_heu_key_rules_slice_get():
    """pylib/numpy_test.py test_slice_get: where the reference asserts equality
    with numpy, the HeMatrix keys select the same elements in the same shape;
    where it warns that they differ, they follow the reference."""
    nparr = np.arange(49).reshape(7, 7)
    harr = _ids((7, 7))
    for key in [np.s_[5:1, 1], np.s_[4:, [1, 2, 2, 3]], np.s_[:4, (5,)], np.s_[-3:-1, (1, 3)], np.s_[1:5:2, -1],
                np.s_[::2, [-7, 5]], [1, 2, 3], 1, np.s_[:]]:
        got, want = _vals(harr[key]), nparr[key]
        assert got.shape == want.shape and np.array_equal(got, want), key
    vec = np.arange(8)
    hvec = _ids((8,))
    assert harr[[1, 2, 3]].shape == (3, 7) and hvec.shape == (8,)
    for key in [[1, 2, 3], np.s_[:], 1, -1, -7]:
        got, want = _vals(hvec[key]), vec[key]
        assert got.shape == np.shape(want) and np.array_equal(got, want), key
    # the two cases the reference marks as differing from numpy
    assert _vals(harr[[1, 2], [0, -1]]).tolist() == [[7, 13], [14, 20]]
    assert _vals(hvec[(0,)]).tolist() == [0]
    # and an empty tuple selects nothing
    assert harr[()].shape == (0, 7)
    with pytest.raises(IndexError):
        harr[7]
    with pytest.raises(IndexError):
        harr[0, -8]
    with pytest.raises(IndexError):
        hvec[0, 0]
    with pytest.raises(TypeError):
        harr[[0.5]]


def test_heu_key_rules_slice_set():
    """pylib/numpy_test.py test_slice_set / test_ciphertext_slice_2d / _1d on
    ciphertext arrays: block assignment, a 1-D patch into a row or a column,
    one element, repeated indices."""
    import torch
    from heu_amd.hnp import CiphertextArray

    nparr = np.arange(49).reshape(7, 7)
    harr = _ids((7, 7))
    key = harr.key

    def patch(values):
        values = np.asarray(values)
        return CiphertextArray(key, torch.from_numpy(values.astype(np.int32).reshape(-1, 1)), values.shape)

    nppatch = np.arange(100, 116).reshape(4, 4)
    nparr[1:5, [0, 1, 5, 6]] = nppatch
    harr[1:5, [0, 1, 5, 6]] = patch(nppatch)
    assert np.array_equal(_vals(harr), nparr)
    nparr[[5, 6], 6:4:-1] = nparr[0:2, [3, 2]]
    harr[[5, 6], 6:4:-1] = harr[0:2, [3, 2]]
    assert np.array_equal(_vals(harr), nparr)
    row = np.arange(200, 207)
    nparr[0, :] = row
    harr[0, :] = patch(row)
    assert np.array_equal(_vals(harr), nparr)
    nparr[:, 5] = row
    harr[:, 5] = patch(row)
    assert np.array_equal(_vals(harr), nparr)
    nparr[1, 1] = 300
    harr[1, 1] = patch(np.array(300))
    assert np.array_equal(_vals(harr), nparr)
    nparr[:, -1] = nparr[0, :]
    harr[:, -1] = harr[0, :]
    assert np.array_equal(_vals(harr), nparr)
    nparr[-2, :] = nparr[0, :]
    harr[-2, :] = harr[0, :]
    assert np.array_equal(_vals(harr), nparr)
    # repeated target indices: the last write wins
    harr[[2, 2], 0] = patch(np.array([[400], [401]]))
    nparr[2, 0] = 401
    assert np.array_equal(_vals(harr), nparr)
    vec = np.arange(100)
    hvec = _ids((100,), key)
    vec[1:3] = vec[98:]
    hvec[1:3] = hvec[98:]
    vec[-100] = 1000
    hvec[-100] = patch(np.array([1000]))
    assert np.array_equal(_vals(hvec), vec)
    with pytest.raises(ValueError):
        harr[0:2, 0:2] = patch(np.arange(3))
    with pytest.raises(TypeError):
        harr[0, 0] = 5
    with pytest.raises(ValueError):
        harr[0, 0] = _ids((), _Key())


def test_heu_key_rules_select_sum():
    """pylib/numpy_test.py test_evaluate (select sums on plaintexts) and
    numpy_test.cc SelectSumWorks' row/column block form."""
    from heu_amd.hnp import Evaluator

    ev = Evaluator(None)
    pt1, pt2 = np.array([[1, 2], [3, 4]]), np.array([[4, 5], [6, 7]])
    assert ev.select_sum(pt1, ([0, 1], 0)) == 4
    assert ev.select_sum(pt1, [0]) == 3
    assert ev.select_sum(pt2, ()) == 0
    assert ev.select_sum(pt2, (1, [0, 1])) == 13
    assert list(ev.batch_select_sum(pt1, [(0, 0), (1, 1)])) == [1, 4]
    assert list(ev.batch_select_sum(pt2, [(), ()])) == [0, 0]
    assert list(ev.batch_select_sum(pt2, [(1, [0, 1]), ([0, 1], 0)])) == [13, 10]
    m = np.arange(900).reshape(30, 30)
    assert ev.select_sum(m, ([], [])) == 0
    assert ev.select_sum(m, ([0], [0])) == 0
    assert ev.select_sum(m, ([0, 1], [0, 1])) == 0 + 1 + 30 + 31
    mask = m[:, 0] % 60 == 0
    assert ev.select_sum(m, (mask, 1)) == m[mask, 1].sum()


def test_plaintexts_beyond_int64():
    """Plaintexts are BigInts in the reference: values past int64 stay Python
    ints through the plaintext paths; values that fit become int64."""
    from heu_amd.hnp import Evaluator, _pt

    ev = Evaluator(None)
    big = np.array([2**100, -(2**70), 5], dtype=object)
    a = _pt(big)
    assert a.dtype == object and a.tolist() == [2**100, -(2**70), 5]
    assert _pt(np.array([1, 2], dtype=object)).dtype == np.int64
    assert _pt(np.array([2**64 - 1], dtype=np.uint64)).tolist() == [2**64 - 1]
    assert ev.add(big, 1).tolist() == [2**100 + 1, -(2**70) + 1, 6]
    assert ev.mul(big, big).tolist() == [2**200, 2**140, 25]
    assert ev.sum(big) == 2**100 - 2**70 + 5
    assert ev.matmul(big, big) == 2**200 + 2**140 + 25
    x = np.array([[2**80], [1], [2**80]], dtype=object)
    got = ev.feature_wise_bucket_sum(x, np.ones(3, np.int8), np.array([[0], [1], [0]], np.int8), 2)
    assert got.tolist() == [[2**81], [1]]
    with pytest.raises(TypeError):
        _pt(np.array([1.5]))
    with pytest.raises(TypeError):
        _pt(np.array(["1"], dtype=object))


def test_schema_names():
    """ParseSchemaType's ZPaillier aliases (phe/base/schema.cc:47-48, any case)
    name this engine's schema; other schemas are refused."""
    from heu_amd.hnp import parse_schema

    for name in ("z-paillier", "ZPaillier", "paillier", "PAILLIER_ZAHLEN", "paillier_z", "hpaillier"):
        assert parse_schema(name) == "z-paillier"
    for name in ("ou", "f-paillier", "elgamal", "mock", ""):
        with pytest.raises(ValueError):
            parse_schema(name)
