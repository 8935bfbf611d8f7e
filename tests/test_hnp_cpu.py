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
