"""CPU tests of the seeded workloads behind the batched and served
configurations and of the full-size chain check (no device needed)."""
import numpy as np
import pytest

I64_MIN, I64_MAX = -(2**63), 2**63 - 1


@pytest.fixture(scope="module")
def cfg():
    from heu_amd import configs

    return configs


def test_workloads_are_seeded_and_in_range(cfg):
    a = cfg.log_uniform_lengths(3, 0, 1, 64, 1, 65536)
    assert np.array_equal(a, cfg.log_uniform_lengths(3, 0, 1, 64, 1, 65536))
    assert not np.array_equal(a, cfg.log_uniform_lengths(3, 1, 1, 64, 1, 65536))
    assert a.min() >= 1 and a.max() <= 65536
    b = cfg.log_uniform_lengths(4, 0, 1, 4096, 256, 262144)
    assert b.min() >= 256 and b.max() <= 262144
    # log-uniform: the median sits near the geometric mean of the bounds
    assert 4096 < np.median(b) < 32768
    w = cfg.int32s(3, 0, 300, 10000)
    assert w.dtype == np.int64 and w.min() >= -(2**31) and w.max() < 2**31
    assert (w < 0).any() and (w > 0).any()
    m = cfg.int64s(3, 0, 100, 10000)
    assert np.array_equal(m, cfg.int64s(3, 0, 100, 10000))
    assert m.min() < -(2**62) and m.max() > 2**62
    r = cfg.r_limbs(3, 0, 200, 100)
    assert r.shape == (100, 8) and r.dtype == np.uint32
    vals = [int.from_bytes(row.tobytes(), "little") for row in r]
    assert all(0 < v < 2**256 for v in vals)


def test_chain_expected_packing(cfg):
    from heu_amd.engine import twos_to_int

    m = np.array([I64_MIN, I64_MAX, -1, 0, 12345, I64_MIN], dtype=np.int64)
    m2 = np.array([I64_MIN, I64_MAX, 1, 5, -12346, I64_MAX], dtype=np.int64)
    k = np.array([-(2**31), 2**31 - 1, -1, 3, 7, -1], dtype=np.int64)
    rows = cfg.chain_expected(m, m2, k)
    assert rows.shape == (6, 4) and rows.dtype == np.uint32
    want = [(-(2**64)) * -(2**31), (2**64 - 2) * (2**31 - 1), 0, 15, -7, 1]
    assert [twos_to_int(row) for row in rows] == want
    # (-2^64)(-2^31) = 2^95: bit 31 of word 2
    assert rows[0].tolist() == [0, 0, 0x80000000, 0]
