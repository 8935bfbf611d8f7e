"""Engine form (heu_to_mform / heu_from_mform / heu_add_mform / heu_sum_mform):
ciphertexts kept on the device as R29 Montgomery limbs between calls.

Checked for both key sizes: the limbs hold c R mod n^2 (R = 2^(29 NL),
interleaved lane layout, below 2 n^2); the round trip gives the canonical
Z-space words back; engine-form ct+ct and segment sums, brought back to
Z-space, equal the Z-space entry points and the oracle (evaluator.cc:30-37,
numpy/evaluator.cc:355-373) bit for bit, at random and edge values, across
chained additions, empty segments and segments longer than a level-0 chunk.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

SHAPES = {2048: (4, 36), 3072: (8, 27)}  # lanes x limbs of the n^2 group


@pytest.fixture(scope="module")
def eng():
    from heu_amd import engine

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return engine


@pytest.fixture(scope="module")
def ctx(eng):
    return eng.Context(0)


@pytest.fixture(scope="module", params=[2048, 3072])
def keys(eng, oracle_mod, ctx, request):
    kd = eng.keygen(request.param, 70 + request.param // 1024)
    return kd, eng.Key.from_keygen(ctx, kd, r_bits=256, window_bits=10), oracle_mod.Key(kd["n"], kd["hs"])


def host(t):
    return t.cpu().numpy().view(np.uint32)


def limbs_value(row, L, K) -> int:
    """R29 interleaved words (word t*L + l holds limb l*K + t) -> integer."""
    v = 0
    for t in range(K):
        for lane in range(L):
            v += int(row[t * L + lane]) << (29 * (lane * K + t))
    return v


def cts(eng, key, kd, n, seed):
    rng = np.random.default_rng(seed)
    m = rng.integers(-(2**62), 2**62, size=n, dtype=np.int64)
    c = host(key.encrypt_i64(m, eng.seeded_r(seed, 2, n, 256))).copy()
    n2 = kd["n"] ** 2
    edges = [1, 2, n2 - 1, n2 - 2, (1 << (n2.bit_length() - 1)) - 1]
    c[: len(edges)] = np.stack([eng.int_to_words(v, key.ct_words) for v in edges])
    return c


def test_round_trip_and_limb_values(eng, keys):
    kd, key, _ = keys
    L, K = SHAPES[key.key_bits]
    assert key.mform_words == L * K
    c = cts(eng, key, kd, 300, 1)
    mf = key.to_mform(c)
    key.ctx.sync()
    n2 = kd["n"] ** 2
    R = 1 << (29 * L * K)
    got = host(mf)
    for i in range(0, 300, 37):
        v = limbs_value(got[i], L, K)
        assert v < 2 * n2 and v % n2 == eng.words_to_int(c[i]) * R % n2
    back = key.from_mform(mf)
    key.ctx.sync()
    assert np.array_equal(host(back), c)


def test_add_chain_equals_z_space_and_oracle(eng, keys):
    kd, key, ok = keys
    a, b = cts(eng, key, kd, 1000, 2), cts(eng, key, kd, 1000, 3)[::-1].copy()
    ma, mb = key.to_mform(a), key.to_mform(b)
    s1 = key.add_mform(ma, mb)
    s2 = key.add_mform(s1, ma)  # (a + b) + a, kept in engine form
    z1, z2 = host(key.from_mform(s1)), host(key.from_mform(s2))
    key.ctx.sync()
    want1 = ok.add(a, b)
    assert np.array_equal(z1, want1)
    assert np.array_equal(z1, host(key.add(a, b)))
    assert np.array_equal(z2, ok.add(want1, a))


def test_sum_equals_z_space_and_oracle(eng, keys):
    kd, key, ok = keys
    c = cts(eng, key, kd, 5000, 4)
    offs = np.array([0, 0, 1, 33, 33, 100, 4096, 5000], dtype=np.int64)
    out = key.sum_mform(key.to_mform(c), offs)
    z = host(key.from_mform(out))
    key.ctx.sync()
    want = ok.dot_i64(c, None, offs)
    assert np.array_equal(z, want)
    assert np.array_equal(z, host(key.sum(c, offs)))
    assert eng.words_to_int(z[0]) == 1 and eng.words_to_int(z[3]) == 1  # empty segments
