"""GPU parity at edge operands, both key sizes, against the oracle.

Random ciphertexts almost never hold the values that stress the lazy 64-bit
accumulators and the final canonicalisation of the Montgomery products: the
largest residues (n^2 - 1, n^2 - 2), the top bit alone (2^(bits-1)), dense
limbs (2^(bits-1) - 1, all 29-bit limbs set below the top; alternating
patterns), and the smallest ones (1, 2).  Every entry point that multiplies
ciphertexts gets them here: ct+ct, ct+pt, ct x int64, ct x scalar of any
length, decryption in both lane shapes (with values sharing a factor with n),
segment sums, dot products and prefix products.  The expected values come from
the oracle (paillier_zahlen restated over OpenSSL BIGNUM) or, where it has no
entry point, from Python's integers.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

I64_MIN, I64_MAX = -(2**63), 2**63 - 1


@pytest.fixture(scope="module")
def eng():
    from heu_amd import engine

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return engine


@pytest.fixture(scope="module")
def ctx(eng):
    return eng.Context(0)


def _keys(eng, oracle_mod, ctx, bits, seed):
    kd = eng.keygen(bits, seed)
    gk = eng.Key.from_keygen(ctx, kd, r_bits=256, window_bits=10)
    ok = oracle_mod.Key(kd["n"], kd["hs"], kd["p"], kd["q"])
    return kd, gk, ok


@pytest.fixture(scope="module")
def k2048(eng, oracle_mod, ctx):
    return _keys(eng, oracle_mod, ctx, 2048, 21)


@pytest.fixture(scope="module")
def k3072(eng, oracle_mod, ctx):
    return _keys(eng, oracle_mod, ctx, 3072, 22)


KEYS = ["k2048", "k3072"]


def host(t):
    return t.cpu().numpy().view(np.uint32)


def edge_values(kd) -> list:
    """Invertible edge residues mod n^2 (n is odd; none shares p or q)."""
    n2 = kd["n"] ** 2
    bits = n2.bit_length()
    words = (bits + 31) // 32
    alt = int.from_bytes(bytes([0xAA]) * (4 * words), "little") % n2
    alt5 = int.from_bytes(bytes([0x55]) * (4 * words), "little") % n2
    vals = [1, 2, n2 - 1, n2 - 2, 1 << (bits - 1), (1 << (bits - 1)) - 1, alt, alt5,
            n2 - (1 << (bits - 2)), (1 << (bits - 1)) + 1]
    limbs29 = (1 << (29 * ((bits - 1) // 29))) - 1  # every 29-bit limb set below the top one
    vals.append(limbs29 % n2)
    out = []
    for v in vals:  # keep the invertible ones (n^2 - 1 = -1 etc. all are)
        import math

        if 0 < v < n2 and math.gcd(v, kd["n"]) == 1 and v not in out:
            out.append(v)
    return out


def rows_of(eng, vals, words):
    return np.stack([eng.int_to_words(v, words) for v in vals])


@pytest.mark.parametrize("which", KEYS)
def test_add_edges(eng, which, request):
    kd, gk, ok = request.getfixturevalue(which)
    ev = edge_values(kd)
    W = gk.ct_words
    a = rows_of(eng, [x for x in ev for _ in ev], W)
    b = rows_of(eng, [y for _ in ev for y in ev], W)
    print(f"[{which}] ct+ct over {len(a)} edge pairs", flush=True)
    got = gk.add(a, b)
    gk.ctx.sync()
    want = ok.add(a, b)
    assert np.array_equal(host(got), want)
    n2 = kd["n"] ** 2
    assert [eng.words_to_int(r) for r in want[:5]] == [(x * y) % n2 for x, y in
                                                       zip([ev[0]] * 5, ev[:5])]


@pytest.mark.parametrize("which", KEYS)
def test_add_i64_and_mul_i64_edges(eng, which, request):
    kd, gk, ok = request.getfixturevalue(which)
    ev = edge_values(kd)
    W = gk.ct_words
    pts = [I64_MIN, I64_MIN + 1, -2, -1, 0, 1, 2, I64_MAX, -(2**31), 2**31 - 1]
    a = rows_of(eng, [x for x in ev for _ in pts], W)
    p = np.array([y for _ in ev for y in pts], dtype=np.int64)
    print(f"[{which}] ct+pt over {len(a)} edge pairs", flush=True)
    got = gk.add_i64(a, p)
    gk.ctx.sync()
    assert np.array_equal(host(got), ok.add_pt_i64(a, p))
    print(f"[{which}] ct x int64 over {len(a)} edge pairs", flush=True)
    got = gk.mul_i64(a, p)
    gk.ctx.sync()
    assert np.array_equal(host(got), ok.mul_i64(a, p))


@pytest.mark.parametrize("which", KEYS)
def test_mul_any_length_edges(eng, which, request):
    kd, gk, ok = request.getfixturevalue(which)
    ev = edge_values(kd)
    n = kd["n"]
    W = gk.ct_words
    scalars = [n // 2, -(n // 2), n - 1, (1 << 200) + 1, -((1 << 127) - 1), 3, -3, 0, 1, -1]
    a = rows_of(eng, [x for x in ev for _ in scalars], W)
    s = [y for _ in ev for y in scalars]
    print(f"[{which}] ct x scalar (up to {n.bit_length()} bits) over {len(a)} edge pairs", flush=True)
    got = gk.mul(a, s)
    gk.ctx.sync()
    assert np.array_equal(host(got), ok.mul(a, s))


@pytest.mark.parametrize("which", KEYS)
def test_decrypt_edges_both_shapes(eng, which, request):
    kd, gk, ok = request.getfixturevalue(which)
    p, q = kd["p"], kd["q"]
    n2 = kd["n"] ** 2
    # ciphertext edge values, plus values sharing the factor p or q (not
    # ciphertexts: c^(p-1) = 0 mod p^2 there, and L(0) truncates to 0)
    vals = edge_values(kd) + [p, q, p * 12345, q * (n2 // q - 1), p * q * 7, n2 - p]
    W = gk.ct_words
    ct = rows_of(eng, vals, W)
    want = ok.decrypt(ct, out_words=gk.n_words)
    shapes = (2, 4) if gk.key_bits == 2048 else (4, 8)
    before = gk.dec_lanes
    for lanes in shapes:
        gk.dec_lanes = lanes
        print(f"[{which}] decrypt {len(vals)} edge values, {lanes} lanes per bignum", flush=True)
        got = gk.decrypt(ct, out_words=gk.n_words)
        gk.ctx.sync()
        assert np.array_equal(host(got), want), f"{lanes} lanes"
    gk.dec_lanes = before
    # the plaintext of a true encryption is recovered at every edge of |m| <= n/2
    half = kd["n"] // 2
    ms = [half, -half, half - 1, -(half - 1), 0, 1, -1]
    r = eng.seeded_r(5, 5, len(ms), 256)
    c = gk.encrypt(ms, r)
    d = gk.decrypt(c, out_words=gk.n_words)
    gk.ctx.sync()
    assert eng.words_to_ints(host(d), signed=True) == ms


@pytest.mark.parametrize("which", KEYS)
def test_sum_dot_prefix_edges(eng, which, request):
    kd, gk, ok = request.getfixturevalue(which)
    ev = edge_values(kd)
    n2 = kd["n"] ** 2
    W = gk.ct_words
    # segments: one value repeated past a level-0 chunk (33 x (n^2 - 1)), each
    # edge value alone, all of them together, an empty one, dense limbs x 40
    seq = [ev[2]] * 33 + ev + ev + [ev[5]] * 40
    offs = np.array([0, 33] + [33 + i + 1 for i in range(len(ev))] + [33 + 2 * len(ev), 33 + 2 * len(ev),
                                                                      len(seq)], dtype=np.int64)
    ct = rows_of(eng, seq, W)
    print(f"[{which}] sum of {len(offs) - 1} edge segments", flush=True)
    got = gk.sum(ct, offs)
    gk.ctx.sync()
    want = ok.dot_i64(ct, None, offs)
    assert np.array_equal(host(got), want)
    for s in range(len(offs) - 1):
        prod = 1
        for v in seq[offs[s]:offs[s + 1]]:
            prod = prod * v % n2
        assert eng.words_to_int(want[s]) == prod
    w = np.array([[I64_MAX, I64_MIN + 1, -1, 1, 0, 2, -(2**31), 2**31 - 1][i % 8] for i in range(len(seq))],
                 dtype=np.int64)
    print(f"[{which}] dot over the same segments with edge weights", flush=True)
    got = gk.dot_i64(ct, w, offs)
    gk.ctx.sync()
    assert np.array_equal(host(got), ok.dot_i64(ct, w, offs))
    print(f"[{which}] prefix products over the same segments", flush=True)
    got = host(gk.prefix_mul(ct, offs))
    gk.ctx.sync()
    for s in range(len(offs) - 1):
        acc = 1
        for j in range(offs[s], offs[s + 1]):
            acc = acc * seq[j] % n2
            assert eng.words_to_int(got[j]) == acc, (s, j)
