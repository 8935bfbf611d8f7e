"""GPU parity: every engine entry point against the CPU oracle (a restatement of
paillier_zahlen over OpenSSL BIGNUM) on the same seeded keys and inputs.

Bit-exact comparisons of Z-space ciphertexts and of decrypted plaintexts.
Mirrors the reference's own cases: paillier_zahlen/paillier_test.cc:40-233
(values {-12345, 123, +-1, +-2, 0, INT64 extremes}, Mul by {2, 1, 0, -1, -2}),
paillier_gpu/paillier_gpu_test.cc:84-165 (EncDec / Add with int64 extremes).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

I64_MIN, I64_MAX = -(2**63), 2**63 - 1
EDGE_I64 = [I64_MIN, I64_MIN + 1, -12345, -1234, -2, -1, 0, 1, 2, 123, 1234, I64_MAX]


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
    return _keys(eng, oracle_mod, ctx, 2048, 0)


@pytest.fixture(scope="module")
def k3072(eng, oracle_mod, ctx):
    return _keys(eng, oracle_mod, ctx, 3072, 1)


def host(t):
    return t.cpu().numpy().view(np.uint32)


def inputs(eng, n, seed=7):
    rng = np.random.default_rng(seed)
    m = rng.integers(I64_MIN, I64_MAX, size=n, dtype=np.int64, endpoint=True)
    m[: len(EDGE_I64)] = EDGE_I64[:n]
    r = eng.seeded_r(seed, 11, n, 256)
    return m, r


@pytest.mark.parametrize("which", ["k2048", "k3072"])
def test_encrypt_decrypt_i64(eng, which, request):
    kd, gk, ok = request.getfixturevalue(which)
    m, r = inputs(eng, 96)
    print(f"[{which}] encrypt_i64 n=96", flush=True)
    ct = gk.encrypt_i64(m, r)
    gk.ctx.sync()
    ref = ok.encrypt_i64(m, r)
    assert np.array_equal(host(ct), ref)
    print(f"[{which}] decrypt n=96", flush=True)
    pt = gk.decrypt(ct, out_words=2)
    gk.ctx.sync()
    assert np.array_equal(host(pt).view(np.int64).ravel(), m)
    full = gk.decrypt(ct)
    gk.ctx.sync()
    assert np.array_equal(host(full), ok.decrypt(ref))


@pytest.mark.parametrize("which", ["k2048", "k3072"])
def test_add_and_add_pt(eng, which, request):
    kd, gk, ok = request.getfixturevalue(which)
    m, r = inputs(eng, 64, 3)
    m2, r2 = inputs(eng, 64, 4)
    a, b = ok.encrypt_i64(m, r), ok.encrypt_i64(m2, r2)
    print(f"[{which}] add n=64", flush=True)
    s = gk.add(a, b)
    gk.ctx.sync()
    assert np.array_equal(host(s), ok.add(a, b))
    dec = ok.decrypt(host(s), out_words=3)
    want = [int(x) + int(y) for x, y in zip(m, m2)]
    assert [eng.twos_to_int(row) for row in dec] == want
    print(f"[{which}] add_i64 n=64", flush=True)
    s2 = gk.add_i64(a, m2)
    gk.ctx.sync()
    assert np.array_equal(host(s2), ok.add_pt_i64(a, m2))


@pytest.mark.parametrize("which", ["k2048", "k3072"])
def test_mul_scalar(eng, which, request):
    kd, gk, ok = request.getfixturevalue(which)
    m, r = inputs(eng, 48, 5)
    a = ok.encrypt_i64(m, r)
    rng = np.random.default_rng(9)
    k = rng.integers(-(2**31), 2**31 - 1, size=48, dtype=np.int64)
    k[:11] = [0, 1, -1, 2, -2, 3, I64_MAX, I64_MIN + 1, I64_MIN, 7, -8]  # I64_MIN: |k| = 2^63, the top window
    print(f"[{which}] mul_i64 n=48", flush=True)
    out = gk.mul_i64(a, k)
    gk.ctx.sync()
    assert np.array_equal(host(out), ok.mul_i64(a, k))


@pytest.mark.parametrize("which", ["k2048", "k3072"])
def test_mul_big_scalar(eng, which, request):
    # Mul(ct, pt) with plaintext scalars of any length (evaluator.cc:99-123):
    # special cases 0 / 1 / -1, int64 edges, and scalars up to n/2.
    kd, gk, ok = request.getfixturevalue(which)
    m, r = inputs(eng, 40, 12)
    a = ok.encrypt_i64(m, r)
    rng = np.random.default_rng(13)
    half = kd["n"] >> 1
    ks = [0, 1, -1, 2, -2, I64_MAX, I64_MIN, 2**64, -(2**64) - 1, half, -half, 2**100 + 3]
    while len(ks) < 40:
        bits = int(rng.integers(1, kd["n"].bit_length() - 1))
        v = int.from_bytes(rng.bytes((bits + 7) // 8), "little") & ((1 << bits) - 1)
        ks.append(-v if rng.integers(2) else v)
    print(f"[{which}] mul (big scalars) n=40", flush=True)
    out = gk.mul(a, ks)
    gk.ctx.sync()
    assert np.array_equal(host(out), ok.mul(a, ks))
    print(f"[{which}] mul (short scalars == mul_i64) n=40", flush=True)
    small = rng.integers(-(2**40), 2**40, size=40, dtype=np.int64)
    small[:3] = [0, 1, -1]
    out2 = gk.mul(a, small.tolist())
    gk.ctx.sync()
    assert np.array_equal(host(out2), ok.mul_i64(a, small))


@pytest.mark.parametrize("which", ["k2048", "k3072"])
def test_negate_sub(eng, which, request):
    kd, gk, ok = request.getfixturevalue(which)
    m, r = inputs(eng, 40, 6)
    m2, r2 = inputs(eng, 40, 8)
    a, b = ok.encrypt_i64(m, r), ok.encrypt_i64(m2, r2)
    print(f"[{which}] negate n=40", flush=True)
    ng = gk.negate(a)
    gk.ctx.sync()
    assert np.array_equal(host(ng), ok.negate(a))
    print(f"[{which}] sub n=40", flush=True)
    sb = gk.sub(a, b)
    gk.ctx.sync()
    assert np.array_equal(host(sb), ok.sub(a, b))


@pytest.mark.parametrize("which", ["k2048", "k3072"])
def test_inverse_edge_values(eng, which, request):
    # Negate inverts mod n^2 (evaluator.cc:90-93): edge values through the
    # single-inversion kernel (one element: its own batch) and as one batch,
    # against the oracle (OpenSSL BN_mod_inverse) and the closed form; then a
    # value sharing the factor p with n^2 must be reported as not invertible
    kd, gk, ok = request.getfixturevalue(which)
    n2 = kd["n"] ** 2
    W = gk.ct_words
    rng = np.random.default_rng(31)
    vals = [1, 2, 3, n2 - 1, n2 - 2, 1 << (n2.bit_length() - 2), (n2 - 1) // 2,
            int(rng.integers(1, 2**62)), int.from_bytes(rng.bytes(8 * W), "little") % n2]
    rows = np.stack([eng.int_to_words(v, W) for v in vals])
    want = ok.negate(rows)
    assert [eng.words_to_int(r) for r in want] == [pow(v, -1, n2) for v in vals]
    for i, v in enumerate(vals):
        print(f"[{which}] inverse of edge value {i}", flush=True)
        got = gk.negate(rows[i : i + 1])
        gk.ctx.sync()
        assert np.array_equal(host(got)[0], want[i]), i
    print(f"[{which}] inverse of all edge values", flush=True)
    got = gk.negate(rows)
    gk.ctx.sync()
    assert np.array_equal(host(got), want)
    from heu_amd._native import HeuError

    print(f"[{which}] inverse of a multiple of p", flush=True)
    with pytest.raises(HeuError):
        gk.negate(np.stack([eng.int_to_words(kd["p"] * 12345, W)]))
        gk.ctx.sync()
    gk.ctx.sync()  # the error was reported once and cleared


@pytest.mark.parametrize("which", ["k2048", "k3072"])
def test_sum_dot_segments(eng, which, request):
    kd, gk, ok = request.getfixturevalue(which)
    n = 150
    m, r = inputs(eng, n, 10)
    a = ok.encrypt_i64(m, r)
    offsets = np.array([0, 0, 1, 3, 40, 41, 41, 120, 150], dtype=np.int64)  # empty, 1, 2, 37, ...
    print(f"[{which}] sum segments={len(offsets) - 1}", flush=True)
    s = gk.sum(a, offsets)
    gk.ctx.sync()
    assert np.array_equal(host(s), ok.dot_i64(a, None, offsets))
    w = np.random.default_rng(11).integers(-(2**31), 2**31 - 1, size=n, dtype=np.int64)
    w[:6] = [0, 1, -1, 2, -2, 0]
    print(f"[{which}] dot segments={len(offsets) - 1}", flush=True)
    d = gk.dot_i64(a, w, offsets)
    gk.ctx.sync()
    assert np.array_equal(host(d), ok.dot_i64(a, w, offsets))


@pytest.mark.parametrize("which", ["k2048", "k3072"])
def test_encrypt_bigint_and_bounds(eng, which, request):
    kd, gk, ok = request.getfixturevalue(which)
    n_half = kd["n"] >> 1
    vals = [n_half, -n_half, n_half - 1, -(n_half - 12345), 2**200 + 7, -(2**1000), 0, 1, -1]
    r = eng.seeded_r(12, 11, len(vals), 256)
    mw = gk.n_words
    print(f"[{which}] encrypt big n={len(vals)}", flush=True)
    ct = gk.encrypt(vals, r)
    gk.ctx.sync()
    ref = ok.encrypt_twos(eng.ints_to_words(vals, mw, signed=True), r)
    assert np.array_equal(host(ct), ref)
    dec = gk.decrypt(ct)
    gk.ctx.sync()
    assert eng.words_to_ints(host(dec), signed=True) == vals
    from heu_amd._native import HeuError

    with pytest.raises(HeuError):
        gk.encrypt([n_half + 1], r[:1])


def test_mont_export_and_randomize(eng, k2048):
    kd, gk, ok = k2048
    m, r = inputs(eng, 32, 13)
    a = ok.encrypt_i64(m, r)
    print("[k2048] to_mont / from_mont", flush=True)
    tm = gk.to_mont(a, 64)
    gk.ctx.sync()
    assert np.array_equal(host(tm), ok.to_mont(a, 64))
    back = gk.from_mont(tm, 64)
    gk.ctx.sync()
    assert np.array_equal(host(back), a)
    print("[k2048] randomize / encrypt zero", flush=True)
    r2 = eng.seeded_r(14, 11, 32, 256)
    rz = gk.randomize(a, r2)
    gk.ctx.sync()
    assert np.array_equal(host(rz), ok.randomize(a, r2))
    z = gk.randomize(None, r2)
    gk.ctx.sync()
    assert np.array_equal(host(z), ok.encrypt_i64(np.zeros(32, np.int64), r2))


def test_errors(eng, ctx, k2048):
    from heu_amd._native import HeuError

    kd, gk, ok = k2048
    pub = eng.Key(ctx, kd["n"], kd["hs"])
    m, r = inputs(eng, 4, 15)
    ct = pub.encrypt_i64(m, r)
    ctx.sync()
    with pytest.raises(HeuError):
        pub.decrypt(ct)
    big_r = np.zeros((4, 9), np.uint32)
    big_r[:, 8] = 1 << 4  # r = 2^260: beyond the 26 x 10-bit table
    pub.encrypt_i64(m, big_r)
    with pytest.raises(HeuError):
        ctx.sync()
    ctx.sync()  # flags cleared
    # segmented reductions (offsets, indices, weights) and elementwise operands
    # that would send a device read past its array are refused on the host,
    # before any launch
    w = np.ones(4, np.int64)
    bad = [lambda: gk.sum(ct, [0, 5]), lambda: gk.sum(ct, [2, 1]), lambda: gk.prefix_mul(ct, [0, 5]),
           lambda: gk.dot_i64(ct, w[:3], [0, 4]), lambda: gk.sum_gather(ct, np.array([0, 4]), [0, 2]),
           lambda: gk.sum_gather(ct, np.array([0, -1]), [0, 2]), lambda: gk.sum_gather(ct, np.array([0]), [0, 2]),
           lambda: gk.dot_gather(ct, np.array([0, 1]), w[:1], [0, 2]),
           # elementwise second operands of the wrong length
           lambda: gk.add(ct, ct[:3]), lambda: gk.sub(ct[:3], ct), lambda: gk.mul_i64(ct, w[:3]),
           lambda: gk.add_i64(ct, w[:2]), lambda: gk.randomize(ct, r[:3]), lambda: gk.encrypt_i64(m, r[:3])]
    for i, call in enumerate(bad):
        with pytest.raises(ValueError):
            call()
            print("not refused:", i, flush=True)
    ok_sum = gk.sum_gather(ct, np.array([3, 0, 3]), [0, 3])
    ctx.sync()
    assert np.array_equal(host(ok_sum), ok.dot_i64(np.stack([host(ct)[i] for i in (3, 0, 3)]), None, [0, 3]))


def test_config1_workload(eng, oracle_mod, ctx):
    # the CPU reference case: 2048-bit key from seed 0, two vectors of 4096
    # int64-uniform plaintexts, r uniform in [1, 2^256) from the seed;
    # encrypt both, ct + ct, decrypt -- every output bit-exact with the oracle
    kd = eng.keygen(2048, 0)
    gk = eng.Key.from_keygen(ctx, kd, r_bits=256, window_bits=0)
    ok = oracle_mod.Key(kd["n"], kd["hs"], kd["p"], kd["q"])
    rng = np.random.default_rng(0)
    m1 = rng.integers(I64_MIN, I64_MAX, size=4096, dtype=np.int64, endpoint=True)
    m2 = rng.integers(I64_MIN, I64_MAX, size=4096, dtype=np.int64, endpoint=True)
    r1, r2 = eng.seeded_r(0, 1, 4096, 256), eng.seeded_r(0, 2, 4096, 256)
    print("config 1: encrypt x2, add, decrypt (4096)", flush=True)
    c1, c2 = gk.encrypt_i64(m1, r1), gk.encrypt_i64(m2, r2)
    s = gk.add(c1, c2)
    d = gk.decrypt(s, out_words=3)
    gk.ctx.sync()
    o1, o2 = ok.encrypt_i64(m1, r1, threads=8), ok.encrypt_i64(m2, r2, threads=8)
    assert np.array_equal(host(c1), o1) and np.array_equal(host(c2), o2)
    os_ = ok.add(o1, o2, threads=8)
    assert np.array_equal(host(s), os_)
    assert np.array_equal(host(d), ok.decrypt(os_, out_words=3, threads=8))
    assert [eng.twos_to_int(row) for row in host(d)] == [int(a) + int(b) for a, b in zip(m1, m2)]
    gk.close()


def test_empty_inputs(eng, ctx, k2048):
    # empty spans (an empty HeMatrix, a worker chunk with no elements) are no-ops
    # that return empty results; a sum or dot over an empty segment is the
    # identity 1, which decrypts to 0 (numpy/evaluator.cc:355-376 on zero terms)
    kd, gk, ok = k2048
    W = gk.ct_words
    m = np.zeros(0, np.int64)
    r = eng.seeded_r(5, 5, 0, 256)
    print("[k2048] empty one-op calls", flush=True)
    ct = gk.encrypt_i64(m, r)
    outs = [ct, gk.decrypt(ct, out_words=2), gk.add(ct, ct), gk.mul_i64(ct, m), gk.negate(ct), gk.add_i64(ct, m),
            gk.randomize(ct, r)]
    gk.ctx.sync()
    assert all(o.shape[0] == 0 for o in outs)
    assert gk.sum(ct, np.zeros(1, np.int64)).shape[0] == 0  # no segments
    one = gk.sum(ct, np.zeros(2, np.int64))  # one empty segment
    dot = gk.dot_i64(ct, m, np.zeros(2, np.int64))
    gk.ctx.sync()
    assert eng.words_to_int(host(one)[0]) == 1 and eng.words_to_int(host(dot)[0]) == 1
    print("[k2048] batch and server with empty tasks", flush=True)
    full_m, full_r = inputs(eng, 5)
    b = eng.Batch(ctx, graph=False)
    e_out = b.encrypt_i64(gk, m, r)
    f_out = b.encrypt_i64(gk, full_m, full_r)
    b.plan()
    b.run()
    gk.ctx.sync()
    assert e_out.shape[0] == 0 and np.array_equal(host(f_out), ok.encrypt_i64(full_m, full_r))
    b.close()
    srv = eng.Server(ctx, workspace_bytes=1 << 30, max_inflight=1)
    t_e, o_e = srv.encrypt_i64(gk, m, r)
    t_f, o_f = srv.encrypt_i64(gk, full_m, full_r)
    srv.wait(t_e, 60000.0)
    srv.wait(t_f, 60000.0)
    assert o_e.shape[0] == 0 and np.array_equal(host(o_f), ok.encrypt_i64(full_m, full_r))
    srv.close()


def test_chain_full_size(eng, ctx, k2048):
    # the headline workload at its full size (1,048,576 elements, one op per
    # launch): Dec((Enc(m) (+) Enc(m')) (x) k) = (m + m') k for every element,
    # a closed form independent of the size; plus a 4,096-element slice of the
    # ciphertexts bit-exact with the oracle (the rest is covered by the closed
    # form: a wrong ciphertext does not decrypt to it)
    from heu_amd.configs import chain_expected

    kd, gk, ok = k2048
    n = 1 << 20
    rng = np.random.default_rng(20)
    m = rng.integers(I64_MIN, I64_MAX, size=n, dtype=np.int64, endpoint=True)
    m[: len(EDGE_I64)] = EDGE_I64
    k = rng.integers(-(2**31), 2**31 - 1, size=n, dtype=np.int64, endpoint=True)
    k[:4] = [0, 1, -1, -(2**31)]
    r = eng.seeded_r(20, 11, n, 256)
    print("full-size chain: encrypt, ct+ct, ct x int32, decrypt (1,048,576)", flush=True)
    c1 = gk.encrypt_i64(m, r)
    c2 = gk.encrypt_i64(np.ascontiguousarray(m[::-1]), r)
    p = gk.mul_i64(gk.add(c1, c2), k)
    d = gk.decrypt(p, out_words=4)
    gk.ctx.sync()
    assert np.array_equal(host(d), chain_expected(m, m[::-1], k))
    assert np.array_equal(host(c1[:4096]), ok.encrypt_i64(m[:4096], r[:4096], threads=8))


@pytest.mark.parametrize("which", ["k2048", "k3072"])
def test_sum_every_chunk_length(eng, ctx, which, request):
    # Z-space sums over enough terms (> 31 x 32,768) that level 0 takes chunks
    # of 32, with segments whose last chunk holds each length 0..31: every
    # entry of the per-chunk R-power table is used.  Each segment's total must
    # decrypt to the plaintext sum; three segments are also bit-exact against
    # the oracle.  The same segments run as a gathered sum (a permuted index
    # list) and as a batch task (work-list kernel).
    kd, gk, ok = request.getfixturevalue(which)
    lens = [32 * 3 + r for r in range(32)] + [0, 1, 31, 33]
    lens.append(1_040_000 - sum(lens))
    offsets = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    n = int(offsets[-1])
    rng = np.random.default_rng(31)
    m = rng.integers(-(2**40), 2**40, size=n, dtype=np.int64)
    r = eng.seeded_r(31, 11, n, 256)
    print(f"[{which}] sum of {n} terms in {len(lens)} segments: encrypt", flush=True)
    ct = gk.encrypt_i64(m, r)
    want = [int(m[a:b].sum()) for a, b in zip(offsets[:-1], offsets[1:])]

    def check(s, what):
        d = gk.decrypt(s, out_words=4)
        gk.ctx.sync()
        got = eng.words_to_ints(host(d), signed=True)
        assert got == want, what

    print("one-op sum", flush=True)
    s = gk.sum(ct, offsets)
    check(s, "one-op sum")
    hs = host(s)
    for j in (0, 17, 31):
        a, b = offsets[j], offsets[j + 1]
        assert np.array_equal(hs[j : j + 1], ok.dot_i64(host(ct[a:b]), None, np.array([0, b - a], np.int64))), j
    print("gathered sum", flush=True)
    perm = rng.permutation(n).astype(np.int64)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(n, dtype=np.int64)
    g = gk.sum_gather(ct[torch.from_numpy(perm).to(ct.device)], torch.from_numpy(inv).to(ct.device), offsets)
    check(g, "gathered sum")
    assert np.array_equal(host(g), hs)
    print("batch sum", flush=True)
    b = eng.Batch(ctx, graph=False)
    bs = b.sum(gk, ct, offsets)
    b.run()
    ctx.sync()
    assert np.array_equal(host(bs), hs)
    b.close()
