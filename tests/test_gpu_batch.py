"""GPU parity of multi-task batches (heu_batch_*): persistent work-list
launches over uneven vectors and mixed 2048/3072-bit keys, eager and
hipGraph-captured, bit-exact against the one-op entry points (themselves
checked against the oracle in test_gpu_parity.py) and, for a subset, against
the oracle directly.
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


@pytest.fixture(scope="module")
def keys(eng, oracle_mod, ctx):
    out = {}
    for bits, seed in ((2048, 21), (3072, 22)):
        kd = eng.keygen(bits, seed)
        out[bits] = (kd, eng.Key.from_keygen(ctx, kd, r_bits=256, window_bits=10),
                     oracle_mod.Key(kd["n"], kd["hs"], kd["p"], kd["q"]))
    return out


def host(t):
    return t.cpu().numpy().view(np.uint32)


LENGTHS = [1, 2, 63, 64, 65, 127, 300, 1000, 7]


def vectors(eng, seed, lengths):
    rng = np.random.default_rng(seed)
    out = []
    for i, n in enumerate(lengths):
        m = rng.integers(I64_MIN, I64_MAX, size=n, dtype=np.int64, endpoint=True)
        k = rng.integers(-(2**31), 2**31 - 1, size=n, dtype=np.int64)
        k[: min(n, 3)] = [0, 1, -1][: min(n, 3)]
        r = eng.seeded_r(seed, 100 + i, n, 256)
        out.append((m, k, r))
    return out


def build_batch(eng, ctx, keys, graph):
    """Every op on both keys, several uneven vectors each; returns the batch,
    its outputs and the inputs the reference calls need."""
    b = eng.Batch(ctx, graph=graph)
    plan = []
    for bits in (2048, 3072):
        kd, gk, ok = keys[bits]
        vs = vectors(eng, bits, LENGTHS)
        for i, (m, k, r) in enumerate(vs):
            ct = b.encrypt_i64(gk, m, r)
            plan.append(("enc", bits, (m, r), ct))
            base = gk.encrypt_i64(m, r)  # resident input for the other ops
            m2, k2, r2 = vs[(i + 1) % len(vs)]
            n = len(m)
            other = gk.encrypt_i64(np.resize(m2, n), np.resize(r2, (n, r2.shape[1])))
            plan.append(("add", bits, (base, other), b.add(gk, base, other)))
            plan.append(("add_i64", bits, (base, k), b.add_i64(gk, base, k)))
            plan.append(("mul", bits, (base, k), b.mul_i64(gk, base, k)))
            plan.append(("dec", bits, (base,), b.decrypt(gk, base, out_words=3)))
            plan.append(("rand", bits, (base, r), b.randomize(gk, base, r)))
        # segmented sum / dot over uneven segments, one empty
        lens = [5, 0, 64, 33, 1, 200]
        offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
        tot = int(offs[-1])
        rng = np.random.default_rng(bits + 7)
        m = rng.integers(-(2**40), 2**40, size=tot, dtype=np.int64)
        w = rng.integers(-(2**31), 2**31 - 1, size=tot, dtype=np.int64)
        w[:4] = [0, 1, -1, 2]
        ct = gk.encrypt_i64(m, eng.seeded_r(bits, 999, tot, 256))
        plan.append(("sum", bits, (ct, offs), b.sum(gk, ct, offs)))
        plan.append(("dot", bits, (ct, w, offs), b.dot_i64(gk, ct, w, offs)))
    return b, plan


def expected(keys, name, bits, args):
    kd, gk, ok = keys[bits]
    if name == "enc":
        return gk.encrypt_i64(*args)
    if name == "add":
        return gk.add(*args)
    if name == "add_i64":
        return gk.add_i64(*args)
    if name == "mul":
        return gk.mul_i64(*args)
    if name == "dec":
        return gk.decrypt(args[0], out_words=3)
    if name == "rand":
        return gk.randomize(*args)
    if name == "sum":
        return gk.sum(*args)
    if name == "dot":
        return gk.dot_i64(*args)
    raise KeyError(name)


@pytest.mark.parametrize("graph", [False, True])
def test_batch_matches_single_ops(eng, ctx, keys, graph):
    print(f"batch graph={graph}: plan", flush=True)
    b, plan = build_batch(eng, ctx, keys, graph)
    st = b.stats()
    print(f"batch stats {st}", flush=True)
    assert st["units"] > 0
    want = {}
    for i, (name, bits, args, out) in enumerate(plan):
        want[i] = host(expected(keys, name, bits, args))
    ctx.sync()
    for rep in range(3):
        for _, _, _, out in plan:
            out.zero_()
        print(f"batch graph={graph}: run {rep}", flush=True)
        b.run()
        ctx.sync()
        for i, (name, bits, args, out) in enumerate(plan):
            assert np.array_equal(host(out), want[i]), (rep, name, bits, i)
    b.close()


def test_batch_against_oracle(eng, ctx, keys):
    # an uneven, mixed-key batch checked directly against the oracle
    b = eng.Batch(ctx, graph=True)
    checks = []
    for bits in (3072, 2048):
        kd, gk, ok = keys[bits]
        for i, n in enumerate([3, 129, 17]):
            m, k, r = vectors(eng, 31 + i, [n])[0]
            ct = b.encrypt_i64(gk, m, r)
            base = ok.encrypt_i64(m, r)
            mul = b.mul_i64(gk, base, k)
            dec = b.decrypt(gk, base, out_words=2)
            checks.append((ok, ("enc", m, r), ct))
            checks.append((ok, ("mul", base, k), mul))
            checks.append((ok, ("dec", m), dec))
    print("oracle batch: run", flush=True)
    b.run()
    ctx.sync()
    for ok, spec, out in checks:
        if spec[0] == "enc":
            assert np.array_equal(host(out), ok.encrypt_i64(spec[1], spec[2]))
        elif spec[0] == "mul":
            assert np.array_equal(host(out), ok.mul_i64(spec[1], spec[2]))
        else:
            assert np.array_equal(host(out).view(np.int64).ravel(), spec[1])
    b.close()


def test_grid_multiplier_does_not_change_results(eng, ctx, keys):
    kd, gk, ok = keys[2048]
    m, k, r = vectors(eng, 5, [777])[0]
    base = gk.encrypt_i64(m, r)
    outs = []
    for mult in (1, 4):
        from heu_amd._native import check, lib

        check(lib().heu_ctx_set_grid_multiplier(ctx.handle, 2048, 0, mult))
        check(lib().heu_ctx_set_grid_multiplier(ctx.handle, 2048, 1, mult))
        b = eng.Batch(ctx)
        o1, o2 = b.mul_i64(gk, base, k), b.decrypt(gk, base, out_words=2)
        b.run()
        ctx.sync()
        outs.append((host(o1), host(o2)))
        b.close()
    check(lib().heu_ctx_set_grid_multiplier(ctx.handle, 2048, 0, 1))
    check(lib().heu_ctx_set_grid_multiplier(ctx.handle, 2048, 1, 1))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    assert np.array_equal(outs[0][1].view(np.int64).ravel(), m)


def test_batch_errors(eng, ctx, keys):
    from heu_amd._native import HeuError

    kd, gk, ok = keys[2048]
    pub = eng.Key(ctx, kd["n"], kd["hs"], r_bits=64, window_bits=8)  # public only
    b = eng.Batch(ctx)
    b.decrypt(pub, gk.encrypt_i64(np.array([1], np.int64), eng.seeded_r(1, 1, 1, 64)))
    with pytest.raises(HeuError):
        b.plan()


def test_context_outlives_its_objects(eng):
    # heu_ctx_destroy before the key and batch made on it: they keep working,
    # and the context is freed with the last of them (any release order)
    from heu_amd._native import check, lib

    c = eng.Context(0)
    kd = eng.keygen(2048, 5)
    k = eng.Key.from_keygen(c, kd, r_bits=64, window_bits=4)
    m = np.array([1, -2, 3], np.int64)
    b = eng.Batch(c)
    ct = b.encrypt_i64(k, m, eng.seeded_r(1, 1, 3, 64))
    b.plan()
    h = c._h
    check(lib().heu_ctx_destroy(h))  # deferred: the key and the batch hold it
    c._h = None
    b.run()
    check(lib().heu_ctx_sync(h))
    d = k.decrypt(ct, out_words=2)
    check(lib().heu_ctx_sync(h))
    assert np.array_equal(host(d).view(np.int64).ravel(), m)
    b.close()
    k.close()  # the last holder: frees the context


@pytest.mark.parametrize("graph", [False, True])
def test_batch_stages_with_barriers(eng, ctx, keys, graph):
    # three stages over both key sizes, each stage reading the previous one's
    # outputs: encrypt | decrypt, ct x k, dot | decrypt(ct x k).  Each stage's
    # shapes run as parallel lanes; the barrier must join them all.
    b = eng.Batch(ctx, graph=graph)
    recs = []
    for bits in (2048, 3072):
        kd, gk, ok = keys[bits]
        for m, k, r in vectors(eng, bits + 5, [300, 65, 1]):
            recs.append([bits, m, k, r, b.encrypt_i64(gk, m, r)])
    b.barrier()
    for rec in recs:
        bits, m, k, r, ct = rec
        gk = keys[bits][1]
        offs = np.array([0, len(m) // 2, len(m)], dtype=np.int64)
        rec += [b.decrypt(gk, ct, out_words=2), b.mul_i64(gk, ct, k), b.dot_i64(gk, ct, k, offs), offs]
    b.barrier()
    for rec in recs:
        bits, m, k, r, ct, dec, mul, dot, offs = rec
        rec.append(b.decrypt(keys[bits][1], mul, out_words=3))
    st = b.stats()
    print(f"stages: graph={graph} {st}", flush=True)
    want = []
    for bits, m, k, r, ct, dec, mul, dot, offs, dmul in recs:
        gk = keys[bits][1]
        c = gk.encrypt_i64(m, r)
        want.append((host(c), host(gk.mul_i64(c, k)), host(gk.dot_i64(c, k, offs))))
    ctx.sync()
    for rep in range(3):
        for rec in recs:
            for t in rec[4:8] + rec[9:]:
                t.zero_()
        print(f"stages: graph={graph} run {rep}", flush=True)
        b.run()
        ctx.sync()
        for (bits, m, k, r, ct, dec, mul, dot, offs, dmul), (wc, wm, wd) in zip(recs, want):
            assert np.array_equal(host(ct), wc), (rep, bits, "encrypt")
            assert np.array_equal(host(dec).view(np.int64).ravel(), m), (rep, bits, "decrypt")
            assert np.array_equal(host(mul), wm), (rep, bits, "mul")
            assert np.array_equal(host(dot), wd), (rep, bits, "dot")
            got = [eng.twos_to_int(row) for row in host(dmul)]
            assert got == [int(a) * int(c) for a, c in zip(m, k)], (rep, bits, "decrypt of mul")
    b.close()
