"""The batched and served configurations at their full shapes, checked against
closed forms and the oracle (not only against the engine's own one-op path).

  config 3: one 3072-bit key, 64 vectors of log-uniform length in [1, 65536],
            int64 plaintexts, int32 weights; encrypt + dot in one hipGraph
            launch after the autotuner picked the key's parameters
            (numpy/evaluator.cc:209-292 MatMul).
  config 4: 16 keys (8 x 2048, 8 x 3072 bits), 16 vectors of log-uniform
            length in [256, 262144] split 40/20/20/20 into encrypt / decrypt /
            ct+ct / ct x int32, plus the axis-0 sum of a 4096 x 256 ciphertext
            matrix (numpy/evaluator.cc:355-376), all in one graph launch.
  config 5: Poisson arrivals against 32 resident keys (24 x 2048, 8 x 3072)
            through the continuous-batching server: every result equals the
            same request run alone and decrypts to its closed form.
The workloads are the seeded ones bench.py times (heu_amd/configs.py).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


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
def oracle():
    from oracle import oracle as O

    O.build()
    return O


def host(t):
    return t.cpu().numpy().view(np.uint32)


def i128_rows(values) -> np.ndarray:
    """Python ints -> 128-bit two's complement rows of 4 uint32 words."""
    raw = b"".join(int(v).to_bytes(16, "little", signed=True) for v in values)
    return np.frombuffer(raw, dtype=np.uint32).reshape(-1, 4)


def test_config3_uneven_3072_encrypt_dot_one_graph(eng, ctx, oracle):
    from heu_amd import configs as C

    kd = eng.keygen(3072, 3 << 16)
    key = eng.Key.from_keygen(ctx, kd, r_bits=C.R_BITS, window_bits=0)
    tune = key.autotune()
    print(f"config 3: autotuned window {key.table_window}, dec lanes {key.dec_lanes}", flush=True)
    lens = C.log_uniform_lengths(3, 0, 1, 64, 1, 65536)
    assert len(lens) == 64 and lens.min() >= 1 and lens.max() <= 65536
    total = int(lens.sum())
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    m_h, w_h = C.int64s(3, 0, 2, total), C.int32s(3, 0, 4, total)
    r_h = C.r_limbs(3, 0, 3, total)
    m, r, w = C._dev(m_h), C._dev(r_h.view(np.int32)), C._dev(w_h)
    ct_all = key._ct_out(total)
    b = eng.Batch(ctx, graph=True)
    for i in range(64):
        a, e = int(offs[i]), int(offs[i + 1])
        b.encrypt_i64(key, m[a:e], r[a:e], out=ct_all[a:e])
    b.barrier()
    dots = b.dot_i64(key, ct_all, w, offs)
    print(f"config 3: {total} elements, one graph launch per run", flush=True)
    b.run()
    ctx.sync()
    first = host(dots).copy()
    b.run()  # a graph replay gives the same bytes
    ctx.sync()
    assert np.array_equal(host(dots), first)
    assert b.stats()["launches"] > 0
    # every dot decrypts to sum m*w (Python integers)
    dec = host(key.decrypt(dots, out_words=4))
    ctx.sync()
    want = i128_rows([C.dot_closed_form(m_h, w_h, offs, i) for i in range(64)])
    assert np.array_equal(dec, want)
    # whole vectors against the oracle: ciphertexts and dots bit-exact
    ok = oracle.Key(kd["n"], kd["hs"])
    ct_h = host(ct_all)
    picks = [int(i) for i in np.argsort(lens)[:3]] + [int(np.argsort(np.abs(lens - 1500))[0])]
    for i in picks:
        a, e = int(offs[i]), int(offs[i + 1])
        assert np.array_equal(ct_h[a:e], ok.encrypt_i64(m_h[a:e], r_h[a:e])), f"vector {i} ciphertexts"
        ref = ok.dot_i64(ct_h[a:e], w_h[a:e], np.array([0, e - a], np.int64))
        assert np.array_equal(first[i:i + 1], ref), f"vector {i} (length {e - a}) dot"
    assert tune["table_window"]["chosen"] == key.table_window
    b.close()
    key.close()


def test_config4_mixed_keys_one_graph(eng, ctx, oracle):
    from heu_amd import configs as C

    kds, keys = [], []
    for i in range(16):
        bits = 2048 if i < 8 else 3072
        kds.append(eng.keygen(bits, (4 << 16) | i))
        keys.append(eng.Key.from_keygen(ctx, kds[-1], r_bits=C.R_BITS, window_bits=0))
    for first in (keys[0], keys[8]):
        first.autotune(eng._native.HEU_TUNE_DEC_LANES | eng._native.HEU_TUNE_GRID)
        for k in keys:
            if k.key_bits == first.key_bits:
                k.dec_lanes = first.dec_lanes
    lens = C.log_uniform_lengths(4, 0, 1, 16, 256, 262144)
    assert lens.min() >= 256 and lens.max() <= 262144
    b = eng.Batch(ctx, graph=True)
    jobs = []
    for i, (key, n) in enumerate(zip(keys, lens)):
        n = int(n)
        a0, a1, a2, a3, a4 = 0, int(0.4 * n), int(0.6 * n), int(0.8 * n), n
        m_h = C.int64s(4, 0, 100 + i, n)
        r_h = C.r_limbs(4, 0, 200 + i, n)
        k_h = C.int32s(4, 0, 300 + i, n)
        m, r, k = C._dev(m_h), C._dev(r_h.view(np.int32)), C._dev(k_h)
        ct = key.encrypt_i64(m, r)
        ct2 = key.encrypt_i64(m.flip(0).contiguous(), r)
        outs = (b.encrypt_i64(key, m[a0:a1], r[a0:a1]), b.decrypt(key, ct[a1:a2], out_words=2),
                b.add(key, ct[a2:a3], ct2[a2:a3]), b.mul_i64(key, ct[a3:a4], k[a3:a4]))
        jobs.append((i, key, (a0, a1, a2, a3, a4), m_h, r_h, k_h, ct, ct2, outs))
    rows, cols = 4096, 256
    mat_m = C.int64s(4, 0, 400, rows * cols) >> 16
    mat_r = C.r_limbs(4, 0, 401, rows * cols)
    mat = keys[0].encrypt_i64(C._dev(mat_m), C._dev(mat_r.view(np.int32)))
    col_offs = np.arange(0, rows * cols + 1, rows, dtype=np.int64)
    sums = b.sum(keys[0], mat, col_offs)
    ctx.sync()
    print(f"config 4: {int(lens.sum())} elements on 16 keys + a {rows}x{cols} axis-0 sum, one graph launch", flush=True)
    b.run()
    ctx.sync()
    for i, key, (a0, a1, a2, a3, a4), m_h, r_h, k_h, ct, ct2, (o_enc, o_dec, o_add, o_mul) in jobs:
        m_rev = m_h[::-1]
        # decrypt-check every output against numpy / Python integers
        d_enc = host(key.decrypt(o_enc, out_words=2)).view(np.int64).ravel()
        assert np.array_equal(d_enc, m_h[a0:a1]), f"vector {i} encrypt"
        assert np.array_equal(host(o_dec).view(np.int64).ravel(), m_h[a1:a2]), f"vector {i} decrypt"
        d_add = host(key.decrypt(o_add, out_words=4))
        want_add = i128_rows(int(x) + int(y) for x, y in zip(m_h[a2:a3], m_rev[a2:a3]))
        assert np.array_equal(d_add, want_add), f"vector {i} ct+ct"
        d_mul = host(key.decrypt(o_mul, out_words=4))
        want_mul = i128_rows(int(x) * int(y) for x, y in zip(m_h[a3:a4], k_h[a3:a4]))
        assert np.array_equal(d_mul, want_mul), f"vector {i} ct x int32"
        # a slice per key against the oracle, bit-exact
        ok = oracle.Key(kds[i]["n"], kds[i]["hs"], kds[i]["p"], kds[i]["q"])
        s = 8
        assert np.array_equal(host(o_enc)[:s], ok.encrypt_i64(m_h[a0:a0 + s], r_h[a0:a0 + s]))
        assert np.array_equal(host(o_dec)[:s], ok.decrypt(host(ct[a1:a1 + s]), out_words=2))
        assert np.array_equal(host(o_add)[:s], ok.add(host(ct[a2:a2 + s]), host(ct2[a2:a2 + s])))
        assert np.array_equal(host(o_mul)[:s], ok.mul_i64(host(ct[a3:a3 + s]), k_h[a3:a3 + s]))
    # the matrix sum: every column decrypts to its column sum; two columns vs the oracle
    d_sum = host(keys[0].decrypt(sums, out_words=2)).view(np.int64).ravel()
    assert np.array_equal(d_sum, mat_m.reshape(cols, rows).sum(axis=1))
    ok0 = oracle.Key(kds[0]["n"], kds[0]["hs"])
    mat_h = host(mat)
    for j in (0, cols - 1):
        ref = ok0.dot_i64(mat_h[j * rows:(j + 1) * rows], None, np.array([0, rows], np.int64))
        assert np.array_equal(host(sums)[j:j + 1], ref)
    b.close()
    for k in keys:
        k.close()


def test_config5_served_poisson_32_keys(eng, ctx):
    from heu_amd import configs as C

    pools = []
    for i in range(32):
        bits = 2048 if i < 24 else 3072
        key = eng.Key.from_keygen(ctx, eng.keygen(bits, (5 << 16) | i), r_bits=C.R_BITS, window_bits=10)
        pools.append(C._Pool(key, 0, i, C.MAX_LEN5))
    ctx.sync()
    rate, count = 45.0, 160
    rq = C._requests5(1, count, len(pools))
    arrivals = np.cumsum(-np.log(1.0 - C.uniform01(5, 1, 20, count)) / rate)
    srv = eng.Server(ctx, workspace_bytes=24 << 30, max_inflight=2)
    print(f"config 5: {count} Poisson requests at {rate:g} req/s over 32 keys", flush=True)
    lat, wall, tickets = C._served(srv, pools, rq, arrivals)
    st = srv.stats()
    srv.close()
    assert st["completed"] == count and st["failed"] == 0 and st["max_batch_requests"] > 1
    print(f"config 5: served p99 {np.percentile(lat, 99):.1f} ms, {st['batches']} batches", flush=True)
    for j, (op, ki, n) in enumerate(rq):
        pool = pools[ki]
        got = tickets[j][1]
        alone = C._call(pool.key, pool, op, n)
        ctx.sync()
        assert torch.equal(alone, got), f"request {j} ({op}, n={n}) differs from its alone run"
        m = pool.m[:n].cpu().numpy()
        if op == "dec":
            assert np.array_equal(host(got), i128_rows(m.tolist())), f"request {j} (dec, n={n})"
            continue
        d = host(pool.key.decrypt(got, out_words=4))
        if op == "enc":
            want = i128_rows(m.tolist())
        elif op == "add":
            m2 = pool.m.flip(0)[:n].cpu().numpy()
            want = i128_rows(int(x) + int(y) for x, y in zip(m, m2))
        elif op == "mul":
            k = pool.k[:n].cpu().numpy()
            want = i128_rows(int(x) * int(y) for x, y in zip(m, k))
        else:
            k = pool.k[:n].cpu().numpy()
            want = i128_rows([C.dot_closed_form(m, k, np.array([0, n]), 0)])
        assert np.array_equal(d, want), f"request {j} ({op}, n={n}) closed form"
    for p in pools:
        p.key.close()
