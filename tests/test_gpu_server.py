"""GPU tests of the request server (heu_server_*): requests of mixed ops,
sizes and keys submitted back to back are co-batched, and every result is
bit-exact the same as the same request run alone (one-op entry points)."""
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
def keys(eng, ctx):
    out = []
    for bits, seed in ((2048, 41), (3072, 42), (2048, 43)):
        kd = eng.keygen(bits, seed)
        out.append(eng.Key.from_keygen(ctx, kd, r_bits=256, window_bits=10))
    return out


def host(t):
    return t.cpu().numpy().view(np.uint32)


def make_requests(eng, keys, count, seed):
    """(op, key, args) with mixed sizes; args are device tensors / arrays."""
    rng = np.random.default_rng(seed)
    reqs = []
    ops = ["dec", "dec"] + list(rng.choice(["enc", "dec", "add", "mul", "dot", "sum"], size=count - 2))
    for i, op in enumerate(ops):
        key = keys[i % len(keys)]
        n = int(np.exp(rng.uniform(0, np.log(3000)))) if i >= 2 else 6000
        m = rng.integers(-(2**62), 2**62, size=n, dtype=np.int64)
        r = eng.seeded_r(seed, i, n, 256)
        ct = key.encrypt_i64(m, r)
        if op == "enc":
            args = (m, r)
        elif op == "dec":
            args = (ct,)
        elif op == "add":
            args = (ct, key.encrypt_i64(m[::-1].copy(), r))
        elif op == "mul":
            args = (ct, rng.integers(-(2**31), 2**31 - 1, size=n, dtype=np.int64))
        else:
            cuts = np.sort(rng.integers(0, n + 1, size=3))
            offs = np.concatenate([[0], cuts, [n]]).astype(np.int64)
            w = rng.integers(-(2**31), 2**31 - 1, size=n, dtype=np.int64)
            args = (ct, w, offs) if op == "dot" else (ct, offs)
        reqs.append((op, key, args))
    return reqs


def alone(key, op, args):
    if op == "enc":
        return key.encrypt_i64(*args)
    if op == "dec":
        return key.decrypt(args[0], out_words=3)
    if op == "add":
        return key.add(*args)
    if op == "mul":
        return key.mul_i64(*args)
    if op == "dot":
        return key.dot_i64(*args)
    return key.sum(*args)


def submit(srv, key, op, args):
    if op == "enc":
        return srv.encrypt_i64(key, *args)
    if op == "dec":
        return srv.decrypt(key, args[0], out_words=3)
    if op == "add":
        return srv.add(key, *args)
    if op == "mul":
        return srv.mul_i64(key, *args)
    if op == "dot":
        return srv.dot_i64(key, *args)
    return srv.sum(key, *args)


def test_alone_equals_mix(eng, ctx, keys):
    reqs = make_requests(eng, keys, 40, 7)
    ctx.sync()
    srv = eng.Server(ctx, workspace_bytes=4 << 30, max_inflight=2)
    print("server: submitting 40 requests", flush=True)
    pending = [(op, key, args, *submit(srv, key, op, args)) for op, key, args in reqs]
    lat = [srv.wait(t, 60000.0) for _, _, _, t, _ in pending]
    st = srv.stats()
    print(f"server stats {st}", flush=True)
    assert st["completed"] == 40 and st["failed"] == 0
    assert st["max_batch_requests"] > 1  # requests were co-batched
    assert all(x >= 0 for x in lat)
    for op, key, args, _, out in pending:
        want = alone(key, op, args)
        ctx.sync()
        assert np.array_equal(host(out), host(want)), op
    srv.close()


def test_urgent_requests_stay_exact(eng, ctx, keys):
    # urgent_ms = 1: almost every request is urgent by the time it is picked, so
    # batches take the oldest request's whole room first; results stay exact
    reqs = make_requests(eng, keys, 24, 8)
    ctx.sync()
    srv = eng.Server(ctx, workspace_bytes=4 << 30, max_inflight=2, urgent_ms=1)
    print("server (urgent_ms=1): submitting 24 requests", flush=True)
    pending = [(op, key, args, *submit(srv, key, op, args)) for op, key, args in reqs]
    for _, _, _, t, _ in pending:
        srv.wait(t, 60000.0)
    st = srv.stats()
    assert st["completed"] == 24 and st["failed"] == 0
    for op, key, args, _, out in pending:
        want = alone(key, op, args)
        ctx.sync()
        assert np.array_equal(host(out), host(want)), op
    srv.close()


def test_server_admission_limits(eng, ctx, keys):
    # a workspace too small for even one part of a big decryption fails that
    # request and keeps serving others
    key = keys[0]
    m = np.arange(-50, 50, dtype=np.int64)
    ct = key.encrypt_i64(m, eng.seeded_r(3, 3, len(m), 256))
    big = key.encrypt_i64(np.zeros(200000, np.int64), eng.seeded_r(3, 4, 200000, 256))
    ctx.sync()
    srv = eng.Server(ctx, workspace_bytes=8 << 20, max_inflight=1, max_workspace_bytes=8 << 20)
    t_big, _ = srv.decrypt(key, big, out_words=2)
    t_ok, out = srv.decrypt(key, ct, out_words=2)
    from heu_amd._native import HeuError

    with pytest.raises(HeuError):
        srv.wait(t_big, 60000.0)
    srv.wait(t_ok, 60000.0)
    assert np.array_equal(host(out).view(np.int64).ravel(), m)
    with pytest.raises(HeuError):
        srv.wait(12345)
    srv.close()


def test_long_request_does_not_block_short_ones(eng, ctx, keys):
    # a long decryption is cut into parts; short requests submitted after it
    # finish long before it does, and every result is exact
    key = keys[1]  # 3072-bit
    n_big = 60000
    big_m = np.arange(n_big, dtype=np.int64) - n_big // 2
    big = key.encrypt_i64(big_m, eng.seeded_r(9, 1, n_big, 256))
    small = [key.encrypt_i64(np.array([i, -i, 7 * i], np.int64), eng.seeded_r(9, 10 + i, 3, 256)) for i in range(8)]
    ctx.sync()
    srv = eng.Server(ctx, max_inflight=2)
    t_big, out_big = srv.decrypt(key, big, out_words=2)
    subs = [srv.decrypt(key, c, out_words=2) for c in small]
    lat_small = [srv.wait(t, 120000.0) for t, _ in subs]
    lat_big = srv.wait(t_big, 120000.0)
    print(f"big {lat_big:.1f} ms, small max {max(lat_small):.1f} ms", flush=True)
    assert max(lat_small) < lat_big
    assert np.array_equal(host(out_big).view(np.int64).ravel(), big_m)
    for i, (_, out) in enumerate(subs):
        assert np.array_equal(host(out).view(np.int64).ravel(), np.array([i, -i, 7 * i], np.int64))
    srv.close()


def test_failed_part_waits_for_running_parts(eng, ctx, keys, monkeypatch):
    # One long decryption is carved into parts of one quantum each
    # (max_batch_tasks=1), so several of its parts run at once on different
    # streams.  The second batch is reported as not enqueued after its
    # launches were issued (HEU_SERVER_FAULT_BATCH, forced on the host; every
    # kernel still runs to its end).  The request fails, but wait() may report
    # that only once no part of it is still on the device: by then every
    # admitted part has written its decryptions, and the rest of the output was
    # never touched.
    key = keys[0]
    n = 120000
    m = (np.arange(n, dtype=np.int64) * 7919) - n
    ct = key.encrypt_i64(m, eng.seeded_r(12, 1, n, 256))
    ctx.sync()
    sentinel = -0x5A5A5A5A
    monkeypatch.setenv("HEU_SERVER_FAULT_BATCH", "2")
    srv = eng.Server(ctx, max_inflight=4, max_batch_tasks=1)
    monkeypatch.delenv("HEU_SERVER_FAULT_BATCH")
    out = torch.full((n, 2), sentinel, dtype=torch.int32, device=ctx.device)
    torch.cuda.synchronize()  # the server's streams do not wait for torch's
    t, _ = srv._record(key, eng._native.HEU_OP_DECRYPT, key._dev(ct, torch.int32), None, n, out, aux=2)
    from heu_amd._native import HeuError

    print("server: one carved decryption, batch 2 fails", flush=True)
    with pytest.raises(HeuError, match="could not enqueue"):
        srv.wait(t, 120000.0)
    got = out.cpu().numpy()  # read at once: no device-wide sync in between
    done = ~np.all(got == sentinel, axis=1)
    x = int(np.argmin(done)) if not done.all() else n
    print(f"decrypted prefix {x} of {n}", flush=True)
    assert 0 < x < n
    assert done[:x].all() and not done[x:].any()
    assert np.array_equal(got[:x].view(np.int64).ravel(), m[:x])
    st = srv.stats()
    assert st["failed"] == 1 and st["batches"] >= 2
    srv.close()


def test_workspace_grows_by_free_hbm(eng, ctx, keys):
    # admission by free device memory: a request whose batch does not fit the
    # initial workspace grows it by a new device segment (hipMemGetInfo shows
    # room), instead of failing; the results stay exact
    key = keys[0]
    n = 200000
    m = (np.arange(n, dtype=np.int64) * 31337) - n
    ct = key.encrypt_i64(m, eng.seeded_r(14, 1, n, 256))
    ctx.sync()
    srv = eng.Server(ctx, workspace_bytes=8 << 20, max_inflight=2)
    assert srv.stats()["workspace_bytes"] == 8 << 20
    print("server: decryption needing more than the initial 8 MiB workspace", flush=True)
    t, out = srv.decrypt(key, ct, out_words=2)
    srv.wait(t, 120000.0)
    st = srv.stats()
    print(f"server workspace grew to {st['workspace_bytes'] >> 20} MiB", flush=True)
    assert st["workspace_bytes"] > 8 << 20 and st["failed"] == 0
    assert np.array_equal(host(out).view(np.int64).ravel(), m)
    srv.close()
