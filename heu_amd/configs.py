"""Seeded workloads of the batched and served configurations, and their runners.

Every random quantity comes from the engine's counter PRNG (`prng_words`)
keyed by (config id << 32 | seed, stream); the streams are listed per config.

  config 3  one 3072-bit key; 64 vectors, lengths log-uniform in [1, 65536];
            int64 plaintexts, r < 2^256, int32 weights; encrypt every vector,
            then one dot product per vector (gradient aggregation).
  config 4  16 keys (8 x 2048, 8 x 3072 bits); 16 vectors, lengths
            log-uniform in [256, 262144], vector i on key i, its elements split
            40/20/20/20 into encrypt / decrypt / ct+ct / ct x int32; plus the
            HeMatrix axis-0 sum of a 4096 x 256 ciphertext matrix (column-major,
            numpy/matrix.h) on key 0.
  config 5  32 resident keys (24 x 2048, 8 x 3072 bits); Poisson arrivals;
            requests of encrypt / decrypt / ct+ct / ct x pt / dot weighted
            30/30/15/15/10, lengths log-uniform in [1, 262144], keys uniform.

Each runner times the plain path (one engine call per op and vector, one op
per launch) beside the batched form and checks that both give the same bytes.
Config 5 also replays its arrivals through the plain batched path (static
batching: whatever has arrived runs as one batch once the device is idle),
whose per-batch time is what served tail latency is measured against.
"""
from __future__ import annotations

import math
import threading
import time

import numpy as np
import torch

from . import engine as E

R_BITS = 256


def words(config: int, seed: int, stream: int, count: int) -> np.ndarray:
    return E.prng_words((config << 32) | seed, stream, count)


def uniform01(config, seed, stream, count) -> np.ndarray:
    return (words(config, seed, stream, count) >> np.uint64(11)).astype(np.float64) / float(1 << 53)


def log_uniform_lengths(config, seed, stream, count, lo, hi) -> np.ndarray:
    u = uniform01(config, seed, stream, count)
    return np.clip(np.floor(np.exp(math.log(lo) + u * (math.log(hi + 1) - math.log(lo)))), lo, hi).astype(np.int64)


def int64s(config, seed, stream, count) -> np.ndarray:
    return words(config, seed, stream, count).view(np.int64)


def int32s(config, seed, stream, count) -> np.ndarray:
    return (words(config, seed, stream, count) >> np.uint64(32)).astype(np.uint32).view(np.int32).astype(np.int64)


def r_limbs(config, seed, stream, count) -> np.ndarray:
    return E.seeded_r((config << 32) | seed, stream, count, R_BITS)


def chain_expected(m: np.ndarray, m2: np.ndarray, k: np.ndarray) -> np.ndarray:
    """(m + m2) * k per element as 128-bit two's complement, 4 uint32 words each:
    what decrypting (Enc(m) (+) Enc(m2)) (x) k must give, from Python integers."""
    vals = ((int(a) + int(b)) * int(c) for a, b, c in zip(m.tolist(), m2.tolist(), k.tolist()))
    raw = b"".join(v.to_bytes(16, "little", signed=True) for v in vals)
    return np.frombuffer(raw, dtype=np.uint32).reshape(len(m), 4)


def dot_closed_form(m: np.ndarray, w: np.ndarray, offs: np.ndarray, i: int) -> int:
    """sum m*w over segment i as a Python integer (exact: int64 products in
    object arithmetic)."""
    a, e = int(offs[i]), int(offs[i + 1])
    return int(np.dot(m[a:e].astype(object), w[a:e].astype(object))) if e > a else 0


def _dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return t if dtype is None else t.to(dtype)


def _elapsed_ms(fn, ctx) -> float:
    torch.cuda.synchronize()
    t = time.perf_counter()
    fn()
    ctx.sync()
    torch.cuda.synchronize()
    return (time.perf_counter() - t) * 1e3


# ---------------------------------------------------------------------------
# config 3
# ---------------------------------------------------------------------------
def run_config3(ctx: E.Context, seed: int = 0, reps: int = 3) -> dict:
    kd = E.keygen(3072, (3 << 16) | seed)
    key = E.Key.from_keygen(ctx, kd, r_bits=R_BITS, window_bits=0)
    tune = key.autotune()
    lens = log_uniform_lengths(3, seed, 1, 64, 1, 65536)
    total = int(lens.sum())
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    m = _dev(int64s(3, seed, 2, total))
    r = _dev(r_limbs(3, seed, 3, total).view(np.int32))
    w = _dev(int32s(3, seed, 4, total))
    ctx.sync()

    # plain path: per vector, encrypt then dot (one op per launch, one call each)
    def plain():
        outs = []
        for i in range(64):
            a, b = int(offs[i]), int(offs[i + 1])
            ct = key.encrypt_i64(m[a:b], r[a:b])
            outs.append(key.dot_i64(ct, w[a:b], np.array([0, b - a], np.int64)))
        return outs

    plain_out = plain()
    ctx.sync()
    plain_ms = min(_elapsed_ms(plain, ctx) for _ in range(reps))

    # batched: 64 encrypt tasks into one buffer, barrier, one 64-segment dot; one graph launch
    ct_all = key._ct_out(total)
    b = E.Batch(ctx, graph=True)
    for i in range(64):
        a, e = int(offs[i]), int(offs[i + 1])
        b.encrypt_i64(key, m[a:e], r[a:e], out=ct_all[a:e])
    b.barrier()
    dots = b.dot_i64(key, ct_all, w, offs)
    b.run()
    ctx.sync()
    st = b.stats()
    batch_ms = min(_elapsed_ms(b.run, ctx) for _ in range(reps))
    same = all(torch.equal(dots[i:i + 1], plain_out[i]) for i in range(64))
    # closed form on every vector: Dec(dot) = sum m*w
    dec = key.decrypt(dots, out_words=4).cpu().numpy().view(np.uint32)
    m_h, w_h = m.cpu().numpy(), w.cpu().numpy()
    closed = all(E.twos_to_int(dec[i]) == dot_closed_form(m_h, w_h, offs, i) for i in range(64))
    b.close()
    key.close()
    return {
        "workload": "paillier3072: 64 uneven vectors (log-uniform 1..65536), encrypt + dot with int32 weights",
        "elements": total,
        "plain_ms": plain_ms,
        "batch_graph_ms": batch_ms,
        "plain_elements_per_s": total / (plain_ms / 1e3),
        "batch_elements_per_s": total / (batch_ms / 1e3),
        "speedup": plain_ms / batch_ms,
        "batch_launches": st["launches"],
        "work_units": st["units"],
        "bit_exact_vs_plain": bool(same),
        "closed_form_check": bool(closed),
        "autotune": tune,
        "work": {"enc": {3072: total}, "dot": {3072: total}},
    }


# ---------------------------------------------------------------------------
# config 4
# ---------------------------------------------------------------------------
def run_config4(ctx: E.Context, seed: int = 0, reps: int = 3) -> dict:
    keys = []
    for i in range(16):
        bits = 2048 if i < 8 else 3072
        keys.append(E.Key.from_keygen(ctx, E.keygen(bits, (4 << 16) | (seed << 8) | i), r_bits=R_BITS, window_bits=0))
    # the autotuner runs on one key per size; its choices apply to the others
    tune = {}
    for first in (keys[0], keys[8]):
        rep = first.autotune()
        tune[first.key_bits] = rep
        for k in keys:
            if k.key_bits == first.key_bits and k is not first:
                if k.table_window != first.table_window:
                    k.table_window = first.table_window
                k.dec_lanes = first.dec_lanes
    lens = log_uniform_lengths(4, seed, 1, 16, 256, 262144)
    jobs = []  # (op, key, inputs)
    total = 0
    work = {op: {2048: 0, 3072: 0} for op in ("enc", "dec", "add", "mul", "sum")}
    for i, (key, n) in enumerate(zip(keys, lens)):
        n = int(n)
        cut = [0, int(0.4 * n), int(0.6 * n), int(0.8 * n), n]
        m = _dev(int64s(4, seed, 100 + i, n))
        r = _dev(r_limbs(4, seed, 200 + i, n).view(np.int32))
        k = _dev(int32s(4, seed, 300 + i, n))
        ct = key.encrypt_i64(m, r)  # resident inputs of decrypt / add / mul
        ct2 = key.encrypt_i64(m.flip(0).contiguous(), r)
        a0, a1, a2, a3, a4 = cut
        jobs.append(("enc", key, (m[a0:a1], r[a0:a1])))
        jobs.append(("dec", key, (ct[a1:a2],)))
        jobs.append(("add", key, (ct[a2:a3], ct2[a2:a3])))
        jobs.append(("mul", key, (ct[a3:a4], k[a3:a4])))
        for op, lo, hi in (("enc", a0, a1), ("dec", a1, a2), ("add", a2, a3), ("mul", a3, a4)):
            work[op][key.key_bits] += hi - lo
        total += n
    # HeMatrix 4096 x 256 (column-major) axis-0 sum on key 0
    rows, cols = 4096, 256
    mat = keys[0].encrypt_i64(_dev(int64s(4, seed, 400, rows * cols) >> 16), _dev(r_limbs(4, seed, 401, rows * cols).view(np.int32)))
    col_offs = np.arange(0, rows * cols + 1, rows, dtype=np.int64)
    work["sum"][2048] += rows * cols
    ctx.sync()

    def plain():
        outs = []
        for op, key, args in jobs:
            if op == "enc":
                outs.append(key.encrypt_i64(*args))
            elif op == "dec":
                outs.append(key.decrypt(args[0], out_words=2))
            elif op == "add":
                outs.append(key.add(*args))
            else:
                outs.append(key.mul_i64(*args))
        outs.append(keys[0].sum(mat, col_offs))
        return outs

    plain_out = plain()
    ctx.sync()
    plain_ms = min(_elapsed_ms(plain, ctx) for _ in range(reps))
    b = E.Batch(ctx, graph=True)
    bouts = []
    for op, key, args in jobs:
        if op == "enc":
            bouts.append(b.encrypt_i64(key, *args))
        elif op == "dec":
            bouts.append(b.decrypt(key, args[0], out_words=2))
        elif op == "add":
            bouts.append(b.add(key, *args))
        else:
            bouts.append(b.mul_i64(key, *args))
    bouts.append(b.sum(keys[0], mat, col_offs))
    b.run()
    ctx.sync()
    st = b.stats()
    batch_ms = min(_elapsed_ms(b.run, ctx) for _ in range(reps))
    same = all(torch.equal(x, y) for x, y in zip(bouts, plain_out))
    b.close()
    for k in keys:
        k.close()
    return {
        "workload": "16 keys (8x2048, 8x3072), 16 uneven vectors (log-uniform 256..262144) split "
                    "40/20/20/20 enc/dec/add/mul, plus a 4096x256 HeMatrix axis-0 sum",
        "elements": total,
        "matrix_sum_terms": rows * cols,
        "plain_ms": plain_ms,
        "batch_graph_ms": batch_ms,
        "plain_elements_per_s": total / (plain_ms / 1e3),
        "batch_elements_per_s": total / (batch_ms / 1e3),
        "speedup": plain_ms / batch_ms,
        "batch_launches": st["launches"],
        "work_units": st["units"],
        "bit_exact_vs_plain": bool(same),
        "autotune": tune,
        "work": work,
    }


# ---------------------------------------------------------------------------
# config 5
# ---------------------------------------------------------------------------
OPS5 = ("enc", "dec", "add", "mul", "dot")
WEIGHTS5 = (0.30, 0.30, 0.15, 0.15, 0.10)
MAX_LEN5 = 262144


class _Pool:
    """Per-key resident inputs; requests use leading slices."""

    def __init__(self, key, seed, i, n):
        self.key = key
        self.m = _dev(int64s(5, seed, 1000 + i, n))
        self.r = _dev(r_limbs(5, seed, 2000 + i, n).view(np.int32))
        self.k = _dev(int32s(5, seed, 3000 + i, n))
        self.ct = key.encrypt_i64(self.m, self.r)
        self.ct2 = key.encrypt_i64(self.m.flip(0).contiguous(), self.r)


def _requests5(seed, count, nkeys):
    u_op = uniform01(5, seed, 10, count)
    lens = log_uniform_lengths(5, seed, 11, count, 1, MAX_LEN5)
    key_idx = (words(5, seed, 12, count) % np.uint64(nkeys)).astype(np.int64)
    cum = np.cumsum(WEIGHTS5)
    ops = [OPS5[int(np.searchsorted(cum, u, side="right"))] if u < cum[-1] else OPS5[-1] for u in u_op]
    return [(op, int(ki), int(n)) for op, ki, n in zip(ops, key_idx, lens)]


def _call(target, pool: _Pool, op, n):
    """Runs / submits one request through `target` (a Key, or a Server taking the key first)."""
    key = pool.key
    server = isinstance(target, (E.Server, E.Batch))
    args = {
        "enc": (pool.m[:n], pool.r[:n]),
        "dec": (pool.ct[:n],),
        "add": (pool.ct[:n], pool.ct2[:n]),
        "mul": (pool.ct[:n], pool.k[:n]),
        "dot": (pool.ct[:n], pool.k[:n], np.array([0, n], np.int64)),
    }[op]
    name = {"enc": "encrypt_i64", "dec": "decrypt", "add": "add", "mul": "mul_i64", "dot": "dot_i64"}[op]
    kw = {"out_words": 4} if op == "dec" else {}
    fn = getattr(target, name)
    return fn(key, *args, **kw) if server else fn(*args, **kw)


def _plain_fifo(ctx, pools, rq, arrivals):
    """The plain path as a server: requests run one engine call each, in
    arrival order, one at a time; latency = completion - arrival."""
    lat = []
    t0 = time.perf_counter()
    for j, (op, ki, n) in enumerate(rq):
        wait = arrivals[j] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        _call(pools[ki].key, pools[ki], op, n)
        ctx.sync()
        lat.append((time.perf_counter() - t0 - arrivals[j]) * 1e3)
    return lat, time.perf_counter() - t0


def _plain_batched(ctx, pools, rq, arrivals):
    """The plain batched path (static batching): whenever the device is idle,
    every request that has arrived is planned into one batch and run to
    completion; requests arriving meanwhile wait for the next batch.  Returns
    latencies (completion - arrival) and each batch's time (plan + run)."""
    lat = [0.0] * len(rq)
    batch_ms = []
    t0 = time.perf_counter()
    j = 0
    while j < len(rq):
        wait = arrivals[j] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        now = time.perf_counter() - t0
        k = j
        while k < len(rq) and arrivals[k] <= now:
            k += 1
        tb = time.perf_counter()
        b = E.Batch(ctx, graph=False)
        for jj in range(j, k):
            op, ki, n = rq[jj]
            _call(b, pools[ki], op, n)
        b.run()
        ctx.sync()
        done = time.perf_counter()
        b.close()
        batch_ms.append((done - tb) * 1e3)
        for jj in range(j, k):
            lat[jj] = (done - t0 - arrivals[jj]) * 1e3
        j = k
    return lat, time.perf_counter() - t0, batch_ms


def _served(srv, pools, rq, arrivals):
    tickets = [None] * len(rq)
    t0 = time.perf_counter()

    def submitter():
        for j, (op, ki, n) in enumerate(rq):
            wait = arrivals[j] - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            tickets[j] = _call(srv, pools[ki], op, n)

    th = threading.Thread(target=submitter)
    th.start()
    lat = []
    for j in range(len(rq)):
        while tickets[j] is None:
            time.sleep(0.0002)
        lat.append(srv.wait(tickets[j][0], 600000.0))
    th.join()
    return lat, time.perf_counter() - t0, tickets


def _summary(rate, lat, wall):
    return {
        "offered_req_per_s": rate,
        "requests": len(lat),
        "throughput_req_per_s": len(lat) / wall,
        "p50_ms": float(np.percentile(lat, 50)),
        "p95_ms": float(np.percentile(lat, 95)),
        "p99_ms": float(np.percentile(lat, 99)),
        "max_ms": float(np.max(lat)),
        "mean_ms": float(np.mean(lat)),
    }


def run_config5(ctx: E.Context, seed: int = 0, rates=(15.0, 30.0, 45.0, 60.0, 90.0, 120.0),
                per_rate: int = 320) -> dict:
    pools = []
    for i in range(32):
        bits = 2048 if i < 24 else 3072
        key = E.Key.from_keygen(ctx, E.keygen(bits, (5 << 16) | (seed << 8) | i), r_bits=R_BITS, window_bits=10)
        pools.append(_Pool(key, seed, i, MAX_LEN5))
    ctx.sync()
    srv = E.Server(ctx, workspace_bytes=24 << 30, max_inflight=2, urgent_ms=150)
    plain, batched, served = [], [], []
    work = {op: {2048: 0, 3072: 0} for op in OPS5}
    check_same = None
    for ri, rate in enumerate(rates):
        rq = _requests5(seed + 1 + ri, per_rate, len(pools))
        for op, ki, n in rq:
            work[op][pools[ki].key.key_bits] += n
        gaps = -np.log(1.0 - uniform01(5, seed + 1 + ri, 20, per_rate)) / rate
        arrivals = np.cumsum(gaps)
        print(f"[config5] {rate:g} req/s: plain FIFO", flush=True)
        lat_p, wall_p = _plain_fifo(ctx, pools, rq, arrivals)
        plain.append(_summary(rate, lat_p, wall_p))
        print(f"[config5] {rate:g} req/s: plain batched", flush=True)
        lat_b, wall_b, batch_ms = _plain_batched(ctx, pools, rq, arrivals)
        row = _summary(rate, lat_b, wall_b)
        row.update({"batches": len(batch_ms), "batch_ms_mean": float(np.mean(batch_ms)),
                    "batch_ms_p50": float(np.percentile(batch_ms, 50)),
                    "batch_ms_p99": float(np.percentile(batch_ms, 99))})
        batched.append(row)
        print(f"[config5] {rate:g} req/s: served", flush=True)
        lat_s, wall_s, tickets = _served(srv, pools, rq, arrivals)
        row = _summary(rate, lat_s, wall_s)
        # the serving target: served p99 <= 2x the plain batched path's per-batch time
        row["p99_over_2x_plain_batch_mean"] = row["p99_ms"] / (2.0 * batched[-1]["batch_ms_mean"])
        row["p99_over_2x_plain_batch_p99"] = row["p99_ms"] / (2.0 * batched[-1]["batch_ms_p99"])
        served.append(row)
        if ri == 0:
            # alone vs in the mix: these requests again, one engine call each
            same = True
            for j, (op, ki, n) in enumerate(rq[:30]):
                alone = _call(pools[ki].key, pools[ki], op, n)
                ctx.sync()
                same &= torch.equal(alone, tickets[j][1])
            check_same = bool(same)
    st = srv.stats()
    srv.close()
    for p in pools:
        p.key.close()
    return {
        "workload": "32 resident keys (24x2048, 8x3072); Poisson arrivals; enc/dec/add/mul/dot 30/30/15/15/10; "
                    "lengths log-uniform 1..262144",
        "plain_fifo": plain,
        "plain_batched": batched,
        "served": served,
        "work": work,
        "requests_in_sweep": per_rate * len(rates),
        "server_batches": st["batches"],
        "max_requests_per_batch": st["max_batch_requests"],
        "alone_equals_mix": check_same,
    }


if __name__ == "__main__":
    # python3 -m heu_amd.configs 3 [4 5]: runs the named configurations alone and prints their results
    import json
    import sys

    ctx = E.Context(0)
    out = {c: globals()[f"run_config{c}"](ctx) for c in (sys.argv[1:] or ["3"])}
    print(json.dumps(out))
