"""GPU tests of the tuning parameters and the autotuner: every decryption
shape (lanes per bignum) and table window gives the oracle's bytes, through the
one-op entry points and the work-list batches, before and after autotuning."""
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


def host(t):
    return t.cpu().numpy().view(np.uint32)


LANES = {2048: (2, 4), 3072: (4, 8)}


@pytest.mark.parametrize("bits", [2048, 3072])
def test_decrypt_shapes_and_table_windows(eng, oracle_mod, ctx, bits):
    kd = eng.keygen(bits, 60 + bits // 1024)
    key = eng.Key.from_keygen(ctx, kd, r_bits=256, window_bits=10)
    ok = oracle_mod.Key(kd["n"], kd["hs"], kd["p"], kd["q"])
    rng = np.random.default_rng(bits)
    n = 300
    m = rng.integers(-(2**63), 2**63 - 1, size=n, dtype=np.int64, endpoint=True)
    r = eng.seeded_r(5, 6, n, 256)
    ref_ct = ok.encrypt_i64(m, r)
    for w in (8, 12, 16, 10):
        print(f"[{bits}] table window {w}", flush=True)
        key.table_window = w
        assert key.table_window == w
        ct = key.encrypt_i64(m, r)
        ctx.sync()
        assert np.array_equal(host(ct), ref_ct), w
    for lanes in LANES[bits]:
        print(f"[{bits}] decrypt lanes {lanes}", flush=True)
        key.dec_lanes = lanes
        assert key.dec_lanes == lanes
        pt = key.decrypt(ref_ct, out_words=2)
        ctx.sync()
        assert np.array_equal(host(pt).view(np.int64).ravel(), m)
        b = eng.Batch(ctx)
        out = b.decrypt(key, ref_ct, out_words=2)
        b.run()
        ctx.sync()
        assert np.array_equal(host(out).view(np.int64).ravel(), m)
        b.close()
    from heu_amd._native import HeuError

    with pytest.raises(HeuError):
        key.dec_lanes = 3
    with pytest.raises(HeuError):
        key.table_window = 0
    key.close()


def test_default_window_follows_table_budget(eng, oracle_mod, ctx, monkeypatch):
    """Without a window, a key takes the widest (10..16 bits) whose table of
    ceil(r_bits / w) * 2^w entries of 576 B fits HEU_TABLE_BYTES (default
    4 GiB); encryption gives the oracle's bytes at each."""
    kd = eng.keygen(2048, 72)
    ok = oracle_mod.Key(kd["n"], kd["hs"], kd["p"], kd["q"])
    m = np.arange(-100, 100, dtype=np.int64)
    rng = np.random.default_rng(72)
    cases = [(256, 64 << 20, 12), (256, None, 16), (1024, None, 16), (1024, 1 << 30, 14), (256, 1 << 20, 10)]
    for r_bits, budget, want in cases:
        if budget is None:
            monkeypatch.delenv("HEU_TABLE_BYTES", raising=False)
        else:
            monkeypatch.setenv("HEU_TABLE_BYTES", str(budget))
        print(f"table budget {budget}, r_bits {r_bits}", flush=True)
        key = eng.Key.from_keygen(ctx, kd, r_bits=r_bits, window_bits=0)
        assert key.table_window == want, (r_bits, budget)
        r = rng.integers(0, 2**32, size=(m.size, r_bits // 32), dtype=np.uint64).astype(np.uint32)
        ct = key.encrypt_i64(m, r)
        assert np.array_equal(host(ct), ok.encrypt_i64(m, r))
        key.close()
    monkeypatch.delenv("HEU_TABLE_BYTES", raising=False)


def test_autotune_keeps_results(eng, oracle_mod, ctx):
    kd = eng.keygen(2048, 70)
    key = eng.Key.from_keygen(ctx, kd, r_bits=256, window_bits=10)
    ok = oracle_mod.Key(kd["n"], kd["hs"], kd["p"], kd["q"])
    print("autotune 2048", flush=True)
    rep = key.autotune()
    print(rep, flush=True)
    assert rep["table_window"]["chosen"] == key.table_window
    assert rep["dec_lanes"]["chosen"] == key.dec_lanes
    assert set(rep["grid_multiplier"]) == {"mul", "decrypt"}
    m = np.arange(-500, 500, dtype=np.int64) * 987654321
    r = eng.seeded_r(7, 8, len(m), 256)
    ct = key.encrypt_i64(m, r)
    ctx.sync()
    assert np.array_equal(host(ct), ok.encrypt_i64(m, r))
    k = np.arange(len(m), dtype=np.int64) - 333
    b = eng.Batch(ctx)
    mul = b.mul_i64(key, ct, k)
    dec = b.decrypt(key, ct, out_words=2)
    b.run()
    ctx.sync()
    assert np.array_equal(host(mul), ok.mul_i64(host(ct), k))
    assert np.array_equal(host(dec).view(np.int64).ravel(), m)
    b.close()
    key.close()
