"""Pins the CPU oracle (oracle/paillier_oracle.c) before it is trusted to
check results: against the golden vectors written from Python closed forms, against
brute force on a toy key, and against the invariants of the reference's own
tests (paillier_zahlen/key_generator_test.cc:21-71, paillier_test.cc:40-233,
phe/encryptor_test.cc:64-86)."""
import json
import math
import os
import random

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def golden():
    with open(os.path.join(HERE, "golden", "paillier_vectors.json")) as f:
        return json.load(f)


def _key(O, k):
    return O.Key(int(k["n"], 16), int(k["hs"], 16), int(k["p"], 16), int(k["q"], 16))


def test_golden_vectors(oracle_mod, golden):
    O = oracle_mod
    for k in golden["keys"]:
        key = _key(O, k)
        n = int(k["n"], 16)
        nw, cw = key.n_words, key.ct_words
        vs = k["vectors"]
        m = [int(v["m"]) for v in vs]
        r = O.ints_to_rows([int(v["r"], 16) for v in vs], 8)
        ct = key.encrypt_twos(np.array([O.int_to_twos(x, nw) for x in m]), r)
        assert O.rows_to_ints(ct) == [int(v["c"], 16) for v in vs]
        dec = key.decrypt(ct)
        assert [O.twos_to_int(row) for row in dec] == m
        c2 = O.ints_to_rows([int(v["c2"], 16) for v in vs], cw)
        assert O.rows_to_ints(key.add(ct, c2)) == [int(v["add"], 16) for v in vs]
        m2 = np.array([int(v["m2"]) for v in vs], np.int64)
        assert O.rows_to_ints(key.add_pt_i64(ct, m2)) == [int(v["add_pt"], 16) for v in vs]
        ks = np.array([int(v["k"]) for v in vs], np.int64)
        assert O.rows_to_ints(key.mul_i64(ct, ks)) == [int(v["mul"], 16) for v in vs]
        assert O.rows_to_ints(key.negate(ct)) == [int(v["negate"], 16) for v in vs]
        # Sub (evaluator.cc:66-68) and Randomize (evaluator.cc:25-28), directly
        assert O.rows_to_ints(key.sub(ct, c2)) == [int(v["sub"], 16) for v in vs]
        r3 = O.ints_to_rows([int(v["r3"], 16) for v in vs], 8)
        assert O.rows_to_ints(key.randomize(ct, r3)) == [int(v["randomize"], 16) for v in vs]
        key.build_table(10, 256)  # the fixed-base table route gives the same values
        assert O.rows_to_ints(key.randomize(ct, r3, use_table=True)) == [int(v["randomize"], 16) for v in vs]
        # int64 fast path agrees where m fits
        fit = [i for i, x in enumerate(m) if -(2**63) <= x < 2**63]
        ct64 = key.encrypt_i64(np.array([m[i] for i in fit], np.int64), r[fit])
        assert np.array_equal(ct64, ct[fit])
        assert n == int(k["p"], 16) * int(k["q"], 16)


def test_keygen_invariants(oracle_mod):
    # key_generator_test.cc:26-71: bit lengths and field relations
    for bits, seed in ((1024, 3), (2048, 0), (3072, 1)):
        kd = oracle_mod.keygen(bits, seed)
        n, p, q, hs = kd["n"], kd["p"], kd["q"], kd["hs"]
        assert n == p * q and n.bit_length() >= bits
        assert abs(p - q).bit_length() >= bits // 2 - 2
        assert math.gcd(p - 1, q - 1) == 2
        assert p % 4 == 3 and q % 4 == 3
        assert 0 < hs < n * n
        lam = (p - 1) * (q - 1) // 2
        assert pow(hs, lam, n * n) == 1
        key = oracle_mod.Key(n, hs, p, q)
        assert key.export("n2") == n * n
        assert key.export("n_half") == n >> 1
        assert key.export("mu") == pow(lam, -1, n)
    # seeded key generation is a function of the seed
    assert oracle_mod.keygen(1024, 9) == oracle_mod.keygen(1024, 9)
    assert oracle_mod.keygen(1024, 9)["n"] != oracle_mod.keygen(1024, 10)["n"]


def test_brute_force_toy_key(oracle_mod):
    O = oracle_mod
    kd = O.keygen(256, 5)
    n, hs, p, q = kd["n"], kd["hs"], kd["p"], kd["q"]
    n2 = n * n
    key = O.Key.from_keygen(kd)
    ms = np.arange(-64, 65, dtype=np.int64)
    rng = random.Random(1)
    rs = [rng.randrange(1, 2**256) for _ in ms]
    ct = key.encrypt_i64(ms, O.ints_to_rows(rs, 8))
    want = [(1 + n * int(m)) % n2 * pow(hs, r, n2) % n2 for m, r in zip(ms, rs)]
    assert O.rows_to_ints(ct) == want
    dec = key.decrypt(ct, out_words=2)
    assert np.array_equal(dec.view(np.int64).ravel(), ms)
    # every product of pairs decrypts to the sum
    sums = key.add(ct, ct[::-1].copy())
    assert np.array_equal(key.decrypt(sums, out_words=2).view(np.int64).ravel(), ms + ms[::-1])


def test_table_equals_direct_powmod(oracle_mod):
    O = oracle_mod
    kd = O.keygen(1024, 4)
    key = O.Key.from_keygen(kd)
    rng = random.Random(2)
    ms = np.array([rng.randrange(-(2**63), 2**63) for _ in range(20)], np.int64)
    rs = O.ints_to_rows([rng.randrange(1, 2**256) for _ in range(20)], 8)
    direct = key.encrypt_i64(ms, rs)
    for unit in (4, 7, 10):
        key.build_table(unit, 256)
        assert np.array_equal(key.encrypt_i64(ms, rs, use_table=True), direct)


def test_min_max_bounds(oracle_mod):
    # paillier_test.cc:99-119 / encryptor_test.cc:64-86: |m| <= n/2 encrypts, beyond throws
    O = oracle_mod
    kd = O.keygen(512, 6)
    key = O.Key.from_keygen(kd)
    n = kd["n"]
    half = n >> 1
    r = O.ints_to_rows([12345], 8)
    nw = key.n_words
    for v in (half, -half):
        ct = key.encrypt_twos(np.array([O.int_to_twos(v, nw)]), r)
        assert O.twos_to_int(key.decrypt(ct)[0]) == v
    for v in (half + 1, -half - 1):
        with pytest.raises(O.OracleError):
            key.encrypt_twos(np.array([O.int_to_twos(v, nw)]), r)
    with pytest.raises(O.OracleError):  # ciphertext >= n^2
        key.decrypt(O.ints_to_rows([n * n], key.ct_words))


def test_mul_special_cases_and_dot(oracle_mod):
    # evaluator.cc:99-123: 0 -> identity, 1 -> a, -1 -> Negate(a)
    O = oracle_mod
    kd = O.keygen(512, 7)
    key = O.Key.from_keygen(kd)
    n2 = kd["n"] ** 2
    ct = key.encrypt_i64(np.array([5, -7, 9], np.int64), O.ints_to_rows([3, 4, 5], 8))
    out = O.rows_to_ints(key.mul_i64(ct, np.array([0, 1, -1], np.int64)))
    c = O.rows_to_ints(ct)
    assert out == [1, c[1], pow(c[2], -1, n2)]
    w = np.array([3, -2, 0], np.int64)
    dot = O.rows_to_ints(key.dot_i64(ct, w, np.array([0, 3, 3], np.int64)))
    assert dot[0] == pow(c[0], 3, n2) * pow(pow(c[1], -1, n2), 2, n2) % n2
    assert dot[1] == 1  # empty segment
    assert key.decrypt(O.ints_to_rows([dot[0]], key.ct_words), out_words=2).view(np.int64)[0, 0] == 15 + 14


def test_mul_big_scalars_closed_form(oracle_mod):
    # Mul by scalars wider than 64 bits: c^k mod n^2, inverse power for k < 0,
    # and Dec(Mul(Enc(m), k)) = m*k when |m*k| <= n/2.
    O = oracle_mod
    kd = O.keygen(512, 9)
    key = O.Key.from_keygen(kd)
    n, n2 = kd["n"], kd["n"] ** 2
    ks = [0, 1, -1, 2**64, -(2**64) - 5, 2**200 + 12345, -(2**230), (n >> 1) - 1]
    ms = [3, -4, 5, -7, 11, 1, -1, 1]
    ct = key.encrypt_i64(np.array(ms, np.int64), O.ints_to_rows(list(range(3, 3 + len(ms))), 8))
    c = O.rows_to_ints(ct)
    out = key.mul(ct, ks)
    assert O.rows_to_ints(out) == [pow(ci, k, n2) for ci, k in zip(c, ks)]
    dec = [O.twos_to_int(r) for r in key.decrypt(out, out_words=key.ct_words // 2)]
    assert dec == [m * k for m, k in zip(ms, ks)]
    small = np.array([2, -3, 0, 1, -1, 7, 9, -9], np.int64)
    assert np.array_equal(key.mul(ct, small.tolist()), key.mul_i64(ct, small))


def test_mont_form_export(oracle_mod):
    O = oracle_mod
    kd = O.keygen(512, 8)
    key = O.Key.from_keygen(kd)
    n2 = kd["n"] ** 2
    ct = key.encrypt_i64(np.array([1, 2], np.int64), O.ints_to_rows([6, 7], 8))
    R = 1 << (64 * ((n2.bit_length() + 63) // 64))
    tm = O.rows_to_ints(key.to_mont(ct, 64))
    assert tm == [c * R % n2 for c in O.rows_to_ints(ct)]
    back = key.to_mont(O.ints_to_rows(tm, key.ct_words), 64, inverse=True)
    assert np.array_equal(back, ct)


def test_threads_do_not_change_results(oracle_mod):
    O = oracle_mod
    kd = O.keygen(1024, 11)
    key = O.Key.from_keygen(kd)
    rng = np.random.default_rng(3)
    ms = rng.integers(-(2**63), 2**63 - 1, size=40, dtype=np.int64)
    rs = O.ints_to_rows([int(x) for x in rng.integers(1, 2**62, size=40)], 8)
    a = key.encrypt_i64(ms, rs, threads=1)
    b = key.encrypt_i64(ms, rs, threads=4)
    assert np.array_equal(a, b)
    assert np.array_equal(key.decrypt(a, threads=1), key.decrypt(a, threads=3))
