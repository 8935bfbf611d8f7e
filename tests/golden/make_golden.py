#!/usr/bin/env python3
"""Writes tests/golden/paillier_vectors.json: known-answer vectors for the
Paillier path computed with Python integers only (pow, modular inverse), i.e.
independently of the OpenSSL-based oracle and of the engine.

Closed forms (paillier_zahlen/README.md:10-13, encryptor.cc:42-49,
decryptor.cc:32-53, evaluator.cc:25-123):
  c        = (1 + n*m mod n^2) * h_s^r mod n^2
  Dec(c)   = L(c^lambda mod n^2) * mu mod n, folded to (-n/2, n/2]
  Add      = c1*c2 mod n^2;  AddPt = c*(1+n*p) mod n^2
  Mul(c,k) = c^k mod n^2 (k < 0: (c^-1)^|k|);  Negate = c^-1 mod n^2
  Sub      = c1 * c2^-1 mod n^2;  Randomize(c, r3) = c * h_s^r3 mod n^2
The keys come from the oracle's seeded key generator (key material only);
their invariants are re-checked here with Python integers.

The reference ships no golden vectors for this path (all of its Paillier
tests are randomized round trips), so these are this project's own.
"""
import json
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def is_probable_prime(n: int) -> bool:
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def decrypt(c, n, p, q):
    lam = (p - 1) * (q - 1) // 2
    mu = pow(lam, -1, n)
    u = pow(c, lam, n * n)
    m = (u - 1) // n * mu % n
    return m - n if m > n // 2 else m


def encrypt(m, r, n, hs):
    n2 = n * n
    return (1 + n * m) % n2 * pow(hs, r, n2) % n2


def key_entry(bits, seed):
    from oracle import oracle as O

    O.build()
    kd = O.keygen(bits, seed)
    n, p, q, hs = kd["n"], kd["p"], kd["q"], kd["hs"]
    lam = (p - 1) * (q - 1) // 2
    assert n == p * q and n.bit_length() == bits
    assert p % 4 == 3 and q % 4 == 3 and is_probable_prime(p) and is_probable_prime(q)
    assert pow(hs, lam, n * n) == 1  # h_s = h^n is killed by lambda
    return kd


def vectors(kd, rng, count, rng_extra):
    n, hs, p, q = kd["n"], kd["hs"], kd["p"], kd["q"]
    n2 = n * n
    out = []
    edge = [-(2**63), -12345, -2, -1, 0, 1, 2, 123, 2**63 - 1, n // 2, -(n // 2)]
    for i in range(count):
        m = edge[i] if i < len(edge) else rng.randrange(-(2**63), 2**63)
        m2 = rng.randrange(-(2**63), 2**63)
        r = rng.randrange(1, 2**256)
        r2 = rng.randrange(1, 2**256)
        k = [0, 1, -1, 2, -2][i] if i < 5 else rng.randrange(-(2**31), 2**31)
        c = encrypt(m, r, n, hs)
        c2 = encrypt(m2, r2, n, hs)
        assert decrypt(c, n, p, q) == m
        mul = pow(c, k, n2) if k >= 0 else pow(pow(c, -1, n2), -k, n2)
        r3 = rng_extra.randrange(1, 2**256)
        out.append({
            "m": str(m), "r": hex(r), "c": hex(c),
            "m2": str(m2), "r2": hex(r2), "c2": hex(c2),
            "add": hex(c * c2 % n2),
            "add_pt": hex(c * ((1 + n * m2) % n2) % n2),
            "k": str(k), "mul": hex(mul),
            "negate": hex(pow(c, -1, n2)),
            "sub": hex(c * pow(c2, -1, n2) % n2),
            "r3": hex(r3), "randomize": hex(c * pow(hs, r3, n2) % n2),
        })
    return out


def main():
    rng = random.Random(20261203)
    data = {"description": __doc__.strip().splitlines()[0], "keys": []}
    for bits, seed, count in ((256, 5, 12), (2048, 0, 12), (3072, 1, 6)):
        kd = key_entry(bits, seed)
        data["keys"].append({
            "bits": bits, "seed": seed,
            "n": hex(kd["n"]), "p": hex(kd["p"]), "q": hex(kd["q"]), "hs": hex(kd["hs"]),
            "vectors": vectors(kd, rng, count, random.Random(1000 + bits)),
        })
    with open(os.path.join(HERE, "paillier_vectors.json"), "w") as f:
        json.dump(data, f, indent=1)


if __name__ == "__main__":
    main()
