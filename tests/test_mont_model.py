"""CPU model of the Montgomery core's exact lane schedule
(heu_amd/csrc/mp_group.hpp: mont_core, seminorm) for every group shape the
engine instantiates, with both carry forms.

The kernels keep each 29x29-bit product in a 64-bit accumulator with no carry
chain, so a wrong bound would wrap silently; random-input GPU tests cannot reach
the worst case.  This file restates the schedule step by step -- the same
products per unrolled iteration (the squaring's tournament included), the
quotient broadcast from group lane 0, the limb shift between lanes, the
middle-slot carry of K = 36 -- and checks:
  * interval bounds: with every operand limb at its largest value (< 2^29 + 65
    after seminorm, 2^29 - 1 for the modulus and the quotient), no accumulator
    reaches 2^64 at any step, and seminorm's output limbs stay < 2^29 + 65;
  * exact runs on dense and random operands: every 64-bit step is exact, group
    lane 0 sends 0, the top lane's carry out of seminorm is 0, and the value
    is a*b*R^-1 mod m (R = 2^(29 NL)), below 2m for inputs below 3m.
"""
import math
import random

import pytest

LIMB = 29
MASK = (1 << LIMB) - 1
U64 = 1 << 64
U32 = 1 << 32
NORM_MUL, NORM_SQR = 31, 27  # kNormEvery, kNormEverySq

# (L, K): lanes per group and limbs per lane of every instantiated shape
SHAPES = {
    "n^2 2048 (4x36)": (4, 36),
    "p^2 2048 (2x36)": (2, 36),
    "p^2 2048 (4x18)": (4, 18),
    "n^2 3072 (8x27)": (8, 27),
    "p^2 3072 (4x27)": (4, 27),
    "p^2 3072 (8x14)": (8, 14),
}
MOD_BITS = {"n^2 2048": 4096, "p^2 2048": 2048, "n^2 3072": 6144, "p^2 3072": 3072}


def sq_cross_here(K, x, y):
    d = (y - x + K) % K
    if K & 1:
        return 1 <= d <= (K - 1) // 2
    return (1 <= d < K // 2) or (d == K // 2 and x < K // 2)


def schedule(K, sqr):
    """Per iteration s (0..K-1 within a block): list of (j, weight) products
    a_i * b[j] * weight into register (j + s) % K, as the unrolled loop issues."""
    out = []
    for s in range(K):
        prods = []
        for j in range(K):
            if not sqr:
                prods.append((j, 1))
            elif j == s:
                prods.append((j, 1))
            elif sq_cross_here(K, s, j):
                prods.append((j, 2))
        out.append(prods)
    return out


def test_squaring_tournament_forms_every_cross_term_once():
    for K in {k for _, k in SHAPES.values()}:
        for x in range(K):
            for y in range(K):
                if x != y:
                    assert sq_cross_here(K, x, y) != sq_cross_here(K, y, x), (K, x, y)
        # an iteration issues (K + 1) / 2 product multiply-adds instead of K
        assert all(len(p) == (K + 1) // 2 for p in schedule(K, True)) or K % 2 == 0
        if K % 2 == 0:
            assert sum(len(p) for p in schedule(K, True)) == K * K // 2 + K // 2


def mont_core(a, b_lanes, m_lanes, mprime, L, K, sqr, bound_mode=False):
    """One Montgomery product as the kernel runs it.  a: NL multiplier limbs
    (the LDS operand); b_lanes[l][t], m_lanes[l][t]: lane l's limbs l*K + t.
    In bound_mode every value is an upper bound (a, b, m, q at their largest)
    and the carries propagate bounds.  Returns (acc[l][t], largest value seen)."""
    kNorm = NORM_SQR if sqr else NORM_MUL
    kMid, kSlot = K // 2, K > kNorm
    acc = [[0] * K for _ in range(L)]
    peak = 0
    sched = schedule(K, sqr)

    def add(l, r, v):
        nonlocal peak
        acc[l][r] += v
        assert acc[l][r] < U64, f"accumulator overflow (lane {l}, reg {r})"
        peak = max(peak, acc[l][r])

    for blk in range(L):
        for s in range(K):
            ai = a[blk * K + s]
            assert (ai << 1) < U32  # the doubled limb of a squaring is a 32-bit operand
            for l in range(L):
                for j, wgt in sched[s]:
                    add(l, (j + s) % K, (ai * wgt) * b_lanes[l][j])
            if bound_mode:
                q = MASK
            else:
                q = (((acc[0][s % K] % U32) * mprime) % U32) & MASK
            for l in range(L):
                for j in range(K):
                    add(l, (j + s) % K, q * m_lanes[l][j])
            out0 = [acc[l][s % K] for l in range(L)]
            for l in range(L):
                add(l, (s + 1) % K, out0[l] >> LIMB)
            if not bound_mode:
                assert out0[0] & MASK == 0, "group lane 0 must send 0"
            for l in range(L):
                if bound_mode:
                    acc[l][s % K] = 0 if (L >= 8 and l == L - 1) else MASK
                elif l == L - 1:
                    # quad rotation brings lane 0's low limb (0); wider groups shift in 0
                    acc[l][s % K] = (out0[0] % U32) & MASK if L <= 4 else 0
                else:
                    acc[l][s % K] = (out0[l + 1] % U32) & MASK
            if kSlot:
                for l in range(L):
                    xr, ar = (kMid + s + 1) % K, (kMid + 1 + s + 1) % K
                    x = acc[l][xr]
                    add(l, ar, (x >> 32) * 8)
                    acc[l][xr] = min(x, U32 - 1) if bound_mode else x % U32
    return acc, peak


def seminorm(acc, L, K, bound_mode=False):
    """Lazy accumulators -> limbs (the value unchanged); returns (limbs, top carry)."""
    out = [[0] * K for _ in range(L)]
    carry = [0] * L
    for l in range(L):
        c = 0
        for t in range(K):
            v = acc[l][t] + c
            assert v < U64
            out[l][t] = MASK if bound_mode else v & MASK
            c = v >> LIMB
        carry[l] = c
    for l in range(L):
        cin = 0 if l == 0 else carry[l - 1]
        v = out[l][0] + cin
        assert v < U64
        out[l][0] = MASK if bound_mode else v & MASK
        out[l][1] += v >> LIMB
    return out, carry[L - 1]


@pytest.mark.parametrize("name", list(SHAPES))
@pytest.mark.parametrize("sqr", [False, True])
def test_worst_case_accumulators(name, sqr):
    L, K = SHAPES[name]
    NL = L * K
    big = (1 << LIMB) + 64  # largest operand limb after seminorm
    a = [big] * NL
    b = [[big] * K for _ in range(L)]
    m = [[MASK] * K for _ in range(L)]
    acc, peak = mont_core(a, b, m, 0, L, K, sqr, bound_mode=True)
    assert peak < U64
    slack = math.log2(U64 / peak)
    print(f"{name} {'sqr' if sqr else 'mul'}: largest accumulator 2^{math.log2(peak):.3f} ({slack:.3f} bits spare)")
    out, top = seminorm(acc, L, K, bound_mode=True)
    # seminorm's output: limb 1 of each lane takes the carry in from the lane
    # below on top of a 29-bit limb, and stays below 2^29 + 65
    assert max(max(row) for row in out) < (1 << LIMB) + 65


def to_lanes(x, L, K):
    limbs = [(x >> (LIMB * j)) & MASK for j in range(L * K)]
    return limbs, [[limbs[l * K + t] for t in range(K)] for l in range(L)]


def lanes_value(rows, L, K):
    return sum(rows[l][t] << (LIMB * (l * K + t)) for l in range(L) for t in range(K))


def lazy_limbs(x, L, K, rng):
    """A representation of x with limbs up to 2^29 + 64 at the positions
    seminorm leaves lazy (limb 1 of each lane), where the value allows it."""
    limbs, _ = to_lanes(x, L, K)
    for l in range(L):
        j = l * K + 1
        if j + 1 < L * K and limbs[j] <= 64 and limbs[j + 1] > 0:
            limbs[j] += 1 << LIMB
            limbs[j + 1] -= 1
    assert sum(v << (LIMB * j) for j, v in enumerate(limbs)) == x
    return limbs, [[limbs[l * K + t] for t in range(K)] for l in range(L)]


def moduli(bits, rng):
    """Odd moduli of `bits` bits: dense limbs, sparse limbs, random."""
    top = 1 << (bits - 1)
    return [(1 << bits) - 1, top + 1, top | (MASK << 200) | 1,
            rng.getrandbits(bits) | top | 1, rng.getrandbits(bits) | top | 1]


@pytest.mark.parametrize("name", list(SHAPES))
def test_exact_schedule_is_montgomery(name):
    L, K = SHAPES[name]
    NL = L * K
    bits = MOD_BITS[name.split(" (")[0]]
    R = 1 << (LIMB * NL)
    rng = random.Random(hash(name) & 0xFFFF)
    for m in moduli(bits, rng):
        assert R >= 4 * m
        mprime = (-pow(m, -1, 1 << LIMB)) % (1 << LIMB)
        _, m_lanes = to_lanes(m, L, K)
        cases = [(2 * m - 1, 2 * m - 1), (3 * m - 1, 2 * m - 2), (m - 1, m - 1), (1, 2 * m - 1),
                 (rng.randrange(3 * m), rng.randrange(3 * m)), (((1 << (bits - 1)) - 1), (1 << (bits + 1)) - 1)]
        for x, y in cases:
            for sqr in (False, True):
                yy = x if sqr else y
                a, _ = lazy_limbs(x, L, K, rng)
                _, b = lazy_limbs(yy, L, K, rng)
                if sqr:
                    b = [[a[l * K + t] for t in range(K)] for l in range(L)]
                acc, _ = mont_core(a, b, m_lanes, mprime, L, K, sqr)
                v = lanes_value(acc, L, K)
                assert v % m == x * yy * pow(R, -1, m) % m
                out, top = seminorm(acc, L, K)
                assert top == 0, "the value fits 29 NL bits"
                assert lanes_value(out, L, K) == v
                assert max(max(r) for r in out) < (1 << LIMB) + 65
                assert v < 2 * m, "inputs below 3m give a result below 2m"
