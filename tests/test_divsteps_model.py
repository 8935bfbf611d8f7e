"""CPU model of the single-value inversion kernel (k_inv_single in
heu_amd/csrc/paillier_device.hpp): Bernstein-Yang divsteps, 60 per batch,
on 128 redundant signed 60-bit limbs held two per lane of a 64-lane wave.

The model restates the kernel's arithmetic exactly -- the batched divsteps on
the low 64 bits (sg_divsteps), the limb update with its two carry rounds and
its one-limb shift (sg_lincomb), the multiple of m that makes d and e divisible
(kd, ke), the second wave that computes the next matrix one batch ahead from
the exact low 180 bits, the resolve and the sign logic -- and checks:
  * every limb, carry and 128-bit product stays inside its C type (int64,
    int128), non-top limbs stay in [-2, 2^60 + 1], and the limb that the shift
    drops is 0 (the division by 2^60 is exact);
  * the matrix of each batch has |u| + |v| <= 2^60, |q| + |r| <= 2^60, and
    equals 60 divsteps taken one at a time on the exact integers;
  * the one-batch-ahead matrices equal those computed from the numbers
    themselves;
  * the result is x^-1 mod m for 4096- and 6144-bit moduli (n^2 of seeded
    keys) at edge and random inputs, and a value sharing a factor with m is
    reported as not invertible.
"""
import random

import pytest

BITS = 60
MASK = (1 << BITS) - 1
LIMBS = 128
LANES = 64
M64 = (1 << 64) - 1
I64 = (-(1 << 63), (1 << 63) - 1)
I128 = (-(1 << 127), (1 << 127) - 1)


def in_i64(x):
    assert I64[0] <= x <= I64[1], f"int64 overflow: {x}"
    return x


def to_i64(u):  # (i64) cast of a u64
    u &= M64
    return u - (1 << 64) if u >> 63 else u


def ctz64(x):
    x &= M64
    return (x & -x).bit_length() - 1 if x else 64


def sg_divsteps(delta, f, g):
    """The kernel's batched divsteps on u64 values; returns (delta, u, v, q, r)."""
    f &= M64
    g &= M64
    uu, vv, qq, rr = 1, 0, 0, 1
    i = BITS
    while True:
        z = ctz64(g | ((M64 << i) & M64))
        g >>= z
        uu = (uu << z) & M64
        vv = (vv << z) & M64
        delta += z
        i -= z
        if i == 0:
            break
        if delta > 0:
            delta = -delta
            f, g = g, (-f) & M64
            uu, qq = qq, (-uu) & M64
            vv, rr = rr, (-vv) & M64
        limit = min(1 - delta, i)
        assert limit >= 1
        m = (M64 >> (64 - limit)) & 63
        f32, g32 = f & 0xFFFFFFFF, g & 0xFFFFFFFF
        w = (f32 * g32 * ((f32 * f32 - 2) & 0xFFFFFFFF)) & 0xFFFFFFFF & m
        g = (g + f * w) & M64
        qq = (qq + uu * w) & M64
        rr = (rr + vv * w) & M64
    return delta, to_i64(uu), to_i64(vv), to_i64(qq), to_i64(rr)


def divsteps_one_at_a_time(delta, f, g, n=BITS):
    """The published divstep (Bernstein-Yang 2019, def. 8.1) on exact integers,
    n times; returns (delta, u, v, q, r) with 2^n f' = u f + v g, 2^n g' = q f + r g."""
    u, v, q, r = 1, 0, 0, 1
    for _ in range(n):
        if delta > 0 and g & 1:
            delta, f, g = 1 - delta, g, (g - f) // 2
            u, v, q, r = 2 * q, 2 * r, q - u, r - v
        elif g & 1:
            delta, f, g = 1 + delta, f, (g + f) // 2
            u, v, q, r = 2 * u, 2 * v, q + u, r + v
        else:
            delta, f, g = 1 + delta, f, g // 2
            u, v, q, r = 2 * u, 2 * v, q, r
    return delta, u, v, q, r


class Wave:
    """A number as 128 redundant limbs, lane-major (lane l holds limbs 2l, 2l+1)."""

    @staticmethod
    def from_int(x):
        return [(x >> (BITS * i)) & MASK for i in range(LIMBS)]

    @staticmethod
    def value(limbs):
        return sum(v << (BITS * i) for i, v in enumerate(limbs))


def sg_lincomb(a, b, x, y, k, ml):
    """out = (a x + b y + k m) / 2^60 as the kernel computes it, limb by limb."""
    P = []
    for i in range(LIMBS):
        p = a * x[i] + b * y[i] + k * ml[i]
        assert I128[0] <= p <= I128[1]
        P.append(p)
    lo = [p & MASK for p in P]
    hi = [p >> BITS for p in P]  # arithmetic shift
    out = [0] * LIMBS
    t = [0] * LIMBS
    for lane in range(LANES):
        i0, i1 = 2 * lane, 2 * lane + 1
        top = lane == LANES - 1
        below = 0 if lane == 0 else in_i64(hi[2 * lane - 1])
        t[i0] = in_i64(lo[i0] + below)
        hi0 = in_i64(hi[i0])
        if top:
            in_i64(P[i1])
            t[i1] = in_i64(P[i1] + hi0)  # the top limb is never split
        else:
            t[i1] = in_i64(lo[i1] + hi0)
    s = [0] * LIMBS
    for lane in range(LANES):
        i0, i1 = 2 * lane, 2 * lane + 1
        top = lane == LANES - 1
        hb1_below = 0 if lane == 0 else (t[2 * lane - 1] >> BITS)
        s[i0] = in_i64((t[i0] & MASK) + hb1_below)
        hb0 = t[i0] >> BITS
        s[i1] = in_i64(t[i1] + hb0) if top else in_i64((t[i1] & MASK) + hb0)
    assert s[0] == 0, "the limb the shift drops must be 0 (exact division by 2^60)"
    for lane in range(LANES):
        i0, i1 = 2 * lane, 2 * lane + 1
        out[i0] = s[i1]
        out[i1] = 0 if lane == LANES - 1 else s[2 * (lane + 1)]
    for i in range(LIMBS - 1):
        assert -2 <= out[i] <= (1 << BITS) + 1, f"limb {i} = {out[i]} outside [-2, 2^60 + 1]"
    return out


def low128(l3):
    return (l3[0] + (l3[1] << BITS) + (l3[2] << (2 * BITS))) & ((1 << 128) - 1)


def inv_single(x, m):
    """k_inv_single on canonical x < m: returns x^-1 mod m, or None if the
    kernel would report the value as not invertible."""
    nbits = m.bit_length()
    nbits = (nbits + 31) // 32 * 32  # 32 * W
    ml = Wave.from_int(m)
    f, g = Wave.from_int(m), Wave.from_int(x)
    d, e = [0] * LIMBS, [1] + [0] * (LIMBS - 1)
    minv = pow(m, -1, 1 << 64)
    batches = ((49 * nbits + 57) // 17 + BITS) // BITS
    first = batches * 7 // 10
    delta, done = 1, 0
    delta, mu, mv, mq, mr = sg_divsteps(delta, low128(f[:3]), low128(g[:3]))
    while True:
        todo = first if done == 0 else (4 if done < batches else 8)
        for _ in range(todo):
            F, G = low128(f[:3]), low128(g[:3])
            # wave 0: the next matrix from the low bits after this batch
            F1 = (mu * F + mv * G) & ((1 << 128) - 1)
            G1 = (mq * F + mr * G) & ((1 << 128) - 1)
            nxt = sg_divsteps(delta, (F1 >> BITS) & M64, (G1 >> BITS) & M64)
            for a, b in ((mu, mv), (mq, mr)):
                assert abs(a) + abs(b) <= 1 << BITS
            # wave 1: apply this batch's matrix
            d0, e0 = d[0] & M64, e[0] & M64
            kd = ((-((mu & M64) * d0 + (mv & M64) * e0) * minv) & M64) & MASK
            ke = ((-((mq & M64) * d0 + (mr & M64) * e0) * minv) & M64) & MASK
            zero = [0] * LIMBS
            fv, gv, dv, ev = Wave.value(f), Wave.value(g), Wave.value(d), Wave.value(e)
            nf = sg_lincomb(mu, mv, f, g, 0, zero)
            ng = sg_lincomb(mq, mr, f, g, 0, zero)
            nd = sg_lincomb(mu, mv, d, e, kd, ml)
            ne = sg_lincomb(mq, mr, d, e, ke, ml)
            assert Wave.value(nf) << BITS == mu * fv + mv * gv
            assert Wave.value(ng) << BITS == mq * fv + mr * gv
            assert Wave.value(nd) << BITS == mu * dv + mv * ev + kd * m
            assert Wave.value(ne) << BITS == mq * dv + mr * ev + ke * m
            f, g, d, e = nf, ng, nd, ne
            # the one-batch-ahead matrix equals the one from the numbers themselves
            assert nxt == sg_divsteps(delta, Wave.value(f) & M64, Wave.value(g) & M64)
            delta, mu, mv, mq, mr = nxt
        done += todo
        fv, gv, dv = Wave.value(f), Wave.value(g), Wave.value(d)
        assert fv % m == dv * x % m  # f = d x mod m throughout
        if gv == 0 and abs(fv) == 1:
            return (dv * fv) % m
        if gv == 0 or done >= 4 * batches:
            return None


@pytest.fixture(scope="module")
def moduli():
    from oracle import oracle as O

    O.build()
    return {bits: O.keygen(bits, seed)["n"] ** 2 for bits, seed in ((2048, 0), (3072, 1))}


def test_batched_divsteps_equal_one_at_a_time():
    rng = random.Random(17)
    for _ in range(3000):
        delta = rng.choice([1, 1, 1, rng.randrange(-100, 101)])
        f = rng.getrandbits(rng.choice([64, 200])) | 1
        g = rng.getrandbits(rng.choice([1, 8, 64, 200]))
        got = sg_divsteps(delta, f & M64, g & M64)
        want = divsteps_one_at_a_time(delta, f, g)
        assert got == want, (delta, f, g)


@pytest.mark.parametrize("bits", [2048, 3072])
def test_inversion_model(bits, moduli):
    m = moduli[bits]
    rng = random.Random(bits)
    nb = m.bit_length()
    vals = [1, 2, 3, m - 1, m - 2, 1 << (nb - 1), (1 << (nb - 1)) - 1, (m - 1) // 2,
            int.from_bytes(bytes([0xAA]) * (nb // 8), "little") % m, rng.randrange(1, m), rng.randrange(1, m)]
    for x in vals:
        assert inv_single(x, m) == pow(x, -1, m), hex(x)[:20]


def test_not_invertible(moduli):
    from oracle import oracle as O

    kd = O.keygen(2048, 0)
    m = moduli[2048]
    assert inv_single(kd["p"] * 12345 % m, m) is None
