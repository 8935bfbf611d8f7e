"""CPU-side checks of the engine library: it loads, exports every entry point
declared in include/heu_paillier.h, the ctypes structures lay out like the
header's, and its host logic (seeded key generation, the counter PRNG, limb
conversions) agrees with the oracle.  No kernel runs."""
import math
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "heu_paillier.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(heu_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def native():
    from heu_amd import _native

    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("engine library not built")
    return _native


def test_library_exports_every_declared_symbol(native):
    L = native.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(native.SIGNATURES), "ctypes signatures and header disagree"


# C struct in the header -> its ctypes mirror in heu_amd/_native.py
_STRUCTS = {"heu_task": "HeuTask", "heu_server_config": "HeuServerConfig", "heu_server_counters": "HeuServerStats"}
_C_SIZES = {"size_t": 8, "uint64_t": 8, "int64_t": 8, "uint32_t": 4, "int32_t": 4}


def header_struct_fields(name):
    """(field, size, align) in declaration order for `typedef struct name {...}`."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = re.search(r"typedef struct " + name + r" \{(.*?)\} " + name + ";", text, re.S).group(1)
    out = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        m = re.match(r"(?:const\s+)?(\w+)\s*(\*?)\s*(.*)", decl)
        ctype, star, names = m.groups()
        for nm in names.split(","):
            nm = nm.strip()
            ptr = bool(star) or nm.startswith("*")
            size = 8 if ptr else _C_SIZES[ctype]
            out.append((nm.lstrip("*"), size, size))
    return out


def test_ctypes_structs_match_header(native):
    import ctypes

    for cname, pyname in _STRUCTS.items():
        fields = header_struct_fields(cname)
        cls = getattr(native, pyname)
        assert [f[0] for f in cls._fields_] == [f[0] for f in fields], cname
        off = 0
        for (fname, size, align), (pname, ptype) in zip(fields, cls._fields_):
            off = (off + align - 1) // align * align
            assert getattr(cls, pname).offset == off and ctypes.sizeof(ptype) == size, (cname, fname)
            off += size
        size = (off + 7) // 8 * 8
        assert ctypes.sizeof(cls) == size, cname


def test_seeded_keygen_matches_oracle(native, oracle_mod):
    from heu_amd import engine as E

    for bits, seed in ((1024, 3), (2048, 0), (3072, 1)):
        a = E.keygen(bits, seed)
        b = oracle_mod.keygen(bits, seed)
        assert all(a[k] == b[k] for k in ("n", "hs", "p", "q")), bits


def test_os_random_keygen_is_valid(native):
    from heu_amd import engine as E

    kd = E.keygen(1024, None)
    n, p, q, hs = kd["n"], kd["p"], kd["q"], kd["hs"]
    assert n == p * q and n.bit_length() == 1024
    assert math.gcd(p - 1, q - 1) == 2 and p % 4 == 3 and q % 4 == 3
    assert pow(hs, (p - 1) * (q - 1) // 2, n * n) == 1
    assert E.keygen(1024, None)["n"] != n


def test_prng_and_r_sampling(native, oracle_mod):
    from heu_amd import engine as E

    w = E.prng_words(5, 7, 50, offset=3)
    assert [int(x) for x in w] == [oracle_mod.lib().orc_prng_word(5, 7, 3 + i) for i in range(50)]
    r = E.seeded_r(1, 11, 1000, 256)
    assert r.shape == (1000, 8)
    vals = [E.words_to_int(row) for row in r]
    assert all(1 <= v < 2**256 for v in vals)
    r100 = E.seeded_r(1, 11, 10, 100)
    assert r100.shape == (10, 4) and all(E.words_to_int(x) < 2**100 for x in r100)
    # the same stream at an offset reproduces the tail
    assert np.array_equal(E.seeded_r(1, 11, 5, 256, offset=3), r[3:8])


def test_limb_conversions(native):
    from heu_amd import engine as E

    vals = [0, 1, -1, 2**255 - 1, -(2**255), 12345678901234567890]
    rows = E.ints_to_words(vals, 8, signed=True)
    assert E.words_to_ints(rows, signed=True) == vals
    with pytest.raises(ValueError):
        E.ints_to_words([-1], 8)


def test_context_without_device_fails_cleanly(native):
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    import ctypes

    h = ctypes.c_void_p()
    rc = native.lib().heu_ctx_create(0, ctypes.byref(h))
    assert rc != native.HEU_OK
    assert native.lib().heu_last_error()
