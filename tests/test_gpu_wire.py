"""The HeMatrix layer's wire format, audit strings and range-checked decryption
on the device, against closed forms.

  serialize / load_from   numpy/matrix.h:221-314 (column-major items, each the
                          variant bytes of a Montgomery-form ciphertext), the
                          interconnection form (numpy/matrix.cc:39-111), and
                          the key round trip of pylib/phe_test.py:279-294 and
                          numpy_test.py:299-321 (serialize on one side, set up
                          from the bytes, decrypt on the other);
  encrypt_with_audit      paillier_zahlen/encryptor.cc:50-54, "p:..,rn:..,c:.."
                          with rn = h_s^r and c in Montgomery form;
  decrypt_in_range        numpy/decryptor.cc:19-45: throws unless every
                          plaintext has at most range_bits bits.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
msgpack = pytest.importorskip("msgpack")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hnp():
    from heu_amd import hnp as H

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return H


@pytest.fixture(scope="module", params=[2048, 3072])
def kit(hnp, request):
    from heu_amd import engine as E

    return hnp.setup(request.param, seed=90 + request.param // 1024, ctx=E.Context(0), r_bits=256, window_bits=10)


def R_of(n2: int) -> int:
    return 1 << (64 * ((n2.bit_length() + 63) // 64))


@pytest.mark.parametrize("shape", [(3, 4), (5,), ()])
def test_matrix_round_trip(hnp, kit, shape):
    from heu_amd import engine as E

    rng = np.random.default_rng(len(shape))
    m = rng.integers(-(2**62), 2**62, size=shape, dtype=np.int64)
    ct = kit.encryptor().encrypt(m, seed=3)
    buf = ct.serialize()
    obj = msgpack.unpackb(buf, raw=True)
    rows = shape[0] if len(shape) else 1
    cols = shape[1] if len(shape) == 2 else 1
    assert obj[:3] == [rows, cols, len(shape)]
    # item k is the C-order element (k % rows) * cols + k // rows, in Montgomery form
    n2 = kit.key.n ** 2
    z = ct.data.cpu().numpy().view(np.uint32)
    for k, item in enumerate(obj[3]):
        assert item[-8:] == (4).to_bytes(8, "little")
        (c_mont,) = msgpack.unpackb(item[:-8], raw=True)
        assert c_mont[0] == 0
        want = E.words_to_int(z[(k % rows) * cols + k // rows]) * R_of(n2) % n2
        assert int.from_bytes(c_mont[1:], "little") == want
    back = hnp.CiphertextArray.load_from(kit, buf)
    assert back.shape == tuple(shape)
    assert torch.equal(back.data, ct.data)
    assert np.array_equal(kit.decryptor().decrypt(back), m)


def test_serialize_ic(hnp, kit):
    # numpy_test.py:323-337: the interconnection bytes differ from the default
    # form, and each format loads back to the same values
    fmt = hnp.MatrixSerializeFormat
    for m in (np.array([[16, 19], [36, 43]], dtype=np.int64), np.array([100], dtype=np.int64),
              np.array(-5, dtype=np.int64)):
        ct = kit.encryptor().encrypt(m, seed=12)
        buf1, buf2 = ct.serialize(fmt.Interconnection), ct.serialize()
        assert buf1 != buf2
        back = hnp.CiphertextArray.load_from(kit, buf1, fmt.Interconnection)
        assert back.shape == m.shape and torch.equal(back.data, ct.data)
        assert np.array_equal(kit.decryptor().decrypt(back), m)
        assert torch.equal(hnp.CiphertextArray.load_from(kit, buf2).data, ct.data)


def test_keys_and_ciphertexts_cross_parties(hnp, kit):
    # client: keys and an encrypted array go out as bytes
    pk_b, sk_b = kit.public_key().serialize(), kit.secret_key().serialize()
    x = np.array([[16, 19], [36, 43]], dtype=np.int64)
    ct_b = kit.encryptor().encrypt(np.array([1, 2], dtype=np.int64), seed=8).serialize()
    # server: a public-key-only kit from the bytes computes x @ ct
    server = hnp.setup(hnp.PublicKey.load_from(pk_b), ctx=kit.key.ctx, r_bits=256, window_bits=10)
    assert not server.key.has_sk
    res_b = server.evaluator().matmul(x, hnp.CiphertextArray.load_from(server, ct_b)).serialize()
    # client again, from the key bytes alone
    client = hnp.setup(hnp.PublicKey.load_from(pk_b), ctx=kit.key.ctx, r_bits=256, window_bits=10,
                       secret_key=hnp.SecretKey.load_from(sk_b))
    res = hnp.CiphertextArray.load_from(client, res_b)
    assert np.array_equal(client.decryptor().decrypt(res), x @ np.array([1, 2]))
    assert hnp.SecretKey.load_from(sk_b) == kit.secret_key()


def test_encrypt_with_audit(hnp, kit):
    from heu_amd import engine as E

    m = np.array([0, 1, -1, 123456789, -(2**63), 2**63 - 1], dtype=np.int64)
    ct, audit = kit.encryptor().encrypt_with_audit(m, seed=11)
    n, hs = kit.key.n, kit.key.hs
    n2 = n * n
    R = R_of(n2)
    r = E.seeded_r(11, 1, len(m), kit.key.r_bits)
    z = ct.data.cpu().numpy().view(np.uint32)
    for i, s in enumerate(audit):
        p_s, rn_s, c_s = s.split(",")
        assert p_s == "p:" + ("-" if m[i] < 0 else "") + format(abs(int(m[i])), "X")
        rv = E.words_to_int(r[i])
        assert rn_s == "rn:" + format(pow(hs, rv, n2) * R % n2, "X")
        c = (1 + n * int(m[i])) % n2 * pow(hs, rv, n2) % n2  # encryptor.cc:42-49
        assert E.words_to_int(z[i]) == c
        assert c_s == "c:" + format(c * R % n2, "X")
    assert np.array_equal(kit.decryptor().decrypt(ct), m)


def test_decrypt_in_range(hnp, kit):
    dec, ev = kit.decryptor(), kit.evaluator()
    m = np.array([5, -7, 2**62, -(2**62)], dtype=np.int64)
    ct = kit.encryptor().encrypt(m, seed=4)
    assert np.array_equal(dec.decrypt_in_range(ct), m)  # default 128 bits
    assert np.array_equal(dec.decrypt_in_range(ct, 63), m)
    with pytest.raises(RuntimeError, match="range check failed"):
        dec.decrypt_in_range(ct, 62)  # |2^62| has 63 bits
    # values beyond 128 bits: (2^62)^3 grows past 2^186
    big = ev.mul(ev.mul(ev.mul(ct, 2**62 - 1), 2**62 - 1), 2**62 - 1)
    with pytest.raises(RuntimeError, match="range check failed"):
        dec.decrypt_in_range(big)
    got = dec.decrypt_in_range(big, 300)
    assert [int(v) for v in got] == [int(v) * (2**62 - 1) ** 3 for v in m]
    small = kit.encryptor().encrypt(np.array([1000, -1023], dtype=np.int64), seed=5)
    assert list(dec.decrypt_in_range(small, 10)) == [1000, -1023]
    with pytest.raises(RuntimeError):
        dec.decrypt_in_range(kit.encryptor().encrypt(np.array([1024], dtype=np.int64), seed=6), 10)
