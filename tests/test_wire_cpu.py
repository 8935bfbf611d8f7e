"""HEU's serialized ZPaillier objects (heu_wire_*, host code, no device).

The bytes are decoded with the msgpack package, an implementation independent
of the engine's writer, and the engine's reader is fed bytes that package
wrote.  Layouts (reference writers): paillier_z::Ciphertext MSGPACK_DEFINE(c_)
(paillier_zahlen/ciphertext.h:37), PublicKey [n, h_s] (public_key.h:64-76),
SecretKey [lambda, mu, p, q] (secret_key.h:64-80), the variant's 8-byte
little-endian schema index, 4 for ZPaillier (phe/base/serializable_types.cc:
50-70, phe/base/schema.h:49), DenseMatrix [rows, cols, ndim, [items as str]]
(numpy/matrix.h:221-314).  Parity unpinned: the inner BigInt bytes (yacl's
msgpack adaptor for BigInt is not part of the reference); they are checked
against this engine's stated layout (sign byte + little-endian magnitude).
"""
import numpy as np
import pytest

msgpack = pytest.importorskip("msgpack")

IDX = (4).to_bytes(8, "little")


@pytest.fixture(scope="module")
def wire():
    from heu_amd import wire as W

    return W


def limbs(v, words):
    return np.frombuffer(v.to_bytes(4 * words, "little"), dtype=np.uint32).copy()


def bigint_bytes(v: int) -> bytes:
    """The stated BigInt layout: sign byte, then the little-endian magnitude."""
    mag = abs(v).to_bytes((abs(v).bit_length() + 7) // 8, "little")
    return bytes([1 if v < 0 else 0]) + mag


def is_str_header(b: int) -> bool:
    return 0xA0 <= b <= 0xBF or b in (0xD9, 0xDA, 0xDB)


@pytest.mark.parametrize("v", [0, 1, 255, 256, (1 << 4095) + 12345, (1 << 4096) - 1, 7 << 3000])
def test_ciphertext_layout(wire, v):
    b = wire.put_ciphertext(limbs(v, 128))
    assert b[-8:] == IDX
    obj = msgpack.unpackb(b[:-8], raw=True)
    assert obj == [bigint_bytes(v)]
    assert b[0] == 0x91 and is_str_header(b[1])  # fixarray(1) holding a str
    assert int.from_bytes(wire.get_ciphertext(b, 128).tobytes(), "little") == v
    # bytes from an independent writer parse to the same value
    other = msgpack.packb([bigint_bytes(v)], use_bin_type=False) + IDX
    assert other == b
    assert int.from_bytes(wire.get_ciphertext(other, 128).tobytes(), "little") == v


def test_ciphertext_rejects_bad_input(wire):
    from heu_amd._native import HeuError

    good = wire.put_ciphertext(limbs(12345, 128))
    for bad in (good[:-1], good[:-8] + (3).to_bytes(8, "little"), good[1:], msgpack.packb([b"\x00\x01", b"\x00"]) + IDX,
                msgpack.packb([b"\x02\x01"], use_bin_type=False) + IDX, b""):
        with pytest.raises(HeuError):
            wire.get_ciphertext(bad, 128)
    # a value wider than the limbs given is refused
    wide = wire.put_ciphertext(limbs((1 << 6000) + 1, 192))
    with pytest.raises(HeuError):
        wire.get_ciphertext(wide, 128)


@pytest.mark.parametrize("rows,cols,ndim", [(3, 1, 1), (2, 3, 2), (1, 1, 0), (300, 1, 1), (0, 1, 1)])
def test_matrix_layout(wire, rows, cols, ndim):
    rng = np.random.default_rng(rows * 7 + cols)
    vals = [int.from_bytes(rng.bytes(512), "little") >> int(rng.integers(0, 4000)) for _ in range(rows * cols)]
    mont = np.stack([limbs(v, 128) for v in vals]) if vals else np.zeros((0, 128), np.uint32)
    b = wire.put_matrix(mont, rows, cols, ndim)
    obj = msgpack.unpackb(b, raw=True)
    assert obj[:3] == [rows, cols, ndim] and len(obj) == 4
    assert [msgpack.unpackb(item[:-8], raw=True) for item in obj[3]] == [[bigint_bytes(v)] for v in vals]
    assert all(item[-8:] == IDX for item in obj[3])
    r, c, d, got = wire.get_matrix(b, 128)
    assert (r, c, d) == (rows, cols, ndim) and np.array_equal(got, mont)
    # an independently written matrix (items as msgpack str) reads back.  The
    # msgpack package writes str of 32..255 bytes as raw16 (0xda) where
    # msgpack-c, like this writer, uses str8 (0xd9); both read the same
    items = [msgpack.packb([bigint_bytes(v)], use_bin_type=False) + IDX for v in vals]
    other = msgpack.packb([rows, cols, ndim, items], use_bin_type=False)
    if all(len(x) < 32 or len(x) >= 256 for x in items + [bigint_bytes(v) for v in vals]):
        assert other == b
    assert np.array_equal(wire.get_matrix(other, 128)[3], mont)


def test_matrix_rejects_bad_shapes(wire):
    from heu_amd._native import HeuError

    item = msgpack.packb([bigint_bytes(5)], use_bin_type=False) + IDX
    for obj in ([2, 2, 2, [item] * 3], [2, 1, 3, [item] * 2], [2, 1, 1], [-1, 1, 1, []]):
        with pytest.raises(HeuError):
            wire.get_matrix(msgpack.packb(obj, use_bin_type=False), 128)


def test_keys(wire):
    from heu_amd import engine as E
    from heu_amd._native import HeuError

    kd = E.keygen(2048, 5)
    n, hs, p, q = kd["n"], kd["hs"], kd["p"], kd["q"]
    pk = wire.put_public_key(n, hs)
    assert pk[-8:] == IDX
    assert msgpack.unpackb(pk[:-8], raw=True) == [bigint_bytes(n), bigint_bytes(hs)]
    assert wire.get_public_key(pk) == (n, hs)
    sk = wire.put_secret_key(p, q)
    assert sk[-8:] == IDX
    lam = (p - 1) * (q - 1) // 2  # key_generator.cc:26-60
    mu = pow(lam, -1, n)
    assert msgpack.unpackb(sk[:-8], raw=True) == [bigint_bytes(x) for x in (lam, mu, p, q)]
    assert wire.get_secret_key(sk) == (p, q)
    # lambda or mu that do not belong to p, q are refused
    tampered = msgpack.packb([bigint_bytes(x) for x in (lam + 2, mu, p, q)], use_bin_type=False) + IDX
    with pytest.raises(HeuError):
        wire.get_secret_key(tampered)
    with pytest.raises(HeuError):
        wire.get_public_key(pk[:-8] + (1).to_bytes(8, "little"))


# ---- interconnection form (numpy/matrix.cc:39-111) ---------------------------
# The org_interconnection .proto is not in the reference tree; the messages
# below restate the field numbers the engine states (wire.cpp), and protobuf's
# own implementation encodes and decodes against them, so the encoding itself
# (tags, varints, packing, nesting, field order) is checked independently.
# Parity unpinned: the field numbers and SCALAR_TYPE_OBJECT's value.
IC_OBJECT = 12


@pytest.fixture(scope="module")
def ic_pb():
    pytest.importorskip("google.protobuf")
    from google.protobuf import descriptor_pb2, descriptor_pool, message_factory

    fd = descriptor_pb2.FileDescriptorProto(name="ic_test.proto", package="ic_test", syntax="proto3")
    nd = fd.message_type.add(name="NdArray")
    F = descriptor_pb2.FieldDescriptorProto
    nd.field.add(name="shape", number=1, type=F.TYPE_INT64, label=F.LABEL_REPEATED)
    nd.field.add(name="items", number=3, type=F.TYPE_BYTES, label=F.LABEL_REPEATED)
    nd.field.add(name="item_buffer", number=2, type=F.TYPE_BYTES, label=F.LABEL_OPTIONAL)  # unknown to the engine
    dep = fd.message_type.add(name="DataExchangeProtocol")
    dep.field.add(name="scalar_type", number=1, type=F.TYPE_INT32, label=F.LABEL_OPTIONAL)
    dep.field.add(name="scalar_type_name", number=2, type=F.TYPE_STRING, label=F.LABEL_OPTIONAL)
    dep.oneof_decl.add(name="container")
    dep.field.add(name="v_other", number=3, type=F.TYPE_BYTES, label=F.LABEL_OPTIONAL, oneof_index=0)
    dep.field.add(name="v_ndarray", number=4, type=F.TYPE_MESSAGE, type_name=".ic_test.NdArray",
                  label=F.LABEL_OPTIONAL, oneof_index=0)
    pool = descriptor_pool.DescriptorPool()
    pool.Add(fd)
    get = message_factory.GetMessageClass
    return get(pool.FindMessageTypeByName("ic_test.DataExchangeProtocol"))


def ic_message(ic_pb, rows, cols, ndim, items, scalar_type=IC_OBJECT):
    m = ic_pb(scalar_type=scalar_type, scalar_type_name="ciphertext")
    m.v_ndarray.SetInParent()
    m.v_ndarray.shape.extend([rows, cols][:ndim])
    m.v_ndarray.items.extend(items)
    return m


@pytest.mark.parametrize("rows,cols,ndim", [(2, 3, 2), (3, 1, 1), (1, 1, 0), (0, 1, 1), (200, 1, 1), (0, 4, 2)])
def test_matrix_ic_layout(wire, ic_pb, rows, cols, ndim):
    rng = np.random.default_rng(rows * 11 + cols + ndim)
    vals = [int.from_bytes(rng.bytes(512), "little") >> int(rng.integers(0, 4000)) for _ in range(rows * cols)]
    mont = np.stack([limbs(v, 128) for v in vals]) if vals else np.zeros((0, 128), np.uint32)
    b = wire.put_matrix_ic(mont, rows, cols, ndim)
    items = [wire.put_ciphertext(limbs(v, 128)) for v in vals]
    # protobuf writes the same bytes for the same message, and reads ours back
    want = ic_message(ic_pb, rows, cols, ndim, items)
    assert want.SerializeToString() == b
    got = ic_pb.FromString(b)
    assert got.scalar_type == IC_OBJECT and got.scalar_type_name == "ciphertext"
    assert got.WhichOneof("container") == "v_ndarray"
    assert list(got.v_ndarray.shape) == [rows, cols][:ndim] and list(got.v_ndarray.items) == items
    r, c, d, back = wire.get_matrix_ic(b, 128)
    assert (r, c, d) == (rows, cols, ndim) and np.array_equal(back, mont)
    # the msgpack form differs (numpy_test.py:323-327)
    assert wire.put_matrix(mont, rows, cols, ndim) != b


def test_matrix_ic_reader_is_lenient_where_protobuf_is(wire, ic_pb):
    vals = [5, 7, 11, 13, 17, 19]
    items = [wire.put_ciphertext(limbs(v, 128)) for v in vals]
    mont = np.stack([limbs(v, 128) for v in vals])
    m = ic_message(ic_pb, 2, 3, 2, items)
    m.v_ndarray.item_buffer = b"ignored"  # a field the engine does not know
    assert np.array_equal(wire.get_matrix_ic(m.SerializeToString(), 128)[3], mont)

    # by hand: fields in reverse order, shape unpacked, a two-byte varint length
    def tag(f, w):
        return bytes([(f << 3) | w])

    nd = b"".join(tag(3, 2) + bytes([len(it)]) + it for it in items) + tag(1, 0) + b"\x02" + tag(1, 0) + b"\x03"
    assert len(nd) < 128
    msg = tag(4, 2) + bytes([len(nd) | 0x80, 0]) + nd + tag(2, 2) + b"\x0aciphertext" + tag(1, 0) + bytes([IC_OBJECT])
    assert ic_pb.FromString(msg).v_ndarray.shape == [2, 3]  # protobuf reads it too
    assert wire.matrix_shape_ic(msg) == (2, 3, 2)
    assert np.array_equal(wire.get_matrix_ic(msg, 128)[3], mont)


def test_matrix_ic_rejects(wire, ic_pb):
    from heu_amd._native import HeuError

    items = [wire.put_ciphertext(limbs(v, 128)) for v in (1, 2, 3)]
    bad = [
        ic_message(ic_pb, 3, 1, 1, items, scalar_type=1).SerializeToString(),  # not OBJECT
        ic_message(ic_pb, 2, 1, 1, items).SerializeToString(),  # shape and len do not match
        ic_message(ic_pb, 3, 1, 1, items[:2] + [b"\x91\xa1\x00"]).SerializeToString(),  # item not a ciphertext
        ic_pb(scalar_type=IC_OBJECT, v_other=b"x").SerializeToString(),  # another container
        ic_pb(scalar_type=IC_OBJECT).SerializeToString(),  # no container
        ic_message(ic_pb, 3, 1, 1, items).SerializeToString()[:-3],  # truncated
        b"\xff\xff\xff",
    ]
    three_d = ic_message(ic_pb, 1, 1, 2, items[:1])
    three_d.v_ndarray.shape.append(1)
    bad.append(three_d.SerializeToString())
    for b in bad:
        with pytest.raises(HeuError):
            wire.get_matrix_ic(b, 128)
    with pytest.raises(HeuError):  # a 0-d matrix is 1 x 1
        wire.put_matrix_ic(np.zeros((2, 128), np.uint32), 2, 1, 0)


def test_shapes_whose_product_wraps_are_refused(wire, ic_pb):
    """rows x cols is compared with the item count without wrapping around:
    2^32 x 2^32 is not 0 items."""
    from heu_amd._native import HeuError

    with pytest.raises(HeuError):
        wire.get_matrix(msgpack.packb([2**32, 2**32, 2, []], use_bin_type=False), 128)
    with pytest.raises(HeuError):
        wire.get_matrix(msgpack.packb([2**62, 4, 2, []], use_bin_type=False), 128)
    with pytest.raises(HeuError):
        wire.get_matrix_ic(ic_message(ic_pb, 2**32, 2**32, 2, []).SerializeToString(), 128)
    # an empty matrix with one zero dimension is still fine
    assert wire.matrix_shape(msgpack.packb([0, 2**40, 2, []], use_bin_type=False)) == (0, 2**40, 2)


def test_secret_key_primes_must_exceed_one(wire):
    """Primes of a serialized secret key must be integers above 1, whatever the
    lambda and mu beside them."""
    from heu_amd._native import HeuError

    for p, q in ((-7, -11), (1, 7), (0, 5)):
        lam = (p - 1) * (q - 1) // 2
        mu = 0
        sk = msgpack.packb([bigint_bytes(x) for x in (lam, mu, p, q)], use_bin_type=False) + IDX
        with pytest.raises(HeuError):
            wire.get_secret_key(sk)
