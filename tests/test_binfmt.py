"""The index .bin format (a Julia-serialized NamedTuple) without Julia.

Pins available without a reference-written file: the tag table's fixed size
(255 entries) and the "7JL" magic (HEADER tag = 0x37).  Everything else is a
round trip through the writer and reader, the byte structure the format rules
fix (back-references for repeated long strings, inline short strings, raw
column-major array data at a computable offset), and agreement with the other
two index containers on the same data.  Parity with files Julia wrote is
unpinned (no such file exists under the reference)."""
import struct

import numpy as np
import pytest

import rotormap_amd as R
from rotormap_amd import binfmt as J
from rotormap_amd import index as I


def _tiny(n=6, m=1, c=5, heads=None, seed=1):
    rng = np.random.default_rng(seed)
    N, rd = 2 * n, 2 * m * 4 ** c
    if heads is None:
        heads = [">chr1 synthetic record one"] * (N // 2) + [">c2"] * (N - N // 2)
    return I.IndexData(rng.integers(0, 256, (N, rd), dtype=np.uint8),
                       rng.random((N, m), dtype=np.float32), heads,
                       np.arange(1, N + 1, dtype=np.int64) * 7,
                       (np.arange(N) % 2).astype(np.uint8), 20_000, 2_000, 5, m, c, 0, n,
                       "/data/ref.fa", embeds16=rng.standard_normal((N, rd)).astype(np.float16))


def test_tag_table_layout():
    assert J.NTAGS == 255                      # the table's fixed size
    assert J.HEADER == 0x37 == ord("7")        # every serialized file starts "7JL"
    assert J.VALUE_TAGS == 68
    assert J.ZERO64_TAG + 32 == 255 and J.ZERO32_TAG + 33 == J.ZERO64_TAG
    assert (J.TYPE_TAG["Symbol"], J.TYPE_TAG["Int64"], J.TYPE_TAG["DataType"],
            J.TYPE_TAG["Tuple"], J.TYPE_TAG["Array"], J.TYPE_TAG["Module"],
            J.TYPE_TAG["String"]) == (1, 8, 16, 20, 21, 31, 33)
    assert J.VALUE_TAG["false"] == 76 and J.VALUE_TAG["true"] == 77 and J.VALUE_TAG["nothing"] == 78
    assert J.TAG_OF_SYMBOL["Core"] == 155 and J.TAG_OF_SYMBOL["Array"] == 80


def test_fp8_layout_round_trip(tmp_path):
    d = _tiny()
    p = tmp_path / "idx.bin"
    I.save_index_bin(d, p)
    e = I.load_index_bin(p)
    assert np.array_equal(e.embeds8, d.embeds8)
    assert np.array_equal(e.norms, d.norms)
    assert e.heads == d.heads
    assert np.array_equal(e.starts, d.starts) and np.array_equal(e.strand, d.strand)
    assert (e.k, e.kstep, e.s, e.m, e.c, e.n_frag, e.source) == (20_000, 2_000, 5, 1, 5, 6, "/data/ref.fa")
    assert e.rec_names == ["chr1", "c2"] and e.rec_of_col.tolist() == [0] * 6 + [1] * 6
    # the field types load_index asserts on (RotorMap/index/load.jl:62-64)
    v, ver = J.jls_load(str(p))
    assert ver == J.SER_VERSION
    assert list(v.keys()) == ["format", "source", "k", "kstep", "s", "m", "c", "normalize", "fp8",
                              "n_frag", "embeds8", "norms", "heads", "starts", "strand"]
    t = dict(zip(v.keys(), v.types))
    assert t["embeds8"] == J.array_type(J.jtype("Float8_E4M3FN"), 2)
    assert t["embeds8"].params[0].module == J.DLFP8
    assert t["norms"] == J.array_type(J.jtype("Float32"), 2)
    assert t["heads"] == J.array_type(J.jtype("String"), 1)
    assert t["fp8"] == J.jtype("Bool") and t["k"] == J.jtype("Int64")
    assert v["embeds8"].shape == (2048, 12) and v["norms"].shape == (1, 12)


def test_fp8_layout_bytes(tmp_path):
    """Header, the raw (rdim, N) column-major e4m3 block at its offset, inline
    short strings and 3-byte back-references for repeated long ones."""
    d = _tiny()
    p = tmp_path / "idx.bin"
    I.save_index_bin(d, p)
    b = p.read_bytes()
    assert b[:8] == bytes([0x37, ord("J"), ord("L"), J.SER_VERSION, 0x04, 0, 0, 0])
    assert b[8] == J.OBJECT and b[9] == J.TYPE_TAG["DataType"]
    blob = d.embeds8.tobytes()                 # N columns of rdim bytes
    off = b.find(blob)
    assert off > 0 and b.find(blob, off + 1) < 0
    # the array header right before the data: ARRAY, eltype backref, dims (2048, 12)
    dims = bytes([J.TYPE_TAG["Tuple"], 2, J.SHORTINT64]) + struct.pack("<i", 2048) + \
        bytes([J.ZERO64_TAG + 12])
    assert b[off - len(dims):off] == dims
    assert b[off - len(dims) - 4] == J.TYPE_TAG["Array"]
    assert b[off - len(dims) - 3] == J.SHORTBACKREF
    long_head = d.heads[0].encode()
    first = b.find(bytes([J.SHARED_REF, J.TYPE_TAG["String"], len(long_head)]) + long_head)
    assert first > 0 and b.count(long_head) == 1
    slot = struct.unpack_from("<H", b, first + 3 + len(long_head) + 1)[0]
    ref = bytes([J.SHORTBACKREF]) + struct.pack("<H", slot)
    assert b[first + 3 + len(long_head):].startswith(ref * 5)
    assert b.count(bytes([J.TYPE_TAG["String"], 3]) + b">c2") == 6


def test_fp16_layout_fields(tmp_path):
    """indexreal.v1: the verbatim (N, rdim) fp16 rows, column-major."""
    d = _tiny()
    p = tmp_path / "idx16.bin"
    I.save_index_bin(d, p, layout="fp16")
    v, _ = J.jls_load(str(p))
    assert v["format"] == "indexreal.v1" and v["fp16"] is True and "embeds8" not in v
    assert v["embeds"].dtype == np.float16 and v["embeds"].shape == (12, 2048)
    assert np.array_equal(np.ascontiguousarray(v["embeds"]), d.embeds16)
    assert dict(zip(v.keys(), v.types))["embeds"] == J.array_type(J.jtype("Float16"), 2)


def test_agrees_with_dir_and_npz_containers(tmp_path):
    d = _tiny(n=9, m=2, c=4, heads=[f">rec{j // 5} {'x' * (j % 3)}" for j in range(18)])
    I.save_index_bin(d, tmp_path / "a.bin")
    I.save_index_dir(d, tmp_path / "a.dir")
    a, b = I.load_index_bin(tmp_path / "a.bin"), I.load_index_dir(str(tmp_path / "a.dir"))
    for f in ("embeds8", "norms", "starts", "strand", "rec_of_col"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    assert a.heads == b.heads and a.rec_names == b.rec_names
    assert (a.k, a.kstep, a.s, a.m, a.c, a.n_frag) == (b.k, b.kstep, b.s, b.m, b.c, b.n_frag)


def test_wide_backrefs_and_scalars(tmp_path):
    """More than 65,535 table slots: later repeats need BACKREF (Int32)."""
    heads = [f">record_{j:06d}" for j in range(70_000)]
    heads += heads[:10] + heads[-10:]
    N = len(heads)
    d = I.IndexData(np.zeros((N, 8), np.uint8), np.ones((N, 1), np.float32), heads,
                    np.arange(N, dtype=np.int64) - 3, np.zeros(N, np.uint8), 16, 2, 2, 1, 1, 0,
                    N // 2, "s" * 300)
    p = tmp_path / "w.bin"
    I.save_index_bin(d, p)
    b = p.read_bytes()
    assert bytes([J.STRUCT_TAG["BACKREF"]]) in b and bytes([J.LONGSTRING]) in b
    v, _ = J.jls_load(str(p))
    assert v["heads"] == heads and v["source"] == "s" * 300
    assert np.array_equal(v["starts"], d.starts)


def test_scalar_encodings():
    import io
    f = io.BytesIO()
    w = J._Writer(f)
    vals = [0, 32, 33, -1, 2 ** 31 - 1, 2 ** 31, -2 ** 40, True, False, None, np.int32(5),
            np.int32(-7), J.Sym("k"), J.Sym("normalize"), J.Sym("normalize"), "short", "a" * 8,
            "a" * 8, (1, 2, 3), J.jtype("Float8_E4M3FN"), J.jtype("Float8_E4M3FN")]
    for v in vals:
        w.value(v)
    b = f.getvalue()
    r = J._Reader(b)
    got = [r.value() for _ in vals]
    assert r.pos == len(b)
    for a, g in zip(vals, got):
        if isinstance(a, (int, np.integer)) and not isinstance(a, bool):
            assert int(a) == int(g) and isinstance(g, (int, np.integer)), (a, g)
        else:
            assert a == g and type(a) is type(g), (a, g)
    assert isinstance(got[10], np.int32) and isinstance(got[11], np.int32)
    assert b[0] == J.ZERO64_TAG and b[1] == J.ZERO64_TAG + 32 and b[2] == J.SHORTINT64
    # :k has a tag of its own; the second :normalize and the second 8-byte
    # string are back-references to their first slots (0 and 1)
    assert bytes([J.TAG_OF_SYMBOL["k"]]) in b
    sym = bytes([J.TYPE_TAG["Symbol"], 9]) + b"normalize"
    i = b.find(sym) + len(sym)
    assert b[i:i + 3] == bytes([J.SHORTBACKREF, 0, 0])
    st = bytes([J.SHARED_REF, J.TYPE_TAG["String"], 8]) + b"a" * 8
    i = b.find(st) + len(st)
    assert b[i:i + 3] == bytes([J.SHORTBACKREF, 1, 0])


def test_bool_arrays_run_length():
    import io
    f = io.BytesIO()
    w = J._Writer(f)
    v = np.array([True] * 300 + [False] * 2 + [True], dtype=bool)
    w.array(v, J.jtype("Bool"))
    b = f.getvalue()
    a = J._Reader(b).value()
    assert np.array_equal(a, v)
    assert b[:3] == bytes([J.TYPE_TAG["Array"], J.VALUE_TAG["Bool"], J.SHORTINT64])
    runs = b[7:]                              # after ARRAY, Bool, length 303 (Int32)
    # (value << 7) | count, runs of at most 127: 127 + 127 + 46 true, 2 false, 1 true
    assert list(runs) == [0xFF, 0xFF, 0x80 | 46, 2, 0x81]


def test_rejects_bad_files(tmp_path):
    d = _tiny()
    p = tmp_path / "idx.bin"
    I.save_index_bin(d, p)
    b = p.read_bytes()
    with pytest.raises(J.BinFormatError):
        J.jls_load(b"7JX" + b[3:])
    with pytest.raises(J.BinFormatError):
        J.jls_load(b[:8] + b[8:] [:-5000])
    with pytest.raises(J.BinFormatError):
        J.jls_load(b[:4] + b"\x00" + b[5:])     # a big-endian or 32-bit stream
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b.replace(b"indexreal.fp8.v1", b"indexreal.fp9.v1"))
    with pytest.raises(R.RotorMapArgumentError):
        I.load_index_bin(bad)
    d.normalize = 1
    I.save_index_bin(d, bad)
    with pytest.raises(R.RotorMapArgumentError):
        I.load_index_bin(bad)


def test_legacy_tag_and_empty_index(tmp_path):
    """Pre-rename "indexflowreal.fp8.v1" files load (RotorMap/index/load.jl:32-36); an index
    with no windows round-trips as (rdim, 0) arrays."""
    d = _tiny()
    p = tmp_path / "legacy.bin"
    fields = [("format", "indexflowreal.fp8.v1"), ("source", "/x"), ("k", d.k), ("kstep", d.kstep),
              ("s", d.s), ("m", d.m), ("c", d.c), ("normalize", 0), ("fp8", True), ("n_frag", d.n_frag),
              ("embeds8", d.embeds8.T), ("norms", np.ascontiguousarray(d.norms).T),
              ("heads", list(d.heads)), ("starts", d.starts), ("strand", d.strand)]
    et = {"embeds8": J.jtype("Float8_E4M3FN"), "norms": J.jtype("Float32"), "heads": J.jtype("String"),
          "starts": J.jtype("Int64"), "strand": J.jtype("UInt8")}
    J.jls_dump_namedtuple(p, fields, et)
    e = I.load_index_bin(p)
    assert np.array_equal(e.embeds8, d.embeds8) and e.heads == d.heads
    z = I.IndexData(np.zeros((0, 2048), np.uint8), np.zeros((0, 1), np.float32), [],
                    np.zeros(0, np.int64), np.zeros(0, np.uint8), 20_000, 2_000, 5, 1, 5, 0, 0, "")
    I.save_index_bin(z, tmp_path / "empty.bin")
    ez = I.load_index_bin(tmp_path / "empty.bin")
    assert ez.embeds8.shape == (0, 2048) and ez.heads == [] and len(ez.starts) == 0
