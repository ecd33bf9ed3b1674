"""RotorMap index .bin files without Julia: a reader and a writer for the
subset of Julia's `Serialization` stream format that an index uses.

RotorMap saves an index as one serialized NamedTuple: `save(data, file)` is
`serialize(f, data)` (RotorMap/common/dna.jl:23-34), `data` is built by
`_index_data_nt` (RotorMap/index/build.jl:755-768) and `load_index` checks it
(RotorMap/index/load.jl:49-73).  The two layouts:

    indexreal.fp8.v1  format, source, k, kstep, s, m, c, normalize, fp8::Bool,
                      n_frag, embeds8::Matrix{Float8_E4M3FN} (rdim, N),
                      norms::Matrix{Float32} (m, N), heads::Vector{String},
                      starts::Vector{Int}, strand::Vector{UInt8}
    indexreal.v1      ... fp16::Bool ... embeds::Matrix{Float16|Float32} (N, rdim)

(N = 2 n_frag; "indexflowreal.*" are the pre-rename tags, RotorMap/index/
load.jl:30-36.)

Stream format (Julia's stdlib Serialization, the version the reference pins:
julia 1.12.2, RotorMap/Manifest.toml:3).  A stream is a header, then one value.
Every value starts with a tag byte, an index into a fixed 255-entry table:

  1-40    type tags: a value of that type follows (Symbol 1, UInt8 3, Int32 6,
          Int64 8, UInt128 11, Float16 12, Float32 13, Float64 14, DataType 16,
          Tuple 20, Array 21, Module 31, String 33, ...)
  41-57   structural tags: UNDEFREF, BACKREF, LONGBACKREF, SHORTBACKREF,
          LONGTUPLE, LONGSYMBOL, LONGEXPR, LONGSTRING, SHORTINT64,
          FULL_DATATYPE, WRAPPER_DATATYPE, OBJECT, REF_OBJECT, FULL_GLOBALREF,
          HEADER (= 0x37, the '7' of the "7JL" magic), IDDICT, SHARED_REF
  68-78   values: (), Bool, Any, Bottom, typeof(Bottom), Type, svec(),
          Tuple{}, false, true, nothing
  79-165  87 frequent symbols (:Any, :Array, :TypeVar, :Box, :Tuple, ...)
  190-222 Int32 0..32;  223-255 Int64 0..32

A tag below 68 written as a value (a type such as `String` used as an array
element type) is preceded by a 0 byte.  Values the writer has seen before
come back as back-references: arrays, DataTypes, symbols longer than 7 bytes
and strings longer than 7 bytes take a slot in a table, numbered in write
order from 0, and a repeat is SHORTBACKREF + UInt16 slot (BACKREF + Int32,
LONGBACKREF + Int64 above that).  Strings longer than 7 bytes are written as
SHARED_REF + STRING.  A NamedTuple is OBJECT + its DataType (DATATYPE, name
symbol, module, Int32 parameter count, parameters) + its fields in order.  A
module is MODULE + (uuid::UInt128 or nothing) + root name + submodule names +
EMPTYTUPLE.  An array is ARRAY + element type (left out for UInt8) + length
(vector) or dims tuple, then either the raw column-major element bytes (bits
element types) or one value per element.

Parity unpinned: no reference-written .bin (and no Julia) exists to check
these bytes against.  The tag numbering is pinned only by the facts that the
table has exactly 255 entries and that HEADER = 0x37 reproduces the "7JL"
magic of serialized files; the rest restates the published format.  The
writer stamps a header version byte (`SER_VERSION`) that Julia's reader must
accept as not newer than its own; the reader accepts any version.
"""
import mmap
import os
import struct

import numpy as np

# ---- the tag table ---------------------------------------------------------
_TYPE_TAGS = ("Symbol", "Int8", "UInt8", "Int16", "UInt16", "Int32", "UInt32", "Int64",
              "UInt64", "Int128", "UInt128", "Float16", "Float32", "Float64", "Char",
              "DataType", "Union", "UnionAll", "TypeName", "Tuple", "Array", "Expr",
              "LineNumberNode", "__LabelNode__", "GotoNode", "QuoteNode", "CodeInfo",
              "TypeVar", "Box", "MethodInstance", "Module", "Task", "String", "SimpleVector",
              "Method", "GlobalRef", "SlotNumber", "TypedSlot", "NewvarNode", "SSAValue")
_STRUCT_TAGS = ("UNDEFREF", "BACKREF", "LONGBACKREF", "SHORTBACKREF", "LONGTUPLE",
                "LONGSYMBOL", "LONGEXPR", "LONGSTRING", "SHORTINT64", "FULL_DATATYPE",
                "WRAPPER_DATATYPE", "OBJECT", "REF_OBJECT", "FULL_GLOBALREF", "HEADER",
                "IDDICT", "SHARED_REF")
_N_RESERVED_TAGS = 8
_VALUE_NAMES = ("()", "Bool", "Any", "Bottom", "TypeofBottom", "Type", "svec()", "Tuple{}",
                "false", "true", "nothing")
_SYMBOLS = ("Any", "Array", "TypeVar", "Box", "Tuple", "Ptr", "return", "call", "::",
            "Function", "=", "==", "===", "gotoifnot", "A", "B", "C", "M", "N", "T", "S", "X",
            "Y") + tuple("abcdefghijklmnopqrstuvwxyz") + (
            "add_int", "sub_int", "mul_int", "add_float", "sub_float", "new", "mul_float",
            "bitcast", "start", "done", "next", "indexed_iterate", "getfield", "meta", "eq_int",
            "slt_int", "sle_int", "ne_int", "push_loc", "pop_loc", "pop", "arrayset",
            "arrayref", "apply_type", "inbounds", "getindex", "setindex!", "Core", "!", "+",
            "Base", "static_parameter", "convert", "colon", "#self#", "#temp#", "tuple", "")
_N_RESERVED_SLOTS = 24
_N_INT_LITERALS = 33

TYPE_TAG = {n: i + 1 for i, n in enumerate(_TYPE_TAGS)}
_base = len(_TYPE_TAGS) + 1
STRUCT_TAG = {n: _base + i for i, n in enumerate(_STRUCT_TAGS)}
_base += len(_STRUCT_TAGS)
RETURNNODE_TAG, GOTOIFNOT_TAG = _base, _base + 1
VALUE_TAGS = _base + 2 + _N_RESERVED_TAGS
VALUE_TAG = {n: VALUE_TAGS + i for i, n in enumerate(_VALUE_NAMES)}
SYMBOL_TAG0 = VALUE_TAGS + len(_VALUE_NAMES)
SYMBOL_OF_TAG = {SYMBOL_TAG0 + i: s for i, s in enumerate(_SYMBOLS)}
TAG_OF_SYMBOL = {s: t for t, s in SYMBOL_OF_TAG.items()}
ZERO32_TAG = SYMBOL_TAG0 + len(_SYMBOLS) + _N_RESERVED_SLOTS
ZERO64_TAG = ZERO32_TAG + _N_INT_LITERALS
NTAGS = ZERO64_TAG + _N_INT_LITERALS - 1

(UNDEFREF, BACKREF, LONGBACKREF, SHORTBACKREF, LONGTUPLE, LONGSYMBOL, LONGSTRING, SHORTINT64,
 OBJECT, HEADER, SHARED_REF) = (STRUCT_TAG[n] for n in (
    "UNDEFREF", "BACKREF", "LONGBACKREF", "SHORTBACKREF", "LONGTUPLE", "LONGSYMBOL",
    "LONGSTRING", "SHORTINT64", "OBJECT", "HEADER", "SHARED_REF"))
EMPTYTUPLE_TAG = VALUE_TAG["()"]

# header: HEADER 'J' 'L' version flags 0 0 0; flags = little-endian (0) | 64-bit (1 << 2)
SER_VERSION = 26
_FLAGS_LE64 = 0x04

# primitive element types an index uses: Julia name -> (numpy dtype, module)
_BITS = {"Int8": ("<i1", "Core"), "UInt8": ("u1", "Core"), "Int16": ("<i2", "Core"),
         "UInt16": ("<u2", "Core"), "Int32": ("<i4", "Core"), "UInt32": ("<u4", "Core"),
         "Int64": ("<i8", "Core"), "UInt64": ("<u8", "Core"), "Float16": ("<f2", "Core"),
         "Float32": ("<f4", "Core"), "Float64": ("<f8", "Core"), "Bool": ("?", "Core"),
         "Float8_E4M3FN": ("u1", "DLFP8Types"), "Float8_E5M2": ("u1", "DLFP8Types")}
# DLFP8Types' package uuid (RotorMap/Manifest.toml:339-342): the module key of
# Float8_E4M3FN in a serialized element type
DLFP8TYPES_UUID = 0xf4c166784a16415b82efed337c5d6c7c
# DataType names whose parameter list is read after the module
_PARAMETRIC = {"Array": 2, "NamedTuple": 2, "Tuple": None}


class BinFormatError(ValueError):
    """The file is not a Julia-serialized value of the supported subset."""


class Sym(str):
    """A Julia Symbol (distinct from a String)."""


class JModule:
    def __init__(self, path, uuid=None):
        self.path, self.uuid = tuple(path), uuid

    def __eq__(self, o):
        return isinstance(o, JModule) and (self.path, self.uuid) == (o.path, o.uuid)

    def __hash__(self):
        return hash(("mod",) + self.path)

    def __repr__(self):
        return ".".join(self.path)


CORE = JModule(("Core",))
DLFP8 = JModule(("DLFP8Types",), DLFP8TYPES_UUID)


class JType:
    """A Julia DataType: name, module and parameters (types or bits values)."""

    def __init__(self, name, module=CORE, params=()):
        self.name, self.module, self.params = name, module, tuple(params)

    def key(self):
        return ("type", self.name, self.module.path,
                tuple(p.key() if isinstance(p, JType) else (type(p).__name__, p)
                      for p in self.params))

    def __eq__(self, o):
        return isinstance(o, JType) and self.key() == o.key()

    def __hash__(self):
        return hash(self.key())

    def __repr__(self):
        if not self.params:
            return self.name
        return f"{self.name}{{{', '.join(map(repr, self.params))}}}"


def jtype(name, *params):
    mod = DLFP8 if _BITS.get(name, ("", "Core"))[1] == "DLFP8Types" else CORE
    return JType(name, mod, params)


def array_type(elty, ndims):
    return JType("Array", CORE, (elty, ndims))


class JNamedTuple(dict):
    """A NamedTuple: field name -> value, in field order; `types` holds the
    field types in the same order."""
    types = ()


# types written as a bare tag (those in the tag table)
_TAGGED_TYPES = {n: TYPE_TAG[n] for n in ("Symbol", "Int8", "UInt8", "Int16", "UInt16", "Int32",
                                           "UInt32", "Int64", "UInt64", "Int128", "UInt128",
                                           "Float16", "Float32", "Float64", "Char", "DataType",
                                           "String", "Module")}
_TAGGED_TYPES.update({"Bool": VALUE_TAG["Bool"], "Any": VALUE_TAG["Any"]})
_TYPE_OF_TAG = {t: n for n, t in _TAGGED_TYPES.items()}


# ---- reader ----------------------------------------------------------------
class _Reader:
    def __init__(self, buf):
        self.buf = buf
        self.mv = memoryview(buf)
        self.pos = 0
        self.table = {}
        self.counter = 0
        self.version = None

    def u8(self):
        b = self.mv[self.pos]
        self.pos += 1
        return b

    def unpack(self, fmt):
        v = struct.unpack_from(fmt, self.buf, self.pos)
        self.pos += struct.calcsize(fmt)
        return v[0]

    def raw(self, n):
        if self.pos + n > len(self.buf):
            raise BinFormatError("unexpected end of file")
        b = bytes(self.mv[self.pos:self.pos + n])
        self.pos += n
        return b

    def slot(self):
        s = self.counter
        self.counter += 1
        return s

    def value(self):
        return self.tagged(self.u8())

    def tagged(self, b):
        if b == 0:                                   # a type-table entry used as a value
            t = self.u8()
            if t in _TYPE_OF_TAG:
                return JType(_TYPE_OF_TAG[t])
            if t == TYPE_TAG["Tuple"]:
                return JType("Tuple")
            raise BinFormatError(f"unsupported value tag {t}")
        if b >= VALUE_TAGS:
            return self.const(b)
        if b == HEADER:
            if self.raw(2) != b"JL":
                raise BinFormatError("bad header magic")
            self.version = self.u8()
            flags = self.u8()
            if flags != _FLAGS_LE64:
                raise BinFormatError(f"unsupported stream flags 0x{flags:02x} (need little-endian, 64-bit)")
            self.raw(3)
            return self.value()
        if b == TYPE_TAG["Symbol"]:
            return self.symbol(self.u8())
        if b == LONGSYMBOL:
            return self.symbol(self.unpack("<i"))
        if b == TYPE_TAG["String"]:
            return self.string(self.u8())
        if b == LONGSTRING:
            return self.string(self.unpack("<q"))
        if b == SHARED_REF:
            s = self.slot()
            v = self.value()
            self.table[s] = v
            return v
        if b == SHORTBACKREF:
            return self.backref(self.unpack("<H"))
        if b == BACKREF:
            return self.backref(self.unpack("<i"))
        if b == LONGBACKREF:
            return self.backref(self.unpack("<q"))
        if b == SHORTINT64:
            return self.unpack("<i")
        if b == TYPE_TAG["Tuple"]:
            return tuple(self.value() for _ in range(self.u8()))
        if b == LONGTUPLE:
            return tuple(self.value() for _ in range(self.unpack("<i")))
        if b == TYPE_TAG["DataType"]:
            return self.datatype()
        if b == TYPE_TAG["Module"]:
            return self.module()
        if b == TYPE_TAG["Array"]:
            return self.array()
        if b == OBJECT:
            t = self.value()
            if not isinstance(t, JType):
                raise BinFormatError("OBJECT tag not followed by a type")
            return self.obj(t)
        name = _TYPE_OF_TAG.get(b)
        if name in _BITS and name != "Bool":          # a bits value of a tagged type
            return np.frombuffer(self.raw(np.dtype(_BITS[name][0]).itemsize),
                                 dtype=_BITS[name][0])[0]
        if name in ("UInt128", "Int128"):
            v = int.from_bytes(self.raw(16), "little", signed=(name == "Int128"))
            return v
        raise BinFormatError(f"unsupported tag {b} at byte {self.pos - 1}")

    def const(self, b):
        if b >= ZERO64_TAG:
            return b - ZERO64_TAG
        if b >= ZERO32_TAG:
            return np.int32(b - ZERO32_TAG)
        if b in SYMBOL_OF_TAG:
            return Sym(SYMBOL_OF_TAG[b])
        name = _VALUE_NAMES[b - VALUE_TAGS] if b - VALUE_TAGS < len(_VALUE_NAMES) else None
        if name == "()":
            return ()
        if name in ("false", "true"):
            return name == "true"
        if name == "nothing":
            return None
        if name in ("Bool", "Any"):
            return JType(name)
        raise BinFormatError(f"unsupported constant tag {b}")

    def symbol(self, n):
        s = Sym(self.raw(n).decode("utf-8", "surrogateescape"))
        if n > 7:
            self.table[self.slot()] = s
        return s

    def string(self, n):
        return self.raw(n).decode("utf-8", "surrogateescape")

    def backref(self, i):
        if i not in self.table:
            raise BinFormatError(f"back-reference to unknown slot {i}")
        return self.table[i]

    def module(self):
        key = self.value()
        root = self.value()
        if not isinstance(root, Sym):
            raise BinFormatError("module name is not a symbol")
        path = [str(root)]
        while True:
            v = self.value()
            if v == ():
                break
            path.append(str(v))
        return JModule(path, key)

    def datatype(self):
        s = self.slot()
        name = self.value()
        mod = self.value()
        if not isinstance(name, Sym) or not isinstance(mod, JModule):
            raise BinFormatError("malformed DataType")
        t = JType(str(name), mod)
        if str(name) in _PARAMETRIC:
            n = self.unpack("<i")
            t.params = tuple(self.value() for _ in range(n))
        elif str(name) not in _BITS:
            raise BinFormatError(f"unsupported DataType {name} (module {mod})")
        self.table[s] = t
        return t

    def array(self):
        s = self.slot()
        d1 = self.value()
        if isinstance(d1, JType):
            elty, d1 = d1, self.value()
        else:
            elty = JType("UInt8")
        if isinstance(d1, (int, np.integer)):
            dims = (int(d1),)
        elif isinstance(d1, tuple):
            dims = tuple(int(x) for x in d1)
        else:
            raise BinFormatError("array dims are neither an Int nor a tuple")
        n = int(np.prod(dims, dtype=np.int64)) if dims else 1
        if elty.name in _BITS and not elty.params:
            dt = np.dtype(_BITS[elty.name][0])
            if elty.name == "Bool":
                a = self.bool_rle(n)
            else:
                nb = n * dt.itemsize
                if self.pos + nb > len(self.buf):
                    raise BinFormatError("array data runs past the end of the file")
                a = np.frombuffer(self.buf, dtype=dt, count=n, offset=self.pos)
                self.pos += nb
            a = a.reshape(dims, order="F")
        else:
            a = [None] * n
            for i in range(n):
                b = self.u8()
                a[i] = None if b == UNDEFREF else self.tagged(b)
        self.table[s] = a
        return a

    def bool_rle(self, n):
        out = np.empty(n, dtype=bool)
        i = 0
        while i < n:
            b = self.u8()
            c = b & 0x7F
            out[i:i + c] = bool(b >> 7)
            i += c
        return out

    def obj(self, t):
        if t.name != "NamedTuple" or len(t.params) != 2:
            raise BinFormatError(f"unsupported object type {t!r}")
        names = t.params[0]
        r = JNamedTuple()
        r.types = tuple(t.params[1].params) if isinstance(t.params[1], JType) else ()
        for nm in names:
            b = self.u8()
            r[str(nm)] = None if b == UNDEFREF else self.tagged(b)
        return r


def jls_load(path_or_bytes):
    """Deserialize one value.  Bits arrays are zero-copy views of the file
    (memory-mapped, read-only) when given a path."""
    if isinstance(path_or_bytes, (bytes, bytearray, memoryview)):
        buf = path_or_bytes
    else:
        with open(path_or_bytes, "rb") as f:
            buf = mmap.mmap(f.fileno(), 0, access=mmap.ACCESS_READ)
    r = _Reader(buf)
    if len(buf) < 8 or buf[0] != HEADER:
        raise BinFormatError("not a Julia-serialized stream (no 7JL header)")
    v = r.value()
    return v, r.version


# ---- writer ----------------------------------------------------------------
class _Writer:
    def __init__(self, f):
        self.f = f
        self.table = {}
        self.counter = 0

    def w(self, b):
        self.f.write(b)

    def tag(self, t):
        self.f.write(bytes((t,)))

    def as_tag(self, t):
        if t < VALUE_TAGS:
            self.f.write(b"\x00")
        self.tag(t)

    def cycle(self, key):
        off = self.table.get(key)
        if off is not None:
            if off <= 0xFFFF:
                self.tag(SHORTBACKREF)
                self.w(struct.pack("<H", off))
            elif off <= 0x7FFFFFFF:
                self.tag(BACKREF)
                self.w(struct.pack("<i", off))
            else:
                self.tag(LONGBACKREF)
                self.w(struct.pack("<q", off))
            return True
        self.table[key] = self.counter
        self.counter += 1
        return False

    def header(self, version):
        self.tag(HEADER)
        self.w(b"JL" + bytes((version, _FLAGS_LE64, 0, 0, 0)))

    def value(self, v):
        if isinstance(v, bool) or isinstance(v, np.bool_):
            self.tag(VALUE_TAG["true" if v else "false"])
        elif v is None:
            self.tag(VALUE_TAG["nothing"])
        elif isinstance(v, Sym):
            self.symbol(v)
        elif isinstance(v, str):
            self.string(v)
        elif isinstance(v, np.int32):
            n = int(v)
            if 0 <= n < _N_INT_LITERALS:
                self.tag(ZERO32_TAG + n)
            else:
                self.tag(TYPE_TAG["Int32"])
                self.w(struct.pack("<i", n))
        elif isinstance(v, (int, np.integer)):
            self.int64(int(v))
        elif isinstance(v, tuple):
            if len(v) <= NTAGS:
                self.tag(TYPE_TAG["Tuple"])
                self.w(bytes((len(v),)))
            else:
                self.tag(LONGTUPLE)
                self.w(struct.pack("<i", len(v)))
            for x in v:
                self.value(x)
        elif isinstance(v, JType):
            self.datatype(v)
        elif isinstance(v, JModule):
            self.module(v)
        else:
            raise BinFormatError(f"cannot serialize {type(v).__name__}")

    def int64(self, n):
        if 0 <= n < _N_INT_LITERALS:
            self.tag(ZERO64_TAG + n)
        elif -2 ** 31 <= n < 2 ** 31:
            self.tag(SHORTINT64)
            self.w(struct.pack("<i", n))
        else:
            self.tag(TYPE_TAG["Int64"])
            self.w(struct.pack("<q", n))

    def symbol(self, s):
        if s in TAG_OF_SYMBOL:
            self.as_tag(TAG_OF_SYMBOL[s])
            return
        b = s.encode("utf-8", "surrogateescape")
        if len(b) > 7 and self.cycle(("sym", bytes(b))):
            return
        if len(b) <= 255:
            self.tag(TYPE_TAG["Symbol"])
            self.w(bytes((len(b),)))
        else:
            self.tag(LONGSYMBOL)
            self.w(struct.pack("<i", len(b)))
        self.w(b)

    def string(self, s):
        b = s.encode("utf-8", "surrogateescape")
        if len(b) > 7:
            if self.cycle(("str", b)):
                return
            self.tag(SHARED_REF)
        if len(b) <= 255:
            self.tag(TYPE_TAG["String"])
            self.w(bytes((len(b),)))
        else:
            self.tag(LONGSTRING)
            self.w(struct.pack("<q", len(b)))
        self.w(b)

    def module(self, m):
        self.tag(TYPE_TAG["Module"])
        if m.uuid is None:
            self.tag(VALUE_TAG["nothing"])
        else:                                        # the package uuid as a UInt128 value
            self.tag(TYPE_TAG["UInt128"])
            self.w(int(m.uuid).to_bytes(16, "little"))
        self.symbol(Sym(m.path[0]))
        for p in m.path[1:]:
            self.symbol(Sym(p))
        self.tag(EMPTYTUPLE_TAG)

    def datatype(self, t):
        if t.name in _TAGGED_TYPES and not t.params:
            self.as_tag(_TAGGED_TYPES[t.name])
            return
        if self.cycle(t.key()):
            return
        self.tag(TYPE_TAG["DataType"])
        self.symbol(Sym(t.name))
        self.module(t.module)
        if t.params:
            self.w(struct.pack("<i", len(t.params)))
            for p in t.params:
                self.value(p)

    def array(self, a, elty):
        self.cycle(("arr", id(a)))
        self.tag(TYPE_TAG["Array"])
        if elty.name != "UInt8":
            self.datatype(elty)
        if isinstance(a, list):
            self.int64(len(a))
            for x in a:
                self.value(x)
            return
        if a.ndim == 1:
            self.int64(a.shape[0])
        else:
            self.value(tuple(int(x) for x in a.shape))
        if elty.name == "Bool":
            self.bool_rle(np.asarray(a).reshape(-1, order="F"))
            return
        dt = np.dtype(_BITS[elty.name][0])
        data = np.asfortranarray(a).view(dt) if a.dtype != dt else np.asfortranarray(a)
        self.w(memoryview(data.reshape(-1, order="F")).cast("B"))

    def bool_rle(self, v):
        if not len(v):
            return
        last, count = bool(v[0]), 1
        out = bytearray()
        for x in v[1:]:
            x = bool(x)
            if x != last or count == 127:
                out.append((int(last) << 7) | count)
                last, count = x, 1
            else:
                count += 1
        out.append((int(last) << 7) | count)
        self.w(bytes(out))


def _field_type(v, hint):
    if hint is not None:
        return hint
    if isinstance(v, bool):
        return JType("Bool")
    if isinstance(v, str):
        return JType("String")
    if isinstance(v, (int, np.integer)):
        return JType("Int64")
    raise BinFormatError(f"no Julia type for field value {type(v).__name__}")


def jls_dump_namedtuple(path, fields, array_eltypes, version=SER_VERSION):
    """Serialize a NamedTuple: `fields` is an ordered list of (name, value);
    array values need an entry in `array_eltypes` (name -> element JType);
    a list value is a Vector of the element type's values."""
    types = []
    for name, v in fields:
        if name in array_eltypes:
            nd = 1 if isinstance(v, list) else v.ndim
            types.append(array_type(array_eltypes[name], nd))
        else:
            types.append(_field_type(v, None))
    nt = JType("NamedTuple", CORE, (tuple(Sym(n) for n, _ in fields), JType("Tuple", CORE, types)))
    with open(path, "wb") as f:
        w = _Writer(f)
        w.header(version)
        w.tag(OBJECT)
        w.datatype(nt)
        for name, v in fields:
            if name in array_eltypes:
                w.array(v, array_eltypes[name])
            else:
                w.value(v)
