"""Index layer: build the e4m3 window database on the GPU, save / load it.

Windows follow RotorMap/index/build.jl:291-324 (_each_ref_window!): per
record, k-windows at 0-based offsets 0, kstep, 2*kstep, ... fully inside the
record (nwin = (L - k) // kstep + 1; shorter records contribute nothing).
Columns per build batch are [fwd nb | rc nb] with shared heads/starts and
strand 0/1 (:649-657); batch_size=None is one batch over all windows (the
tiny harness layout, RotorMap/common/harness.jl:177-245).  Each column is
e4m3(fp32(fp16(encoder output))) (:730), written by the encoder kernel
straight into the resident layout.
"""
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .encode import RopeEncoder, encode_frag_real_batch, rdim as enc_rdim, revcomp_frag_words

FP8_FORMAT = "indexreal.fp8.v1"


@dataclass
class IndexData:
    """The fields of an indexreal.fp8.v1 index (RotorMap/index/build.jl:755-768);
    embeds8 is stored (N, rdim) row-major = the reference's (rdim, N)
    column-major matrix."""
    embeds8: np.ndarray
    norms: np.ndarray     # (N, m) = the reference's (m, N) column-major
    heads: list
    starts: np.ndarray    # 1-based window starts
    strand: np.ndarray    # 0 fwd, 1 rc
    k: int
    kstep: int
    s: int
    m: int
    c: int
    normalize: int = 0
    n_frag: int = 0
    source: str = ""
    format: str = FP8_FORMAT
    rec_of_col: np.ndarray = None   # 0-based record index per column
    rec_names: list = field(default_factory=list)
    # the classic fp16 rows ("indexreal.v1" embeds, (N, rdim)) when built with
    # keep_f16: the FP16 path's verbatim database (RotorMap/index/load.jl:96-116)
    embeds16: np.ndarray = None

    @property
    def encoder(self):
        return RopeEncoder(k=self.k, s=self.s, m=self.m, c=self.c)


def window_table(ref, k, kstep):
    """(rec index, 0-based start) of every window in file order."""
    recs, starts = [], []
    for j, r in enumerate(ref.records):
        if r.length < k:
            continue
        nwin = (r.length - k) // kstep + 1
        recs.append(np.full(nwin, j, dtype=np.int64))
        starts.append(np.arange(nwin, dtype=np.int64) * kstep)
    if not recs:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    return np.concatenate(recs), np.concatenate(starts)


def build_index(ref, re, kstep=None, batch_size=None, device=0, normalize=0, chunk=1 << 16,
                keep_f16=False):
    """Encode every window of `ref` (both strands) on the GPU into e4m3 columns
    (and, keep_f16, the fp16 rows of the classic layout from the same encoder
    pass: e4m3 = e4m3(fp32(fp16)) of them)."""
    import torch
    k = re.k
    kstep = kstep or max(k // 10, 1)
    rd = enc_rdim(re)
    W = (k + 15) // 16
    wrec, wst = window_table(ref, k, kstep)
    nwin = len(wrec)
    if nwin == 0:
        raise L.RotorMapArgumentError(L.RM_ERR_ARG, "no k-window fits any record")
    bs = nwin if batch_size is None else int(batch_size)
    # column of window g: batch b = g // bs holds columns [2*b*bs, ...): fwd
    # rows first (nb of them), then the rc rows
    g = np.arange(nwin, dtype=np.int64)
    b = g // bs
    nb = np.minimum(bs, nwin - b * bs)
    col_f = 2 * b * bs + (g - b * bs)
    col_r = col_f + nb
    N = 2 * nwin
    dev = f"cuda:{device}"
    emb = torch.empty(N, rd, dtype=torch.uint8, device=dev)
    emb16 = torch.empty(N, rd, dtype=torch.float16, device=dev) if keep_f16 else None
    norms = torch.empty(N, re.m, dtype=torch.float32, device=dev)
    recw = [torch.from_numpy(r.words.view(np.int32)).to(dev) for r in ref.records]
    for lo in range(0, nwin, chunk):
        hi = min(nwin, lo + chunk)
        n = hi - lo
        fw = torch.empty(n * W, dtype=torch.int32, device=dev)
        for j in np.unique(wrec[lo:hi]):
            sel = np.where(wrec[lo:hi] == j)[0]
            st = torch.from_numpy(wst[lo:hi][sel]).to(dev)
            tmp = torch.empty(len(sel) * W, dtype=torch.int32, device=dev)
            L.check(L.lib().rm_extract_windows(L.ctx(device), L.ptr(tmp), L.ptr(recw[j]), L.ptr(st),
                                               len(sel), k, L.stream_of(device)))
            idx = torch.from_numpy(sel).to(dev)
            fw.view(n, W)[idx] = tmp.view(len(sel), W)
        rc = torch.empty_like(fw)
        revcomp_frag_words(rc, fw, k, n)
        for words, cols in ((fw, col_f[lo:hi]), (rc, col_r[lo:hi])):
            e = torch.empty(n, rd, dtype=torch.uint8, device=dev)
            e16 = torch.empty(n, rd, dtype=torch.float16, device=dev) if keep_f16 else None
            nn = torch.empty(n, re.m, dtype=torch.float32, device=dev)
            encode_frag_real_batch(e, nn.t(), re, words, normalize, dest2=e16)
            ci = torch.from_numpy(cols).to(dev)
            emb[ci] = e
            if keep_f16:
                emb16[ci] = e16
            norms[ci] = nn
    torch.cuda.synchronize(device)
    names = [r.name for r in ref.records]
    rec_of_col = np.empty(N, dtype=np.int64)
    starts = np.empty(N, dtype=np.int64)
    strand = np.empty(N, dtype=np.uint8)
    rec_of_col[col_f] = wrec
    rec_of_col[col_r] = wrec
    starts[col_f] = wst + 1
    starts[col_r] = wst + 1
    strand[col_f] = 0
    strand[col_r] = 1
    heads = [ref.records[j].header for j in rec_of_col.tolist()]
    return IndexData(emb.cpu().numpy(), norms.cpu().numpy(), heads, starts, strand, k, kstep,
                     re.s, re.m, re.c, normalize, nwin, "", FP8_FORMAT, rec_of_col, names,
                     emb16.cpu().numpy() if keep_f16 else None)


def save_index(d, path):
    """npz container of the indexreal.fp8.v1 fields (no pickled objects)."""
    uniq = sorted(set(d.heads), key=d.heads.index)
    hid = np.array([uniq.index(h) for h in d.heads], dtype=np.int32) if len(uniq) < 64 else \
        np.unique(np.array(d.heads), return_inverse=True)[1].astype(np.int32)
    if len(uniq) >= 64:
        uniq = list(np.unique(np.array(d.heads)))
    np.savez(path, embeds8=d.embeds8, norms=d.norms, head_ids=hid,
             head_table=np.array(uniq, dtype=np.str_), starts=d.starts, strand=d.strand,
             meta=np.array([d.k, d.kstep, d.s, d.m, d.c, d.normalize, d.n_frag], dtype=np.int64),
             format=np.array(d.format), source=np.array(d.source))


def load_index(path):
    """Inverse of save_index; asserts normalize = 0 like RotorMap/index/load.jl:57-59."""
    z = np.load(path, allow_pickle=False)
    k, kstep, s, m, c, normalize, n_frag = z["meta"].tolist()
    if str(z["format"]) != FP8_FORMAT:
        raise L.RotorMapArgumentError(L.RM_ERR_ARG, f"not an {FP8_FORMAT} index: {path}")
    if normalize != 0:
        raise L.RotorMapArgumentError(L.RM_ERR_ARG, "the mapping flow assumes a normalize = 0 index")
    table = [str(x) for x in z["head_table"]]
    heads = [table[i] for i in z["head_ids"].tolist()]
    names = []
    for h in heads:
        nm = h.split()[0][1:]
        if nm not in names:
            names.append(nm)
    rec_of_col = np.array([names.index(h.split()[0][1:]) for h in heads], dtype=np.int64)
    rd = 2 * m * 4 ** c
    e8 = z["embeds8"]
    assert e8.shape == (2 * n_frag, rd), "embeds8 shape mismatch"
    return IndexData(e8, z["norms"], heads, z["starts"], z["strand"], k, kstep, s, m, c, normalize,
                     n_frag, str(z["source"]), FP8_FORMAT, rec_of_col, names)


def _names_and_recs(heads):
    names, idx, rec = [], {}, np.empty(len(heads), dtype=np.int64)
    for i, h in enumerate(heads):
        nm = h.split()[0][1:] if h.split() else ""
        if nm not in idx:
            idx[nm] = len(names)
            names.append(nm)
        rec[i] = idx[nm]
    return names, rec


def save_index_dir(d, path):
    """Write d in the raw-array layout of julia/export_index.jl (meta.json,
    embeds8.u8, norms.f32, starts.i64, strand.u8, heads.txt)."""
    import json
    import os
    os.makedirs(path, exist_ok=True)
    N, rd = d.embeds8.shape
    norms = np.ascontiguousarray(d.norms, dtype=np.float32)  # (N, m) = the (m, N) column-major
    meta = dict(format=d.format, source=d.source, k=d.k, kstep=d.kstep, s=d.s, m=d.m, c=d.c,
                normalize=d.normalize, n_frag=d.n_frag, rdim=rd, N=N,
                norms_shape=[norms.shape[1], norms.shape[0]])
    with open(os.path.join(path, "meta.json"), "w") as f:
        json.dump(meta, f)
    np.ascontiguousarray(d.embeds8, dtype=np.uint8).tofile(os.path.join(path, "embeds8.u8"))
    norms.tofile(os.path.join(path, "norms.f32"))
    np.asarray(d.starts, dtype=np.int64).tofile(os.path.join(path, "starts.i64"))
    np.asarray(d.strand, dtype=np.uint8).tofile(os.path.join(path, "strand.u8"))
    with open(os.path.join(path, "heads.txt"), "w", encoding="latin-1") as f:
        for h in d.heads:
            f.write(h + "\n")


def load_index_dir(path, device=None):
    """Read an index exported by julia/export_index.jl (or save_index_dir).
    An "indexreal.v1" export (fp16 rows) is quantised to the e4m3 database
    on the GPU, e4m3(fp32(fp16)) as index_database_host does (RotorMap/index/
    load.jl:130-143); normalize must be 0 (RotorMap/index/load.jl:57-59)."""
    import json
    import os
    with open(os.path.join(path, "meta.json")) as f:
        meta = json.load(f)
    fmt, N, rd = meta["format"], int(meta["N"]), int(meta["rdim"])
    if fmt not in (FP8_FORMAT, "indexreal.v1"):
        raise L.RotorMapArgumentError(L.RM_ERR_ARG, f"not an indexreal index: {fmt}")
    if meta["normalize"] != 0:
        raise L.RotorMapArgumentError(L.RM_ERR_ARG, "the mapping flow assumes a normalize = 0 index")
    if rd != 2 * meta["m"] * 4 ** meta["c"]:
        raise L.RotorMapArgumentError(L.RM_ERR_SHAPE, "rdim does not match 2*m*4^c")
    if fmt == FP8_FORMAT:
        e8 = np.fromfile(os.path.join(path, "embeds8.u8"), dtype=np.uint8)
        if e8.size != N * rd:
            raise L.RotorMapArgumentError(L.RM_ERR_SHAPE, "embeds8.u8 size mismatch")
        e8 = e8.reshape(N, rd)
    else:
        e16 = np.fromfile(os.path.join(path, "embeds16.f16"), dtype=np.float16)
        if e16.size != N * rd:
            raise L.RotorMapArgumentError(L.RM_ERR_SHAPE, "embeds16.f16 size mismatch")
        e8 = _quantize_rows(e16.reshape(N, rd), device)
    m_, n_ = meta.get("norms_shape", [meta["m"], N])
    norms = np.fromfile(os.path.join(path, "norms.f32"), dtype=np.float32)
    norms = norms.reshape(n_, m_) if m_ == meta["m"] else norms.reshape(m_, n_).T
    starts = np.fromfile(os.path.join(path, "starts.i64"), dtype=np.int64)
    strand = np.fromfile(os.path.join(path, "strand.u8"), dtype=np.uint8)
    with open(os.path.join(path, "heads.txt"), encoding="latin-1") as f:
        heads = f.read().split("\n")[:N]
    if not (len(heads) == len(starts) == len(strand) == N):
        raise L.RotorMapArgumentError(L.RM_ERR_SHAPE, "meta length mismatch")
    names, rec = _names_and_recs(heads)
    return IndexData(e8, norms, heads, starts, strand, int(meta["k"]), int(meta["kstep"]),
                     int(meta["s"]), int(meta["m"]), int(meta["c"]), 0, int(meta["n_frag"]),
                     meta.get("source", ""), FP8_FORMAT, rec, names)


def _quantize_rows(rows, device=None):
    """(N, rdim) fp16 or fp32 rows -> (N, rdim) e4m3 bytes on the GPU, the
    quantisation of index_to_f8 (RotorMap/index/load.jl:76-94): e4m3 of the
    fp32 value (fp16 rows widen exactly first)."""
    import torch
    from .gemm import f32_to_f8
    dev = torch.device(f"cuda:{device or 0}")
    src = torch.from_numpy(np.ascontiguousarray(rows)).to(dev).reshape(-1)
    dst = torch.empty(src.numel(), dtype=torch.uint8, device=dev)
    f32_to_f8(dst, src)
    return dst.cpu().numpy().reshape(rows.shape)


FP8_BIN_FORMATS = (FP8_FORMAT, "indexflowreal.fp8.v1")
ROW_BIN_FORMATS = ("indexreal.v1", "indexflowreal.v1")


def save_index_bin(d, path, layout="fp8", source=None):
    """Write `d` as RotorMap's index .bin: one Julia-serialized NamedTuple with
    the fields, order and types of `_index_data_nt` (RotorMap/index/
    build.jl:755-768).  layout "fp8" writes indexreal.fp8.v1 (embeds8, the
    (rdim, N) e4m3 matrix); "fp16" writes indexreal.v1 with the verbatim fp16
    rows (`d.embeds16`, (N, rdim)), which needs an index built with keep_f16."""
    from . import binfmt as J
    N, rd = d.embeds8.shape
    if N != 2 * d.n_frag:
        raise L.RotorMapArgumentError(L.RM_ERR_SHAPE, "embeds8 rows != 2 * n_frag")
    if not (len(d.heads) == len(d.starts) == len(d.strand) == N):
        raise L.RotorMapArgumentError(L.RM_ERR_SHAPE, "meta length mismatch")
    norms = np.ascontiguousarray(d.norms, dtype=np.float32)   # (N, m) = Julia (m, N)
    src = d.source if source is None else source
    fields = [("format", FP8_FORMAT if layout == "fp8" else "indexreal.v1"), ("source", src),
              ("k", int(d.k)), ("kstep", int(d.kstep)), ("s", int(d.s)), ("m", int(d.m)),
              ("c", int(d.c)), ("normalize", int(d.normalize))]
    et = {"norms": J.jtype("Float32"), "heads": J.jtype("String"), "starts": J.jtype("Int64"),
          "strand": J.jtype("UInt8")}
    if layout == "fp8":
        fields += [("fp8", True), ("n_frag", int(d.n_frag)),
                   ("embeds8", np.ascontiguousarray(d.embeds8, dtype=np.uint8).T)]
        et["embeds8"] = J.jtype("Float8_E4M3FN")
    elif layout == "fp16":
        if d.embeds16 is None:
            raise L.RotorMapArgumentError(L.RM_ERR_ARG, "layout fp16 needs the fp16 rows (keep_f16)")
        fields += [("fp16", True), ("n_frag", int(d.n_frag)),
                   ("embeds", np.asfortranarray(d.embeds16, dtype=np.float16))]
        et["embeds"] = J.jtype("Float16")
    else:
        raise L.RotorMapArgumentError(L.RM_ERR_ARG, f"unknown layout {layout!r}")
    fields += [("norms", norms.T), ("heads", [str(h) for h in d.heads]),
               ("starts", np.asarray(d.starts, dtype=np.int64)),
               ("strand", np.asarray(d.strand, dtype=np.uint8))]
    J.jls_dump_namedtuple(path, fields, et)


def load_index_bin(path, device=None):
    """Read RotorMap's index .bin (a Julia-serialized NamedTuple) without
    Julia, with the checks of load_index (RotorMap/index/load.jl:49-73).

    indexreal.fp8.v1: embeds8 is a zero-copy view of the memory-mapped file.
    indexreal.v1: the (N, rdim) rows are kept (`embeds16` when fp16) and
    quantised to the e4m3 database on the GPU, as index_database does
    (RotorMap/index/load.jl:156-161)."""
    from . import binfmt as J
    d, _ = J.jls_load(path)
    if not isinstance(d, J.JNamedTuple):
        raise L.RotorMapArgumentError(L.RM_ERR_ARG, f"not an index NamedTuple: {path}")
    fmt = d.get("format")
    if fmt not in FP8_BIN_FORMATS + ROW_BIN_FORMATS:
        raise L.RotorMapArgumentError(L.RM_ERR_ARG,
                                      f"not an indexreal.v1/indexreal.fp8.v1 index: {path}")
    if d["normalize"] != 0:
        raise L.RotorMapArgumentError(
            L.RM_ERR_ARG, f"the mapping flow assumes a normalize = 0 index (got {d['normalize']})")
    k, kstep, s, m, c, n_frag = (int(d[x]) for x in ("k", "kstep", "s", "m", "c", "n_frag"))
    rd = 2 * m * 4 ** c
    N = 2 * n_frag
    types = dict(zip(d.keys(), d.types)) if hasattr(d, "types") else {}
    emb16 = None
    if fmt in FP8_BIN_FORMATS:
        t = types.get("embeds8")
        if t is not None and t != J.array_type(J.jtype("Float8_E4M3FN"), 2):
            raise L.RotorMapArgumentError(L.RM_ERR_ARG, f"embeds8 eltype mismatch ({t!r})")
        e8 = d["embeds8"]
        if e8.shape != (rd, N):
            raise L.RotorMapArgumentError(L.RM_ERR_SHAPE, "embeds8 shape mismatch")
        e8 = e8.T
    else:
        rows = d["embeds"]
        want = np.float16 if d.get("fp16") else np.float32
        if rows.dtype != want:
            raise L.RotorMapArgumentError(L.RM_ERR_ARG, "embeds eltype mismatch")
        if rows.shape != (N, rd):
            raise L.RotorMapArgumentError(L.RM_ERR_SHAPE, "embeds shape mismatch")
        rows = np.ascontiguousarray(rows)
        if rows.dtype == np.float16:
            emb16 = rows
        e8 = _quantize_rows(rows, device)
    heads, starts, strand = d["heads"], np.asarray(d["starts"]), np.asarray(d["strand"])
    if not (len(heads) == len(starts) == len(strand) == N):
        raise L.RotorMapArgumentError(L.RM_ERR_SHAPE, "meta length mismatch")
    norms = np.asarray(d["norms"])
    norms = norms.T if norms.shape == (m, N) else norms
    names, rec = _names_and_recs(heads)
    return IndexData(e8, norms, list(heads), starts, strand, k, kstep, s, m, c, 0, n_frag,
                     str(d.get("source", "")), FP8_FORMAT, rec, names, emb16)
