"""Search layer: the resident index, the fused GEMM + top-k engine, the
streamed flow and the recall scorer.

Mirrors RotorMap/search/engine_fp8.jl (TopKEngine :290-337, batch_gemm_topk!
:399-437, map_and_count :654-733) and its FP16 twin
RotorMap/search/engine_fp16.jl; engine `:hip` is this library.  The streamed
flows (topk_flow, rope_topk_stream) are in stream.py.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .encode import RopeEncoder, rdim as enc_rdim

MODES = {"fp8": L.RM_MODE_FP8, "fp8_rescore": L.RM_MODE_FP8_RESCORE, "fp16": L.RM_MODE_FP16}


def _mode(m):
    if isinstance(m, int):
        return m
    if m not in MODES:
        raise L.RotorMapArgumentError(L.RM_ERR_ARG, f"mode must be one of {sorted(MODES)}")
    return MODES[m]


class ResidentIndex:
    """The e4m3 database resident in HBM: B is (N, rdim) uint8 row-major
    (= the reference's (rdim, N) column-major embeds8), host numpy or device
    torch.  fp16=True also keeps the exact fp16 widening (index_database_f16,
    RotorMap/index/load.jl:130-143) for the FP16 path."""

    def __init__(self, B, fp16=False, device=0):
        """B (N, rdim): uint8 e4m3 rows, or float16 rows -- the classic fp16
        index kept verbatim for the FP16 mode (rm_index_create_f16), with its
        e4m3 quantisation for the FP8 modes."""
        self.device = device
        N, rd = B.shape
        verbatim = str(B.dtype) in ("float16", "torch.float16")
        flags = L.RM_INDEX_FP16 if (fp16 and not verbatim) else 0
        if isinstance(B, np.ndarray):
            B = np.ascontiguousarray(B)
            flags |= L.RM_INDEX_HOST
        h = C.c_void_p()
        if verbatim:
            L.check(L.lib().rm_index_create_f16(L.ctx(device), L.ptr(B), rd, N, flags, C.byref(h)))
        else:
            if str(B.dtype) not in ("uint8", "torch.uint8"):
                raise L.RotorMapArgumentError(L.RM_ERR_ARG, f"index rows must be uint8 e4m3 or float16, not {B.dtype}")
            L.check(L.lib().rm_index_create(L.ctx(device), L.ptr(B), rd, N, flags, C.byref(h)))
        self.handle = h
        self.cols, self.rdim = int(N), int(rd)
        self.fp16 = bool(fp16 or verbatim)
        self.verbatim_f16 = verbatim

    def set_rescore(self, dtype="e4m3"):
        """Columns the FP32 rescoring (mode fp8_rescore) reads: "e4m3" (the
        database the candidates were ranked on) or "fp16" (the index's fp16
        columns: for a float16-built index the verbatim ones)."""
        code = {"e4m3": L.RM_E4M3, "fp16": L.RM_F16}[dtype]
        L.check(L.lib().rm_index_set_rescore(self.handle, code))

    @property
    def rescore(self):
        return {L.RM_E4M3: "e4m3", L.RM_F16: "fp16"}[L.lib().rm_index_rescore(self.handle)]

    def close(self):
        if self.handle:
            L.lib().rm_index_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class TopKBatch:
    """One streamed batch's result (RotorMap/search/engine_fp8.jl:265-271):
    vals (nb, k) float32 descending, locs (nb, k) int32 1-based global column
    ids (0 = empty), norms (m, nb), heads, first (1-based)."""
    vals: np.ndarray
    locs: np.ndarray
    norms: np.ndarray = None
    heads: list = field(default_factory=list)
    first: int = 1


class TopKEngine:
    """Reusable fused GEMM + top-k state over one resident database.

    k: top-k size; kprime: candidates kept before FP32 rescoring
    (mode "fp8_rescore"); mode: "fp8_rescore" | "fp8" | "fp16";
    rows_cap: largest batch (reads) per call (the reference's requirement
    1 <= k <= N of RotorMap/search/engine_fp8.jl:313-314 is checked)."""

    def __init__(self, B, k=20, kprime=32, mode="fp8_rescore", rows_cap=2 ** 13, device=0):
        self.mode = _mode(mode)
        self.index = B if isinstance(B, ResidentIndex) else ResidentIndex(
            B, fp16=(self.mode == L.RM_MODE_FP16), device=device)
        N = self.index.cols
        if not (1 <= k <= N):
            raise L.RotorMapArgumentError(L.RM_ERR_ARG, f"need 1 <= k <= N (got k = {k}, N = {N})")
        self.k, self.kprime, self.rows_cap, self.device = int(k), int(max(kprime, k)), int(rows_cap), device

    @property
    def mode_name(self):
        return {v: k for k, v in MODES.items()}[self.mode]

    def _out(self, M):
        import torch
        v = torch.empty(self.k, M, dtype=torch.float32, device=f"cuda:{self.device}").t()
        l = torch.empty(self.k, M, dtype=torch.int32, device=f"cuda:{self.device}").t()
        return v, l

    def batch_gemm_topk(self, Q, Q32=None, stream=None):
        """Top-k of queries Q (M, rdim) (e4m3 uint8 rows; fp16 for mode fp16)
        against the whole database in one fused launch.  Returns device
        (vals, locs) of shape (M, k), column-major like D_val/D_loc."""
        M = Q.shape[0]
        v, l = self._out(M)
        L.check(L.lib().rm_topk_batch(
            L.ctx(self.device), self.index.handle, L.ptr(Q), Q.stride(0), L.ptr(Q32),
            Q32.stride(0) if Q32 is not None else 0, M, self.k, self.kprime, self.mode,
            L.ptr(v), L.ptr(l), M, stream or L.stream_of(self.device)))
        return v, l

    def map_words(self, re, words, n, norms=None, stream=None):
        """Fused encode -> GEMM -> top-k (-> rescore) of n packed fragments
        (device uint32/int32 tensor of n*ceil(k/16) words)."""
        if enc_rdim(re) != self.index.rdim:
            raise L.RotorMapArgumentError(
                L.RM_ERR_SHAPE, f"B has {self.index.rdim} rows; the encoder's real dim is {enc_rdim(re)}")
        v, l = self._out(n)
        L.check(L.lib().rm_map_batch(
            L.ctx(self.device), self.index.handle, L.ptr(words), n, re.k, re.s, re.m, re.c,
            self.k, self.kprime, self.mode, L.ptr(v), L.ptr(l), n, L.ptr(norms),
            stream or L.stream_of(self.device)))
        return v, l

    def map_batches(self, re, words, offs, stream=None):
        """Uneven batches (row offsets offs, nb+1 entries) in one call: one
        work list over the concatenation, cut into launches as one batch of
        that many rows would be (launch_rows)."""
        offs = np.ascontiguousarray(offs, dtype=np.int64)
        n = int(offs[-1])
        v, l = self._out(n)
        L.check(L.lib().rm_map_batches(
            L.ctx(self.device), self.index.handle, L.ptr(words), L.ptr(offs), len(offs) - 1,
            re.k, re.s, re.m, re.c, self.k, self.kprime, self.mode, L.ptr(v), L.ptr(l), n,
            stream or L.stream_of(self.device)))
        return v, l

    def autotune(self, bucket_rows):
        """Time tile shapes, split counts and XCD groupings per batch-size
        bucket at this engine's k / kprime; the context keeps the fastest.
        Returns (tile rows, splits, query groups, ms per launch) per bucket."""
        b = np.ascontiguousarray(bucket_rows, dtype=np.int64)
        tile = np.zeros(len(b), dtype=np.int32)
        best = np.zeros(len(b), dtype=np.int32)
        qg = np.zeros(len(b), dtype=np.int32)
        ms = np.zeros(len(b), dtype=np.float32)
        L.check(L.lib().rm_autotune(L.ctx(self.device), self.index.handle, L.ptr(b), len(b),
                                    self.mode, self.k, self.kprime, L.ptr(tile), L.ptr(best),
                                    L.ptr(qg), L.ptr(ms)))
        return tile, best, qg, ms

    def launch_rows(self, M):
        """Rows per fused launch for a batch of M rows (0: one launch); each
        launch runs with plan(launch rows)."""
        r = C.c_int64()
        kp = self.kprime if self.mode == L.RM_MODE_FP8_RESCORE else self.k
        L.check(L.lib().rm_launch_rows(L.ctx(self.device), self.index.handle, int(M), self.mode, kp,
                                       C.byref(r)))
        return r.value

    def plan(self, M):
        """(tile rows, splits, query groups) one launch of M rows runs with."""
        t, s, g = C.c_int32(), C.c_int32(), C.c_int32()
        kp = self.kprime if self.mode == L.RM_MODE_FP8_RESCORE else self.k
        L.check(L.lib().rm_plan_get(L.ctx(self.device), self.index.handle, int(M), self.mode, kp,
                                    C.byref(t), C.byref(s), C.byref(g)))
        return t.value, s.value, g.value

    def set_plan(self, M, tile_rows=0, splits=0, qgroup=0):
        """Pin (or with all zeros clear) the plan of M's size bucket."""
        L.check(L.lib().rm_plan_set(L.ctx(self.device), int(M), self.mode, int(tile_rows),
                                    int(splits), int(qgroup)))


def count_mapped(locs, read_rec, read_start, col_rec, col_start, dbk, ktop=None):
    """The recall metric of map_and_count (RotorMap/search/engine_fp8.jl:
    692-724): a read maps iff some of its ktop hits lies in the read's record
    and its dbk-window [db_start, db_start+dbk-1] intersects the read window
    [start, start+dbk-1].  Returns total, correct, the first-hit rank
    histogram and the best-overlap totals (inter_sum, inter_cnt)."""
    locs = np.asarray(locs)
    ktop = ktop or locs.shape[1]
    locs = locs[:, :ktop].astype(np.int64)
    valid = locs > 0
    c = np.where(valid, locs - 1, 0)
    rec = np.asarray(col_rec)[c]
    st = np.asarray(col_start)[c]
    rs = np.asarray(read_start)[:, None]
    same = valid & (rec == np.asarray(read_rec)[:, None])
    inter = np.minimum(rs + dbk - 1, st + dbk - 1) - np.maximum(rs, st) + 1
    hit = same & (inter > 0)
    anyhit = hit.any(axis=1)
    first = np.argmax(hit, axis=1)
    rank_hist = np.bincount(first[anyhit], minlength=ktop)[:ktop]
    best = np.where(hit, inter, 0).max(axis=1)
    return dict(total=int(len(locs)), correct=int(anyhit.sum()), rank_hist=rank_hist,
                inter_sum=int(best[anyhit].sum()), inter_cnt=int((best[anyhit] > 0).sum()),
                hit=anyhit)
