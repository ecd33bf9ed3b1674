"""Streaming flows: FASTA file -> batches of encodings or of top-k hits.

Mirrors RotorMap/encode/stream.jl (RopeRealBatch :83-98,
rope_encode_real_stream :143-248) and RotorMap/search/engine_fp8.jl
(topk_flow :498-547, rope_topk_stream :576-594).  The reference pipes host
parser threads -> pinned gather -> H2D -> encoder -> D2H through channels;
here the FASTA text is parsed and 2-bit packed on the GPU
(fasta.iter_fasta_gpu), so the packed words never leave the device:

  rope_encode_real_stream  file -> RopeRealBatch (host embeddings, norms,
                           headers, 1-based index of the batch's first read)
  topk_flow                RopeRealBatch source -> TopKBatch: the reference's
                           second stage (fp16 embeddings quantised to e4m3 on
                           the device, then one fused GEMM + top-k launch)
  rope_topk_stream         file -> TopKBatch; by default the fused path
                           (encode -> GEMM -> top-k (-> rescore) on the
                           device, rm_map_batch), or with fused=False the two
                           reference stages above

Batches come in file order (the GPU parser keeps it); every row is
self-describing through its header, as in the reference.
"""
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .encode import RopeEncoder, encode_frag_real_batch, rdim as enc_rdim
from .fasta import iter_fasta_gpu
from .search import TopKBatch, TopKEngine


@dataclass
class RopeRealBatch:
    """One encoded batch (RotorMap/encode/stream.jl:83-98): embeds (nb, 2D)
    host matrix (float32 or float16, column-major like the reference's),
    norms (m, nb) float32, heads (nb headers), first (1-based index of the
    batch's first fragment in the stream)."""
    embeds: np.ndarray
    norms: np.ndarray
    heads: list
    first: int


def _check_encoder_args(re, k, batch_size, normalize):
    if k < re.s:
        raise L.RotorMapArgumentError(L.RM_ERR_ARG, f"k = {k} < s = {re.s}: no s-mer windows")
    if k != re.k:
        raise L.RotorMapArgumentError(L.RM_ERR_ARG, f"k = {k} must equal the encoder baseline re.k = {re.k}")
    if batch_size < 1:
        raise L.RotorMapArgumentError(L.RM_ERR_ARG, "batch_size must be >= 1")
    if normalize not in (0, 1, 2, 3):
        raise L.RotorMapArgumentError(L.RM_ERR_ARG, "normalize must be 0..3")


def rope_encode_real_stream(re: RopeEncoder, file, k=None, batch_size=2 ** 13, normalize=0,
                            fp16=False, device=0, chunk_bytes=1 << 30):
    """Yield RopeRealBatch for every batch of <= batch_size fragments of
    `file` (records trimmed to k = re.k bases, shorter ones skipped): the
    reference's rope_encode_real_stream (RotorMap/encode/stream.jl:143-248)."""
    import torch
    k = re.k if k is None else int(k)
    _check_encoder_args(re, k, batch_size, normalize)
    rd = enc_rdim(re)
    dt = torch.float16 if fp16 else torch.float32
    dev = f"cuda:{device}"
    for words, heads, first in iter_fasta_gpu(file, k, batch_size, chunk_bytes, device):
        nb = words.shape[0]
        dest = torch.zeros(rd, nb, dtype=dt, device=dev).t()  # (nb, 2D) column-major
        norms = torch.zeros(nb, re.m, dtype=torch.float32, device=dev).t()  # (m, nb)
        encode_frag_real_batch(dest, norms, re, words.reshape(-1), normalize)
        emb = dest.cpu().numpy()  # (nb, 2D), Fortran order
        yield RopeRealBatch(emb, np.ascontiguousarray(norms.cpu().numpy()), list(heads), first)


def topk_flow(source, eng: TopKEngine, re: RopeEncoder = None):
    """Per batch of `source`, one fused GEMM + top-k launch against the
    engine's resident index, yielding TopKBatch (host vals (nb, k) float32
    descending, locs (nb, k) int32 1-based global column ids, norms, heads,
    first).  RotorMap/search/engine_fp8.jl:498-547.

    source yields either RopeRealBatch with float16 embeddings (the
    reference's stage: the embeddings are quantised to e4m3 on the device,
    e4m3(fp16), and rescoring uses their fp32 widening) or (words, heads,
    first) word batches, device or host uint32/int32 (nb, ceil(k/16)), which
    take the fused encode -> GEMM -> top-k path and need `re`."""
    import torch
    dev = f"cuda:{eng.device}"
    rd = eng.index.rdim
    for bat in source:
        if isinstance(bat, RopeRealBatch):
            emb = bat.embeds
            if emb.dtype != np.float16:
                raise L.RotorMapArgumentError(
                    L.RM_ERR_ARG, f"batch eltype {emb.dtype} != float16 (build the rope stream with "
                                  "fp16=True; this stage quantises the embeddings to e4m3 on the GPU)")
            nb = emb.shape[0]
            if nb == 0:
                continue
            if nb > eng.rows_cap:
                raise L.RotorMapArgumentError(L.RM_ERR_ARG, f"batch height {nb} > engine rows_cap {eng.rows_cap}")
            if emb.shape[1] != rd:
                raise L.RotorMapArgumentError(L.RM_ERR_SHAPE, f"embeddings have {emb.shape[1]} columns; the index rdim is {rd}")
            q16 = torch.from_numpy(np.ascontiguousarray(emb)).to(dev)  # (nb, rdim) rows
            q32 = None
            if eng.mode == L.RM_MODE_FP16:
                q = q16
            else:
                q = torch.empty(nb, rd, dtype=torch.uint8, device=dev)
                L.check(L.lib().rm_quantize_e4m3(L.ctx(eng.device), L.ptr(q), L.ptr(q16), L.RM_F16,
                                                 nb * rd, L.stream_of(eng.device)))
                if eng.mode == L.RM_MODE_FP8_RESCORE:
                    q32 = q16.float()
            v, l = eng.batch_gemm_topk(q, q32)
            norms, heads, first = bat.norms, bat.heads, bat.first
        else:
            words, heads, first = bat
            nb = words.shape[0]
            if nb == 0:
                continue
            if re is None:
                raise L.RotorMapArgumentError(L.RM_ERR_ARG, "word batches need the encoder `re`")
            if nb > eng.rows_cap:
                raise L.RotorMapArgumentError(L.RM_ERR_ARG, f"batch height {nb} > engine rows_cap {eng.rows_cap}")
            if isinstance(words, np.ndarray):
                words = torch.from_numpy(np.ascontiguousarray(words).view(np.int32)).to(dev)
            nrm = torch.empty(nb, re.m, dtype=torch.float32, device=dev).t()
            v, l = eng.map_words(re, words.reshape(-1), nb, nrm)
            norms = nrm.cpu().numpy()
        yield TopKBatch(v.cpu().numpy(), l.cpu().numpy(), norms, list(heads), first)


def rope_topk_stream(re: RopeEncoder, file, B, k=20, kprime=32, mode="fp8_rescore",
                     batch_size=2 ** 13, rows_cap=None, normalize=0, fused=True, device=0,
                     chunk_bytes=1 << 30):
    """The full mapping flow on a FASTA `file` against the resident index B
    (ResidentIndex, TopKEngine or (N, rdim) e4m3 rows): yields TopKBatch per
    batch of <= batch_size reads (RotorMap/search/engine_fp8.jl:576-594).
    fused (default): words -> rm_map_batch on the device; fused=False: the
    reference's two stages, rope_encode_real_stream(fp16=True) -> topk_flow."""
    rows_cap = batch_size if rows_cap is None else int(rows_cap)
    if isinstance(B, TopKEngine):
        eng = B
        if eng.rows_cap < batch_size:
            raise L.RotorMapArgumentError(L.RM_ERR_ARG, f"batch_size {batch_size} > engine rows_cap {eng.rows_cap}")
    else:
        eng = TopKEngine(B, k=k, kprime=kprime, mode=mode, rows_cap=rows_cap, device=device)
    if eng.index.rdim != enc_rdim(re):
        raise L.RotorMapArgumentError(
            L.RM_ERR_SHAPE, f"B has {eng.index.rdim} rows; the encoder's real embedding dim is "
                            f"2*m*4^c = {enc_rdim(re)}")
    _check_encoder_args(re, re.k, batch_size, normalize)
    if fused:
        if normalize != 0:
            raise L.RotorMapArgumentError(L.RM_ERR_ARG, "the fused flow encodes with normalize = 0 "
                                                        "(the mapping index's mode); use fused=False")
        src = iter_fasta_gpu(file, re.k, batch_size, chunk_bytes, device)
        return topk_flow(src, eng, re)
    src = rope_encode_real_stream(re, file, re.k, batch_size, normalize, fp16=True, device=device,
                                  chunk_bytes=chunk_bytes)
    return topk_flow(src, eng)
