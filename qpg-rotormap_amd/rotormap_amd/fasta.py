"""FASTA reads parsed and 2-bit packed on the GPU (rm_fasta_reads): the
host mirror of fasta_reads (RotorMap/fasta/reader.jl:321-350).

fasta_reads_gpu(data, k)      one buffer (bytes / uint8 array / path) ->
                              (device words (n, ceil(k/16)) int32, headers)
iter_fasta_gpu(path, k, ...)  the file in chunks cut at line-start '>' ->
                              batches of at most batch_size reads, in file
                              order, as (words, headers, first) with `first`
                              the 1-based index of the batch's first read
"""
import ctypes as C

import numpy as np

from . import _lib as L


def _as_bytes(data):
    """A writable contiguous uint8 view (a copy only when the input is read-only)."""
    if isinstance(data, (bytes, bytearray, memoryview)):
        a = np.frombuffer(data, np.uint8)
    elif isinstance(data, np.ndarray):
        a = np.ascontiguousarray(data, dtype=np.uint8).reshape(-1)
    else:
        return np.fromfile(data, np.uint8)
    return a if a.flags.writeable else a.copy()


def fasta_reads_gpu(data, k, device=0, stream=None):
    """Parse `data` on the GPU; returns (words, headers): words a device
    int32 tensor (n, ceil(k/16)) of packed reads, headers the reads' header
    lines ('>' included, line terminator dropped) as str."""
    import torch
    raw_h = _as_bytes(data)
    W = (k + 15) // 16
    cap = int(np.count_nonzero(raw_h == ord(">")))  # >= the records of the buffer
    dev = torch.device(f"cuda:{device}")
    if raw_h.size == 0 or cap == 0:
        return torch.empty(0, W, dtype=torch.int32, device=dev), []
    raw = torch.from_numpy(raw_h).to(dev)
    words = torch.empty(cap, W, dtype=torch.int32, device=dev)
    hoff = torch.empty(cap, dtype=torch.int64, device=dev)
    hlen = torch.empty(cap, dtype=torch.int32, device=dev)
    n = C.c_int64()
    L.check(L.lib().rm_fasta_reads(L.ctx(device), L.ptr(raw), raw_h.size, int(k), L.ptr(words), cap,
                                   L.ptr(hoff), L.ptr(hlen), C.byref(n),
                                   stream or L.stream_of(device)))
    n = n.value
    ho, hl = hoff[:n].cpu().numpy(), hlen[:n].cpu().numpy()
    heads = [raw_h[o:o + l].tobytes().decode("latin-1") for o, l in zip(ho, hl)]
    return words[:n], heads


def _last_record_start(buf):
    """Offset of the last line-start '>' in buf (0 if none after byte 0)."""
    pos = np.flatnonzero(buf == ord(">"))
    for p in pos[::-1]:
        if p > 0 and buf[p - 1] in (10, 13):
            return int(p)
    return 0


def iter_fasta_gpu(path, k, batch_size=1 << 16, chunk_bytes=1 << 30, device=0):
    """Stream `path` through the GPU parser in chunks cut at a line-start
    '>' (the split the parser allows), yielding (words, headers, first)
    batches of at most batch_size reads."""
    import torch
    first = 1
    pend_w, pend_h = [], []
    carry = np.zeros(0, np.uint8)
    with open(path, "rb") as f:
        while True:
            chunk = np.empty(chunk_bytes, np.uint8)
            chunk = chunk[:f.readinto(chunk)]
            eof = chunk.size == 0
            buf = np.concatenate([carry, chunk]) if carry.size else chunk
            if eof:
                cut = buf.size
            else:
                cut = _last_record_start(buf)
                if cut == 0:  # no record boundary yet: read more
                    carry = buf
                    continue
            if cut:
                w, h = fasta_reads_gpu(buf[:cut], k, device)
                pend_w.append(w)
                pend_h.extend(h)
            carry = buf[cut:].copy()
            nw = sum(x.shape[0] for x in pend_w)
            while nw >= batch_size or (eof and nw):
                allw = torch.cat(pend_w) if len(pend_w) > 1 else pend_w[0]
                take = min(batch_size, nw)
                yield allw[:take], pend_h[:take], first
                first += take
                pend_w, pend_h = [allw[take:]], pend_h[take:]
                nw -= take
            if eof:
                return
