"""Encoder layer: the RopeEncoder record and the HIP batch encoder.

Mirrors RotorMap/encode/ropeencoder.jl (RopeEncoder, dim) and
RotorMap/encode/kernel.jl:298-335 (encode_frag_real_batch!,
encode_frag_real_batch); the kernel is encode.hip's rope_encode_kernel.
"""
from . import _lib as L


class RopeEncoder:
    """Encoder parameters k, s, m, c (RotorMap/encode/ropeencoder.jl:16-28)."""

    def __init__(self, k, s=5, m=1, c=0):
        self.k, self.s, self.m, self.c = int(k), int(s), int(m), int(c)

    def __repr__(self):
        return f"RopeEncoder(k={self.k}, s={self.s}, m={self.m}, c={self.c})"

    def __eq__(self, o):
        return isinstance(o, RopeEncoder) and (self.k, self.s, self.m, self.c) == (o.k, o.s, o.m, o.c)


def dim(re):
    """Complex dimension D (RotorMap/encode/ropeencoder.jl:30)."""
    return re.m * 4 ** re.c if re.c != 0 else re.m * 4 ** re.s


def rdim(re):
    """Real [Re; Im] dimension 2D."""
    return 2 * dim(re)


def _dtype_code(t):
    import torch
    if t.dtype == torch.float32:
        return L.RM_F32
    if t.dtype == torch.float16:
        return L.RM_F16
    if t.dtype in (torch.uint8, getattr(torch, "float8_e4m3fn", None)):
        return L.RM_E4M3
    raise L.RotorMapArgumentError(L.RM_ERR_ARG, f"unsupported output dtype {t.dtype}")


def encode_frag_real_batch(dest, dest_norms, re, dnas, normalize=0, counts=None, dest2=None,
                           stream=None):
    """Encode n fragments (device uint32 words, n*ceil(k/16)) into dest (n, 2D).

    dest may have any strides (the reference's column-major (n, 2D) is a
    transposed view); dtype float32, float16 or e4m3 (uint8 /
    float8_e4m3fn).  dest_norms (m, n) float32 (the reference's layout, i.e. a
    tensor whose element (i, f) sits at f*m + i), or None.  counts (n, 4^c)
    int32 receives the integer s-mer key histogram.  Asserts mirror
    RotorMap/encode/kernel.jl:305-315.
    """
    if re.s > 16:
        raise L.RotorMapArgumentError(L.RM_ERR_ARG, "the encoder requires s <= 16")
    if re.c > re.s:
        raise L.RotorMapArgumentError(L.RM_ERR_ARG, "the compact c-mer must not exceed the s-mer")
    n = dest.shape[0]
    if dest.shape[1] != rdim(re):
        raise L.RotorMapArgumentError(L.RM_ERR_SHAPE, "dest must be (n, 2*m*4^c)")
    W = (re.k + 15) // 16
    if dnas.numel() != n * W:
        raise L.RotorMapArgumentError(L.RM_ERR_SHAPE, "dnas must hold n * cld(k, 16) words")
    if dest_norms is not None:
        if tuple(dest_norms.shape) != (re.m, n) or dest_norms.stride() != (1, re.m):
            raise L.RotorMapArgumentError(L.RM_ERR_SHAPE, "dest_norms must be (m, n) column-major")
    dev = dest.device.index or 0
    st = stream if stream is not None else L.stream_of(dev)
    d2 = (L.ptr(dest2), _dtype_code(dest2), dest2.stride(0), dest2.stride(1)) if dest2 is not None \
        else (None, 0, 0, 0)
    L.check(L.lib().rm_encode_batch(
        L.ctx(dev), L.ptr(dnas), n, re.k, re.s, re.m, re.c, normalize,
        L.ptr(dest), _dtype_code(dest), dest.stride(0), dest.stride(1), *d2,
        L.ptr(dest_norms), L.ptr(counts), st))
    return dest, dest_norms


def encode_frag_real_batch_alloc(dtype, re, dnas, normalize=0):
    """Allocating variant: (embeds (n, 2D) column-major, norms (m, n))."""
    import torch
    n = dnas.numel() // ((re.k + 15) // 16)
    dest = torch.zeros(rdim(re), n, dtype=dtype, device=dnas.device).t()
    norms = torch.zeros(n, re.m, dtype=torch.float32, device=dnas.device).t()
    return encode_frag_real_batch(dest, norms, re, dnas, normalize)


def revcomp_frag_words(drc, dnas, k, n, stream=None):
    """Reverse complement n k-base fragments (RotorMap/index/build.jl:448-476)."""
    W = (k + 15) // 16
    if dnas.numel() != n * W or drc.numel() != n * W:
        raise L.RotorMapArgumentError(L.RM_ERR_SHAPE, "words must hold n * cld(k, 16) words")
    dev = dnas.device.index or 0
    L.check(L.lib().rm_revcomp_words(L.ctx(dev), L.ptr(drc), L.ptr(dnas), n, k,
                                     stream if stream is not None else L.stream_of(dev)))
    return drc


def phase_table(k, s, m):
    """The encoder's integer phasor table (host numpy int32, (m, k-s+1, 2))."""
    import numpy as np
    nw = k - s + 1
    out = np.empty((m, nw, 2), dtype=np.int32)
    L.check(L.lib().rm_phase_table(k, s, m, L.ptr(out)))
    return out
