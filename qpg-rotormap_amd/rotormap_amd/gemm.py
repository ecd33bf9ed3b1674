"""GEMM layer: the :hip engine's chunk GEMM, the running top-k merge and the
e4m3 conversions (RotorMap/gemm/fp8_lt.jl, topk_kernels.jl, fp8_convert.jl)."""
from . import _lib as L
from .encode import _dtype_code


def _dev(t):
    return t.device.index or 0


def f32_to_f8(dst, src, stream=None):
    """dst (uint8, n) = e4m3 satfinite of src (fp32 or fp16, n) -- f32_to_f8!
    (RotorMap/gemm/fp8_convert.jl:100-103)."""
    if dst.numel() != src.numel():
        raise L.RotorMapArgumentError(L.RM_ERR_SHAPE, "dst and src sizes differ")
    d = _dev(src)
    L.check(L.lib().rm_quantize_e4m3(L.ctx(d), L.ptr(dst), L.ptr(src), _dtype_code(src),
                                     src.numel(), stream or L.stream_of(d)))
    return dst


def f8_to_f32(dst, src, stream=None):
    """dst (fp32, n) = exact value of e4m3 bytes src -- f8_to_f32!"""
    d = _dev(src)
    L.check(L.lib().rm_dequantize_e4m3(L.ctx(d), L.ptr(dst), L.ptr(src), src.numel(),
                                       stream or L.stream_of(d)))
    return dst


def fp8_gemm_hip(C, A, B, wi, K=None, b_col0=0, stream=None):
    """C (M x wi) = A (M x K) * B[:, b_col0 : b_col0+wi] -- the :hip twin of
    fp8_gemm_lt! (RotorMap/gemm/fp8_lt.jl:128-168).

    C: fp32/fp16 tensor of shape (M, >= wi) in any column-major layout (the
    reference's C buffer is (rows_cap, w) column-major: a transposed view).
    A: (M, K) e4m3 (uint8) or fp16, any strides.  B: (N, K) row-major
    tensor = the reference's (K, N) column-major database.
    """
    M = A.shape[0]
    K = K or A.shape[1]
    if C.stride(0) != 1:
        raise L.RotorMapArgumentError(L.RM_ERR_SHAPE, "C must be column-major (stride(0) == 1)")
    ab = _dtype_code(A)
    esz = 1 if ab == L.RM_E4M3 else 2
    d = _dev(A)
    bp = B.data_ptr() + b_col0 * B.stride(0) * esz
    L.check(L.lib().rm_gemm_chunk(L.ctx(d), L.ptr(C), _dtype_code(C), C.stride(1), L.ptr(A),
                                  A.stride(0), A.stride(1), L.ptr(bp), B.stride(0), ab, M, wi, K,
                                  stream or L.stream_of(d)))
    return C


def launch_seg_rowtopk_merge(D_val, D_loc, C, offset, k, stream=None):
    """Merge the per-row top-k of chunk C (M x wi, column-major) into the
    running D_val/D_loc ((rows_cap, k) column-major) -- launch_seg_rowtopk_merge!
    (RotorMap/gemm/topk_kernels.jl:331-347).  Fresh ids offset+1 .. offset+wi."""
    M, wi = C.shape
    d = _dev(C)
    L.check(L.lib().rm_rowtopk_merge(L.ctx(d), L.ptr(D_val), L.ptr(D_loc), D_val.stride(1),
                                     L.ptr(C), _dtype_code(C), C.stride(1), M, wi, offset, k,
                                     stream or L.stream_of(d)))
