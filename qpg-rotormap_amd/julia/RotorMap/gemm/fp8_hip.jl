# ==============================================================================
# gemm/fp8_hip.jl -- the `:hip` GEMM engine beside `:lt` (gemm/fp8_lt.jl) and
# `:mma` (gemm/fp8_ptx.jl): hand-written gfx950 MFMA kernels in
# librotormap_hip.so, reached through ccall.  Requires common/hip.jl.
#
#   fp8_gemm_hip!(C, A, Bp, wi)     the chunk GEMM of _gemm_chunk! with the
#                                   contract of fp8_gemm_lt! (:128-168):
#                                   D = A^T-layout x B chunk, e4m3 x e4m3 with
#                                   fp32 accumulation, C (rows_cap, wi) Float16
#                                   or Float32 column-major (ld = rows_cap), A
#                                   the (rows_cap, K) column-major staging a8,
#                                   Bp the chunk's first column of the (K, N)
#                                   database (column stride K)
#   launch_seg_rowtopk_merge_hip!   the running row top-k merge of a C chunk
#                                   (topk_kernels.jl:331-347)
#   f32_to_f8_hip!, f8_to_f32_hip!  fp8_convert.jl:100-108 (satfinite RNE)
#   hip_padmul                      the engine's batch-height multiple
#
# The engine plumbing in search/engine_fp8.jl that selects it is three lines
# (shown in search/engine_hip.jl's header): `engine in (:mma, :lt, :hip)`
# (:315), `padmul(:hip)` (:275) and a `:hip` branch of `_gemm_chunk!`
# (:379-387) calling fp8_gemm_hip!.  With the fused path (search/engine_hip.jl)
# the C chunk never exists: scores stay in the kernel.
# ==============================================================================

# Any batch height works (ragged tiles read zeros past the last row); 16 keeps
# the reference's 16-element leading-dimension multiple of the cuBLASLt path.
hip_padmul() = 16

"""
    fp8_gemm_hip!(C, A, Bp, wi; K = size(A, 2), stream = C_NULL)

C[:, 1:wi] = A * B[:, lo:lo+wi-1] with B's chunk given by its device base
pointer `Bp` (column stride K bytes).  A: (M, K) column-major e4m3 (element
(r, k) at r + k M); C: (M, wi) Float16 or Float32, leading dimension size(C, 1).
"""
function fp8_gemm_hip!(C, A, Bp, wi::Integer; K::Integer = size(A, 2),
                       stream::Ptr{Cvoid} = C_NULL, device::Integer = 0)
    M = size(A, 1)
    ldc = size(C, 1)
    ldc >= M || throw(ArgumentError("C's leading dimension $ldc < M = $M"))
    size(C, 2) >= wi || throw(ArgumentError("C has $(size(C, 2)) columns < wi = $wi"))
    K % 64 == 0 || throw(ArgumentError("the :hip GEMM needs K % 64 == 0 (K = $K)"))
    hip_check(ccall((:rm_gemm_chunk, HIP_LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Int64, Ptr{Cvoid}, Int64, Int64, Ptr{Cvoid},
                     Int64, Cint, Int64, Int64, Int64, Ptr{Cvoid}),
                    hip_ctx(device), hip_devptr(C), hip_dtype(eltype(C)), ldc,
                    hip_devptr(A), 1, M,      # column-major A: row stride 1, k stride M
                    hip_devptr(Bp), K, HIP_E4M3, M, wi, K, stream))
    return C
end

"""
    launch_seg_rowtopk_merge_hip!(D_val, D_loc, C, offset, wi; k = size(D_val, 2))

Merge the first wi columns of the chunk C (rows_cap, >= wi) into the running
per-row top-k (D_val Float32, D_loc Int32 1-based global column ids, both
(rows_cap, k) column-major); fresh columns are offset+1 .. offset+wi.
"""
function launch_seg_rowtopk_merge_hip!(D_val, D_loc, C, offset::Integer, wi::Integer;
                                       k::Integer = size(D_val, 2), stream::Ptr{Cvoid} = C_NULL,
                                       device::Integer = 0)
    M = size(C, 1)
    size(D_val) == size(D_loc) || throw(ArgumentError("D_val / D_loc shapes differ"))
    size(D_val, 1) >= M || throw(ArgumentError("D_val has fewer rows than C"))
    hip_check(ccall((:rm_rowtopk_merge, HIP_LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Cint, Int64, Int64,
                     Int64, Int64, Cint, Ptr{Cvoid}),
                    hip_ctx(device), hip_devptr(D_val), hip_devptr(D_loc), size(D_val, 1),
                    hip_devptr(C), hip_dtype(eltype(C)), M, M, wi, offset, k, stream))
    return D_val, D_loc
end

"""f32_to_f8_hip!(dst, src): e4m3 satfinite RNE of Float32 / Float16 `src` (device)."""
function f32_to_f8_hip!(dst, src; stream::Ptr{Cvoid} = C_NULL, device::Integer = 0)
    length(dst) == length(src) || throw(ArgumentError("length mismatch"))
    hip_check(ccall((:rm_quantize_e4m3, HIP_LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Int64, Ptr{Cvoid}),
                    hip_ctx(device), hip_devptr(dst), hip_devptr(src), hip_dtype(eltype(src)),
                    length(src), stream))
    return dst
end

"""f8_to_f32_hip!(dst, src): exact e4m3 -> Float32 (device)."""
function f8_to_f32_hip!(dst, src; stream::Ptr{Cvoid} = C_NULL, device::Integer = 0)
    length(dst) == length(src) || throw(ArgumentError("length mismatch"))
    hip_check(ccall((:rm_dequantize_e4m3, HIP_LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cvoid}),
                    hip_ctx(device), hip_devptr(dst), hip_devptr(src), length(src), stream))
    return dst
end
