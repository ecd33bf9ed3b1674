# ==============================================================================
# encode/kernel_hip.jl -- the `:hip` encoder: encode_frag_real_batch! on an
# MI355X through librotormap_hip.so (rm_encode_batch).
#
# Drop-in for encode_frag_real_batch! (RotorMap/encode/kernel.jl:298-325):
# same arguments, same argument checks, same output layout -- dest (n, 2D)
# column-major (the real [Re; Im] split of kernel_rope_frag_real, :284-287),
# dest_norms (m, n) Float32 (the reference's norms layout, :189-213).  The
# kernel accumulates exact integer phasors (deterministic; the reference's
# float LDS atomics are not) and normalises in Float32 like the reference's.
# Requires common/hip.jl.
# ==============================================================================

"""
    encode_frag_real_batch_hip!(dest, dest_norms, re, dnas; normalize = 0, stream = C_NULL)

Encode the n fragments of `dnas` (a device vector of n * cld(k, 16) UInt32
words, fragment-contiguous) into `dest` (n, 2 m 4^c; Float32, Float16 or
e4m3 elements) and `dest_norms` (m, n) Float32 (or `nothing`).
"""
function encode_frag_real_batch_hip!(dest, dest_norms, re, dnas; normalize::Integer = 0,
                                     stream::Ptr{Cvoid} = C_NULL, device::Integer = 0)
    (; k, s, m, c) = re
    s <= 16 || throw(ArgumentError("the encoder needs s <= 16 (got s = $s)"))
    c <= s || throw(ArgumentError("the compact c-mer must not exceed the s-mer (c = $c, s = $s)"))
    n = size(dest, 1)
    D = m * 4^c
    size(dest, 2) == 2 * D ||
        throw(ArgumentError("dest must be (n, 2*m*4^c) = ($n, $(2D)), got $(size(dest))"))
    W = cld(k, 16)
    length(dnas) == n * W ||
        throw(ArgumentError("dnas must hold n * cld(k, 16) = $(n * W) words, got $(length(dnas))"))
    dest_norms === nothing || size(dest_norms) == (m, n) ||
        throw(ArgumentError("dest_norms must be (m, n) = ($m, $n)"))
    normalize in 0:3 || throw(ArgumentError("normalize must be 0..3"))
    n == 0 && return dest
    hip_check(ccall((:rm_encode_batch, HIP_LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Cint, Cint, Cint, Cint, Cint,
                     Ptr{Cvoid}, Cint, Int64, Int64, Ptr{Cvoid}, Cint, Int64, Int64,
                     Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                    hip_ctx(device), hip_devptr(dnas), n, k, s, m, c, normalize,
                    hip_devptr(dest), hip_dtype(eltype(dest)), 1, n,  # column-major (n, 2D)
                    C_NULL, HIP_F32, 0, 0,
                    hip_devptr(dest_norms), C_NULL, stream))
    return dest
end

"""The reverse complement of n fragments (revcomp_frag_words!, RotorMap/index/build.jl:448-476)."""
function revcomp_frag_words_hip!(drc, dnas, k::Integer, n::Integer; stream::Ptr{Cvoid} = C_NULL,
                                 device::Integer = 0)
    length(dnas) == length(drc) == n * cld(k, 16) ||
        throw(ArgumentError("words must hold n * cld(k, 16) words"))
    hip_check(ccall((:rm_revcomp_words, HIP_LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Cint, Ptr{Cvoid}),
                    hip_ctx(device), hip_devptr(drc), hip_devptr(dnas), n, k, stream))
    return drc
end
