# ==============================================================================
# common/hip.jl -- the process-wide handle of librotormap_hip.so for the `:hip`
# engine: library path, status -> exception mapping, one context per device.
#
# Include first (before encode/kernel_hip.jl, gemm/fp8_hip.jl and
# search/engine_hip.jl), the way common/ precedes the other layers
# (RotorMap/ARCHITECTURE.md:76-113).  It plays the role of the cuBLASLt
# handle + workspace globals of gemm/fp8_lt.jl (cublaslt_init, :47-93): a
# context owns its workspace, phasor tables and tile plans, and is created
# lazily, once per device.
#
# Device arrays: any GPU array whose `pointer` is a device address (AMDGPU.jl
# ROCArray on an MI355X box) is passed as a raw pointer; the library never
# frees caller memory.  Host arrays use the entry points that stage through
# the context (see RotorMapHIP.jl for the host-pointer-only module).
# ==============================================================================

const HIP_LIB = get(ENV, "ROTORMAP_HIP_LIB",
                    joinpath(@__DIR__, "..", "..", "..", "rotormap_amd", "librotormap_hip.so"))

const HIP_OK, HIP_ERR_ARG, HIP_ERR_SHAPE = 0, 1, 2
const HIP_F32, HIP_F16, HIP_E4M3 = Cint(0), Cint(1), Cint(2)
const HIP_MODES = Dict(:fp8 => Cint(0), :fp8_rescore => Cint(1), :fp16 => Cint(2))
const HIP_INDEX_HOST, HIP_INDEX_FP16 = Cint(1), Cint(2)

hip_last_error() = unsafe_string(ccall((:rm_last_error, HIP_LIB), Cstring, ()))

# a failed call throws like the reference's layers: ArgumentError for bad
# arguments and shapes (RotorMap/search/engine_fp8.jl:313-326), else an
# ErrorException carrying the library's message (gemm/fp8_lt.jl:55)
function hip_check(rc::Integer)
    rc == HIP_OK && return nothing
    msg = hip_last_error()
    rc in (HIP_ERR_ARG, HIP_ERR_SHAPE) && throw(ArgumentError(msg))
    error("librotormap_hip: " * msg)
end

const _HIP_CTX = Dict{Int,Ptr{Cvoid}}()
const _HIP_CTX_LOCK = ReentrantLock()

"""hip_ctx(device = 0): the device's context (created on first use, kept for the process)."""
function hip_ctx(device::Integer = 0)
    lock(_HIP_CTX_LOCK) do
        get!(_HIP_CTX, Int(device)) do
            out = Ref{Ptr{Cvoid}}(C_NULL)
            hip_check(ccall((:rm_ctx_create, HIP_LIB), Cint, (Cint, Ref{Ptr{Cvoid}}), device, out))
            out[]
        end
    end
end

"""Raw device address of a GPU array (or C_NULL for `nothing`)."""
hip_devptr(x) = x === nothing ? C_NULL : Ptr{Cvoid}(UInt(pointer(x)))
hip_devptr(p::Ptr) = Ptr{Cvoid}(p)

"""Element-type code of an encoder / GEMM output array."""
function hip_dtype(::Type{T}) where {T}
    T === Float32 && return HIP_F32
    T === Float16 && return HIP_F16
    sizeof(T) == 1 && return HIP_E4M3  # Float8_E4M3FN (or its UInt8 bits)
    throw(ArgumentError("unsupported element type $T (Float32, Float16 or e4m3)"))
end

hip_version() = unsafe_string(ccall((:rm_version, HIP_LIB), Cstring, ()))
