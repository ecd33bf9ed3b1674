# RotorMapHIP.jl -- Julia binding of librotormap_hip.so (include/rotormap_hip.h)
# for RotorMap's mapping path on an MI355X: the `:hip` engine.
#
# Host pointers only: the index is uploaded from the host database matrix
# (index_database_host, RotorMap/index/load.jl:49-73), reads go in as the
# parser's packed words (FastaFragment.words, RotorMap/fasta/reader.jl:46-50)
# and (vals, locs) come back as host matrices, so a Julia process on an AMD
# box needs neither CUDA.jl nor AMDGPU.jl.  Results use the reference's
# conventions: vals (nb, k) Float32 descending, locs (nb, k) Int32 1-based
# global column ids, 0 = empty (RotorMap/search/engine_fp8.jl:255-271).
#
# Usage from RotorMap (after `include("RotorMapHIP.jl"); using .RotorMapHIP`):
#   for bat in rope_topk_stream_hip(re, reads_path, Bh; k = 20)   # tools/map.jl:128-142
#       ... bat.vals, bat.locs, bat.heads ...
#   end
module RotorMapHIP

export HipContext, HipIndex, HipServer, rope_topk_stream_hip, map_words_hip!,
       hip_encode_batch!, hip_version, fasta_reads_hip

const LIB = get(ENV, "ROTORMAP_HIP_LIB",
                joinpath(@__DIR__, "..", "rotormap_amd", "librotormap_hip.so"))

const RM_OK, RM_ERR_ARG, RM_ERR_SHAPE, RM_ERR_NOMEM, RM_ERR_HIP, RM_ERR_STATE = 0, 1, 2, 3, 4, 5
const RM_F32, RM_F16, RM_E4M3 = 0, 1, 2
const MODES = Dict(:fp8 => 0, :fp8_rescore => 1, :fp16 => 2)
const RM_INDEX_HOST, RM_INDEX_FP16 = 1, 2

hip_version() = unsafe_string(ccall((:rm_version, LIB), Cstring, ()))
last_error() = unsafe_string(ccall((:rm_last_error, LIB), Cstring, ()))

# status -> the reference's exception types (ArgumentError for arguments and
# shapes, ErrorException otherwise), same wording as the library message
function check(rc::Integer)
    rc == RM_OK && return nothing
    msg = last_error()
    rc in (RM_ERR_ARG, RM_ERR_SHAPE) && throw(ArgumentError(msg))
    error("librotormap_hip: " * msg)
end

mutable struct HipContext
    ptr::Ptr{Cvoid}
    device::Int
end

function HipContext(device::Integer = 0)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:rm_ctx_create, LIB), Cint, (Cint, Ref{Ptr{Cvoid}}), device, out))
    ctx = HipContext(out[], device)
    finalizer(c -> (c.ptr != C_NULL && ccall((:rm_ctx_destroy, LIB), Cint, (Ptr{Cvoid},), c.ptr);
                    c.ptr = C_NULL), ctx)
    return ctx
end

"""
The e4m3 database resident in HBM, uploaded from a host matrix `Bh` of
size (rdim, N) with 1-byte elements (Float8_E4M3FN or UInt8 bits): column j
is window j, exactly the `.bin`'s `embeds8` (RotorMap/index/build.jl:730).
`fp16 = true` also keeps the exact fp16 widening for the FP16 path
(index_database_f16, RotorMap/index/load.jl:130-143).
"""
mutable struct HipIndex
    ptr::Ptr{Cvoid}
    ctx::HipContext
    rdim::Int
    N::Int
end

function HipIndex(ctx::HipContext, Bh::AbstractMatrix; fp16::Bool = false)
    sizeof(eltype(Bh)) == 1 || throw(ArgumentError("database elements must be 1-byte e4m3"))
    B = Bh isa Matrix ? Bh : Matrix(Bh)
    rdim, N = size(B)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    flags = RM_INDEX_HOST | (fp16 ? RM_INDEX_FP16 : 0)
    GC.@preserve B check(ccall((:rm_index_create, LIB), Cint,
                               (Ptr{Cvoid}, Ptr{UInt8}, Int64, Int64, Cint, Ref{Ptr{Cvoid}}),
                               ctx.ptr, Ptr{UInt8}(pointer(B)), rdim, N, flags, out))
    idx = HipIndex(out[], ctx, rdim, N)
    finalizer(i -> (i.ptr != C_NULL && ccall((:rm_index_destroy, LIB), Cint, (Ptr{Cvoid},), i.ptr);
                    i.ptr = C_NULL), idx)
    return idx
end

"""
Continuous-batching server over a resident index: requests (host words) are
packed row by row into launches of at most `max_rows` fragments, `n_streams`
launches in flight.  The whole encode -> fused GEMM + top-k' -> FP32 rescore
-> top-k flow runs inside (`mode` = :fp8_rescore, :fp8 or :fp16).
"""
mutable struct HipServer
    ptr::Ptr{Cvoid}
    index::HipIndex
    k::Int
    W::Int
end

const POLICIES = Dict(:fifo => Cint(0), :srpt => Cint(1))  # RM_SCHED_FIFO, RM_SCHED_SRPT

function HipServer(index::HipIndex, kmer::Integer, s::Integer, m::Integer, c::Integer;
                   k::Integer = 20, kprime::Integer = 32, mode::Symbol = :fp8_rescore,
                   max_rows::Integer = 2^14, n_streams::Integer = 2, hbm_budget::Integer = 0,
                   policy::Symbol = :fifo, lowload::Tuple{Integer,Integer} = (0, 0))
    haskey(MODES, mode) || throw(ArgumentError("mode must be :fp8, :fp8_rescore or :fp16"))
    haskey(POLICIES, policy) || throw(ArgumentError("policy must be :fifo or :srpt"))
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:rm_server_create, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint, Cint, Cint, Cint, Cint, Int64, Cint,
                 Int64, Ref{Ptr{Cvoid}}),
                index.ctx.ptr, index.ptr, kmer, s, m, c, k, kprime, MODES[mode], max_rows,
                n_streams, hbm_budget, out))
    srv = HipServer(out[], index, Int(k), cld(Int(kmer), 16))
    finalizer(v -> (v.ptr != C_NULL && ccall((:rm_server_destroy, LIB), Cint, (Ptr{Cvoid},), v.ptr);
                    v.ptr = C_NULL), srv)
    # :srpt fills each launch with the requests that have the fewest rows left
    check(ccall((:rm_server_set_policy, LIB), Cint, (Ptr{Cvoid}, Cint), srv.ptr, POLICIES[policy]))
    # lowload = (launch_rows, backlog_rows): while the unlaunched backlog is
    # small, one launch in flight holding whole requests (see the header)
    check(ccall((:rm_server_set_lowload, LIB), Cint, (Ptr{Cvoid}, Int64, Int64), srv.ptr,
                lowload[1], lowload[2]))
    return srv
end

# enqueue nb fragments (words: W x nb, fragment-contiguous) without a copy;
# the caller keeps `words` alive until `wait!` returns for the id
function submit!(srv::HipServer, words::Matrix{UInt32}, nb::Integer)
    size(words, 1) == srv.W ||
        throw(ArgumentError("words must be $(srv.W) x n (cld(k, 16) words per fragment)"))
    id = Ref{Int64}(0)
    check(ccall((:rm_server_submit_ref, LIB), Cint, (Ptr{Cvoid}, Ptr{UInt32}, Int64, Ref{Int64}),
                srv.ptr, pointer(words), nb, id))
    return id[]
end

# wait for request `id` of nb rows: (vals (nb, k) Float32, locs (nb, k) Int32)
function wait!(srv::HipServer, id::Integer, nb::Integer)
    vt = Matrix{Float32}(undef, srv.k, nb)  # the library writes (nb x k) row-major
    lt = Matrix{Int32}(undef, srv.k, nb)
    check(ccall((:rm_server_wait, LIB), Cint,
                (Ptr{Cvoid}, Int64, Cint, Ptr{Float32}, Ptr{Int32}, Ptr{Cdouble}, Ptr{Cdouble},
                 Ptr{Cdouble}),
                srv.ptr, id, -1, vt, lt, C_NULL, C_NULL, C_NULL))
    return permutedims(vt), permutedims(lt)
end

"""
map_words_hip!(srv, words, nb) -> (vals, locs): one synchronous batch of nb
packed fragments (`words` W x nb UInt32, the layout rope_encode_real_stream
gathers, RotorMap/encode/stream.jl:161-170).
"""
function map_words_hip!(srv::HipServer, words::Matrix{UInt32}, nb::Integer)
    GC.@preserve words begin
        id = submit!(srv, words, nb)
        return wait!(srv, id, nb)
    end
end

"""
rope_topk_stream_hip(re, file, Bh; k, kprime, mode, batch_size, n_streams)

The `:hip` twin of rope_topk_stream / rope_topk_stream_compact
(RotorMap/search/engine_fp8.jl:576-594, engine_compact.jl:415): parses
`file` with the reference's own `fasta_reads` (passed as `reader`, default
`Main.fasta_reads`), maps each batch on the GPU and yields named tuples with
the TopKBatch fields (vals, locs, norms, heads, first).  `norms` is an empty
(m, 0)-shaped placeholder: the fused path does not return encoder norms
(tools/map.jl does not read them).  Two batches are in flight: batch i + 1
is submitted before batch i is collected.
"""
function rope_topk_stream_hip(re, file::String, Bh::AbstractMatrix;
                              k::Integer = 20, kprime::Integer = 32,
                              mode::Symbol = :fp8_rescore, batch_size::Integer = 2^13,
                              n_streams::Integer = 2, device::Integer = 0,
                              reader = getfield(Main, :fasta_reads),
                              parts::Integer = Threads.nthreads(), out_cap::Integer = 2)
    ctx = HipContext(device)
    idx = HipIndex(ctx, Bh; fp16 = (mode === :fp16))
    srv = HipServer(idx, re.k, re.s, re.m, re.c; k, kprime, mode, max_rows = batch_size,
                    n_streams)
    W = cld(re.k, 16)
    out = Channel{Any}(out_cap)
    task = Threads.@spawn try
        frags = reader(file; k = re.k, parts)
        pending = nothing  # (id, words, nb, heads, first)
        first = 1
        done = false
        while !done
            words = Matrix{UInt32}(undef, W, batch_size)
            heads = String[]
            nb = 0
            for f in frags
                nb += 1
                copyto!(view(words, :, nb), f.words)
                push!(heads, f.header)
                nb == batch_size && break
            end
            done = nb < batch_size
            cur = nb > 0 ? (submit!(srv, words, nb), words, nb, heads, first) : nothing
            first += nb
            if pending !== nothing
                id, wds, pnb, ph, pf = pending
                vals, locs = GC.@preserve wds wait!(srv, id, pnb)
                put!(out, (vals = vals, locs = locs, norms = zeros(Float32, re.m, 0),
                           heads = ph, first = pf))
            end
            pending = cur
        end
        if pending !== nothing
            id, wds, pnb, ph, pf = pending
            vals, locs = GC.@preserve wds wait!(srv, id, pnb)
            put!(out, (vals = vals, locs = locs, norms = zeros(Float32, re.m, 0),
                       heads = ph, first = pf))
        end
    catch e
        close(out, e)
        rethrow()
    finally
        isopen(out) && close(out)
        finalize(srv)
    end
    bind(out, task)
    return out
end

"""
fasta_reads_hip(file; k, chunk_bytes, device) -> Channel

The GPU parse of `file` (rm_fasta_reads_host): the same fragments as the
reference's fasta_reads (RotorMap/fasta/reader.jl:321-350) -- named tuples
(header, len = k, words = cld(k, 16) UInt32), kept records in file order.
The file is read in chunks cut at a line-start '>'.  Usable as the `reader`
of rope_topk_stream_hip.
"""
function fasta_reads_hip(file::AbstractString; k::Integer = 20_000, parts::Integer = 1,
                         chunk_bytes::Integer = 1 << 30, device::Integer = 0)
    ctx = HipContext(device)
    W = cld(Int(k), 16)
    out = Channel{Any}(64)
    task = Threads.@spawn try
        open(file, "r") do io
            carry = UInt8[]
            while true
                chunk = read(io, chunk_bytes)
                eof_ = isempty(chunk)
                buf = isempty(carry) ? chunk : vcat(carry, chunk)
                cut = length(buf)
                if !eof_
                    cut = 0
                    for p in length(buf):-1:2  # last line-start '>'
                        if buf[p] == UInt8('>') && (buf[p - 1] == UInt8('\n') || buf[p - 1] == UInt8('\r'))
                            cut = p - 1
                            break
                        end
                    end
                    if cut == 0
                        carry = buf
                        continue
                    end
                end
                if cut > 0
                    cap = count(==(UInt8('>')), view(buf, 1:cut)) + 1
                    words = Matrix{UInt32}(undef, W, cap)
                    off = Vector{Int64}(undef, cap)
                    len = Vector{Int32}(undef, cap)
                    n = Ref{Int64}(0)
                    GC.@preserve buf check(ccall((:rm_fasta_reads_host, LIB), Cint,
                        (Ptr{Cvoid}, Ptr{UInt8}, Int64, Cint, Ptr{UInt32}, Int64, Ptr{Int64},
                         Ptr{Int32}, Ref{Int64}),
                        ctx.ptr, pointer(buf), cut, k, words, cap, off, len, n))
                    for i in 1:n[]
                        hdr = String(buf[off[i] + 1:off[i] + len[i]])
                        put!(out, (header = hdr, len = Int(k), words = words[:, i]))
                    end
                end
                carry = buf[cut + 1:end]
                eof_ && break
            end
        end
    catch e
        close(out, e)
        rethrow()
    finally
        isopen(out) && close(out)
    end
    bind(out, task)
    return out
end

"""
hip_encode_batch!(ctx, dest_ptr, dest_dtype, norms_ptr, re, words_ptr, n; normalize, stream)

The encoder drop-in for encode_frag_real_batch! (RotorMap/encode/kernel.jl:
298-325) on device pointers held by another HIP-aware package: words (n x W
UInt32, fragment-contiguous) -> dest (n, 2 m 4^c) COLUMN-major (element
(r, j) at r + j n, `dest_dtype` RM_F32 / RM_F16 / RM_E4M3), norms (m, n).
"""
function hip_encode_batch!(ctx::HipContext, dest::Ptr{Cvoid}, dest_dtype::Integer,
                           norms::Ptr{Float32}, re, words::Ptr{UInt32}, n::Integer;
                           normalize::Integer = 0, stream::Ptr{Cvoid} = C_NULL)
    check(ccall((:rm_encode_batch, LIB), Cint,
                (Ptr{Cvoid}, Ptr{UInt32}, Int64, Cint, Cint, Cint, Cint, Cint,
                 Ptr{Cvoid}, Cint, Int64, Int64, Ptr{Cvoid}, Cint, Int64, Int64,
                 Ptr{Float32}, Ptr{Int32}, Ptr{Cvoid}),
                ctx.ptr, words, n, re.k, re.s, re.m, re.c, normalize,
                dest, dest_dtype, 1, n, C_NULL, 0, 0, 0, norms, C_NULL, stream))
end

end # module
