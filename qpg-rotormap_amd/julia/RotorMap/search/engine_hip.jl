# ==============================================================================
# search/engine_hip.jl -- the `:hip` streaming engine: the fused GEMM + top-k
# (+ FP32 rescoring) over a resident index on an MI355X, with the reference's
# TopKEngine / batch_gemm_topk! / topk_flow / rope_topk_stream surface
# (RotorMap/search/engine_fp8.jl:290-594).  Requires common/hip.jl (and, for
# the chunk-wise `:hip` engine, gemm/fp8_hip.jl).
#
# Two ways in:
#
# 1. The chunk engine inside the reference's TopKEngine.  engine_fp8.jl
#    changes in three places:
#      :275      padmul(engine) returns hip_padmul() for :hip
#      :315      the engine check accepts (:mma, :lt, :hip)
#      :379-387  _gemm_chunk! gets a third branch:
#                  elseif eng.engine === :hip
#                      fp8_gemm_hip!(buf, eng.a8, Bp, wi; K = size(eng.B, 1))
#    and batch_gemm_topk! then runs unchanged (C chunks + the row top-k).
#
# 2. The fused engine below: one launch per batch over all N columns, scores
#    never leave the kernel, optional FP32 rescoring of the top-k'.  Its
#    batches are the same TopKBatch named fields, so map_and_count
#    (engine_fp8.jl:654-733) and tools/map.jl's match table consume them
#    unchanged.
# ==============================================================================

"""
    HipTopKEngine(B; k = 20, kprime = 32, mode = :fp8_rescore, rows_cap = 2^13,
                  fp16 = mode === :fp16, host = false, rescore = :e4m3, device = 0)

The resident database for the fused path: `B` (rdim, N) e4m3 columns (a
device array, or a host matrix with host = true), as index_database returns
it (RotorMap/index/load.jl:156-161); fp16 = true also keeps the exact fp16
widening (index_database_f16, :130-143).  A Float16 `B` is the classic fp16
index (index_to_f16, :96-116), kept verbatim for :fp16 and quantised to
e4m3(fp32(fp16)) on the device for the FP8 modes (index_to_f8, :76-89).
mode :fp8 (top-k of the e4m3 GEMM), :fp8_rescore (top-kprime, then FP32
rescoring to k) or :fp16.  rescore = :fp16 (Float16 `B`, or fp16 = true)
rescores the kprime candidates against the fp16 columns instead of the e4m3
ones.
"""
mutable struct HipTopKEngine
    index::Ptr{Cvoid}
    device::Int
    k::Int
    kprime::Int
    mode::Symbol
    rows_cap::Int
    rdim::Int
    N::Int
end

function HipTopKEngine(B; k::Integer = 20, kprime::Integer = 32, mode::Symbol = :fp8_rescore,
                       rows_cap::Integer = 2^13, fp16::Bool = mode === :fp16, host::Bool = false,
                       rescore::Symbol = :e4m3, device::Integer = 0)
    rdim, N = size(B)
    1 <= k <= N || throw(ArgumentError("need 1 <= k <= N (got k = $k, N = $N)"))
    haskey(HIP_MODES, mode) || throw(ArgumentError("mode must be :fp8, :fp8_rescore or :fp16"))
    rescore in (:e4m3, :fp16) || throw(ArgumentError("rescore must be :e4m3 or :fp16"))
    esz = sizeof(eltype(B))
    esz in (1, 2) || throw(ArgumentError("B must hold e4m3 (1-byte) or Float16 elements"))
    src = host ? Ptr{Cvoid}(pointer(B)) : hip_devptr(B)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    if esz == 2  # the verbatim fp16 index (+ its e4m3 quantisation)
        GC.@preserve B hip_check(ccall((:rm_index_create_f16, HIP_LIB), Cint,
                                       (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Cint, Ref{Ptr{Cvoid}}),
                                       hip_ctx(device), src, rdim, N,
                                       host ? HIP_INDEX_HOST : Cint(0), out))
    else
        flags = (host ? HIP_INDEX_HOST : Cint(0)) | (fp16 ? HIP_INDEX_FP16 : Cint(0))
        GC.@preserve B hip_check(ccall((:rm_index_create, HIP_LIB), Cint,
                                       (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Cint, Ref{Ptr{Cvoid}}),
                                       hip_ctx(device), src, rdim, N, flags, out))
    end
    eng = HipTopKEngine(out[], Int(device), Int(k), Int(max(kprime, k)), mode, Int(rows_cap),
                        Int(rdim), Int(N))
    finalizer(e -> (e.index != C_NULL && ccall((:rm_index_destroy, HIP_LIB), Cint, (Ptr{Cvoid},),
                                               e.index); e.index = C_NULL), eng)
    rescore === :fp16 && hip_check(ccall((:rm_index_set_rescore, HIP_LIB), Cint,
                                         (Ptr{Cvoid}, Cint), eng.index, HIP_F16))
    return eng
end

"""
    hip_autotune!(eng, bucket_rows) -> Vector{NamedTuple}

Time the candidate tile plans (query rows per tile, column splits, XCD query
grouping) for batches of each size in `bucket_rows` on the engine's index and
keep the fastest per size bucket in the device's context; later launches of
that bucket use it.  Results never depend on the plan.  Returns the chosen
(tile, splits, qgroup, ms) per bucket.
"""
function hip_autotune!(eng::HipTopKEngine, bucket_rows::AbstractVector{<:Integer})
    nb = length(bucket_rows)
    rows = Int64.(bucket_rows)
    tile, sp, qg = zeros(Int32, nb), zeros(Int32, nb), zeros(Int32, nb)
    ms = zeros(Float32, nb)
    hip_check(ccall((:rm_autotune, HIP_LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Int64}, Cint, Cint, Cint, Cint,
                     Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Float32}),
                    hip_ctx(eng.device), eng.index, rows, nb, HIP_MODES[eng.mode], eng.k,
                    eng.kprime, tile, sp, qg, ms))
    return map(i -> (tile = Int(tile[i]), splits = Int(sp[i]), qgroup = Int(qg[i]), ms = ms[i]), 1:nb)
end

"""
    hip_plan!(eng, M; tile = 0, splits = 0, qgroup = 0)

Pin the tile plan of M's size bucket (tile 64 / 128 / 256 query rows or 0 for
the default, splits 0 for the built-in rule); tile = splits = 0 clears it.
"""
hip_plan!(eng::HipTopKEngine, M::Integer; tile::Integer = 0, splits::Integer = 0,
          qgroup::Integer = 0) =
    hip_check(ccall((:rm_plan_set, HIP_LIB), Cint, (Ptr{Cvoid}, Int64, Cint, Cint, Cint, Cint),
                    hip_ctx(eng.device), M, HIP_MODES[eng.mode], tile, splits, qgroup))

"""
    batch_gemm_topk_hip!(D_val, D_loc, eng, Q; Q32 = nothing, stream = C_NULL)

One fused launch (batch_gemm_topk!, engine_fp8.jl:399-437): Q (nb, rdim)
device rows -- e4m3 for the FP8 modes, Float16 for :fp16 -- against all N
columns; Q32 (nb, rdim) Float32 rows for :fp8_rescore.  D_val / D_loc
(>= nb, k) column-major receive the descending top-k and 1-based columns.
"""
function batch_gemm_topk_hip!(D_val, D_loc, eng::HipTopKEngine, Q; Q32 = nothing,
                              stream::Ptr{Cvoid} = C_NULL)
    nb = size(Q, 1)
    nb <= eng.rows_cap || throw(ArgumentError("batch height $nb > engine rows_cap $(eng.rows_cap)"))
    eng.mode === :fp8_rescore && Q32 === nothing &&
        throw(ArgumentError(":fp8_rescore needs the Float32 queries Q32"))
    # Q is row-major (nb, rdim): a Julia (rdim, nb) matrix viewed transposed
    hip_check(ccall((:rm_topk_batch, HIP_LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Int64, Cint, Cint,
                     Cint, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cvoid}),
                    hip_ctx(eng.device), eng.index, hip_devptr(Q), eng.rdim, hip_devptr(Q32),
                    eng.rdim, nb, eng.k, eng.kprime, HIP_MODES[eng.mode], hip_devptr(D_val),
                    hip_devptr(D_loc), size(D_val, 1), stream))
    return D_val, D_loc
end

"""
    map_words_hip!(D_val, D_loc, D_norms, eng, re, dwords, nb)

The fused flow of one batch of nb packed fragments (device words): encode ->
GEMM -> top-k (-> rescore) in one call (rm_map_batch), D_norms (m, nb) or
nothing.
"""
function map_words_hip!(D_val, D_loc, D_norms, eng::HipTopKEngine, re, dwords, nb::Integer;
                        stream::Ptr{Cvoid} = C_NULL)
    nb <= eng.rows_cap || throw(ArgumentError("batch height $nb > engine rows_cap $(eng.rows_cap)"))
    length(dwords) == nb * cld(re.k, 16) ||
        throw(ArgumentError("dwords must hold nb * cld(k, 16) words"))
    hip_check(ccall((:rm_map_batch, HIP_LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Cint, Cint, Cint, Cint, Cint, Cint,
                     Cint, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Ptr{Cvoid}),
                    hip_ctx(eng.device), eng.index, hip_devptr(dwords), nb, re.k, re.s, re.m, re.c,
                    eng.k, eng.kprime, HIP_MODES[eng.mode], hip_devptr(D_val), hip_devptr(D_loc),
                    size(D_val, 1), hip_devptr(D_norms), stream))
    return D_val, D_loc
end

"""
    topk_flow_hip(source, eng; alloc, to_host, out_cap = 2) -> Channel

topk_flow (engine_fp8.jl:498-547) for the fused engine: for every batch of
`source` (fields embeds (nb, rdim) Float16 host, norms, heads, first -- a
RopeRealBatch{Float16} stream), upload the embeddings with `alloc` (a
function host matrix -> device array, e.g. AMDGPU's ROCArray), quantise them
to e4m3 on the device (upload_fp16_as_f8!'s e4m3(fp16)), run one fused
launch and emit (vals, locs, norms, heads, first) with host matrices
(`to_host` = Array for AMDGPU).  `alloc(T, dims...)` makes an empty device
array.
"""
function topk_flow_hip(source, eng::HipTopKEngine; alloc, to_host, out_cap::Integer = 2)
    out = Channel{Any}(out_cap)
    task = Threads.@spawn try
        for bat in source
            emb = bat.embeds
            eltype(emb) == Float16 ||
                throw(ArgumentError("batch eltype $(eltype(emb)) != Float16 (build the rope " *
                                    "stream with fp16 = true)"))
            nb = size(emb, 1)
            nb == 0 && continue
            nb <= eng.rows_cap || throw(ArgumentError("batch height $nb > engine rows_cap $(eng.rows_cap)"))
            q16 = alloc(permutedims(emb))            # (rdim, nb): row-major (nb, rdim) rows
            Dv = alloc(Float32, nb, eng.k)
            Dl = alloc(Int32, nb, eng.k)
            if eng.mode === :fp16
                batch_gemm_topk_hip!(Dv, Dl, eng, q16)
            else
                q8 = alloc(UInt8, eng.rdim, nb)
                f32_to_f8_hip!(q8, q16; device = eng.device)
                q32 = eng.mode === :fp8_rescore ? alloc(Float32.(permutedims(emb))) : nothing
                batch_gemm_topk_hip!(Dv, Dl, eng, q8; Q32 = q32)
            end
            put!(out, (vals = to_host(Dv), locs = to_host(Dl), norms = bat.norms,
                       heads = bat.heads, first = bat.first))
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
    rope_topk_stream_hip(re, file, eng; reader, alloc, to_host, batch_size = eng.rows_cap)

rope_topk_stream (engine_fp8.jl:576-594) for the fused engine, with the
encoder fused in as well: `reader(file; k)` yields fragments (header,
words) -- the reference's fasta_reads or RotorMapHIP.fasta_reads_hip --
gathered into batches of packed words, uploaded (`alloc`) and mapped in one
rm_map_batch call each; emits (vals, locs, norms, heads, first).
"""
function rope_topk_stream_hip(re, file::AbstractString, eng::HipTopKEngine; reader, alloc,
                              to_host, batch_size::Integer = eng.rows_cap, out_cap::Integer = 2)
    batch_size <= eng.rows_cap || throw(ArgumentError("batch_size > engine rows_cap"))
    W = cld(re.k, 16)
    out = Channel{Any}(out_cap)
    task = Threads.@spawn try
        first = 1
        words = Matrix{UInt32}(undef, W, batch_size)
        heads = String[]
        flush!(nb) = begin
            dw = alloc(vec(words[:, 1:nb]))
            Dv = alloc(Float32, nb, eng.k)
            Dl = alloc(Int32, nb, eng.k)
            Dn = alloc(Float32, re.m, nb)
            map_words_hip!(Dv, Dl, Dn, eng, re, dw, nb)
            put!(out, (vals = to_host(Dv), locs = to_host(Dl), norms = to_host(Dn),
                       heads = copy(heads), first = first))
            first += nb
            empty!(heads)
        end
        nb = 0
        for f in reader(file; k = re.k)
            nb += 1
            copyto!(view(words, :, nb), f.words)
            push!(heads, f.header)
            if nb == batch_size
                flush!(nb)
                nb = 0
            end
        end
        nb > 0 && flush!(nb)
    catch e
        close(out, e)
        rethrow()
    finally
        isopen(out) && close(out)
    end
    bind(out, task)
    return out
end
