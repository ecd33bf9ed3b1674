# export_index.jl -- write a RotorMap index .bin (Julia-serialized NamedTuple,
# RotorMap/index/build.jl:755-768) as raw little-endian arrays that
# rotormap_amd.load_index_dir reads without Julia:
#
#   meta.json      format, source, k, kstep, s, m, c, normalize, n_frag, rdim, N,
#                  norms_shape
#   embeds8.u8     "indexreal.fp8.v1": the (rdim, N) e4m3 matrix as stored,
#                  i.e. N columns of rdim bytes
#   embeds16.f16   "indexreal.v1": the (N, rdim) fp16 rows, row-major
#   norms.f32      the norms matrix, column-major as stored
#   starts.i64     1-based window starts (N)
#   strand.u8      0 fwd / 1 rc (N)
#   heads.txt      one verbatim header per column
#
# Run inside the RotorMap workbench (its layers included, so `load_index`,
# RotorMap/index/load.jl:49-73, is defined):
#   include("export_index.jl"); export_index("index.bin", "index.rmidx")

function _json_str(s::AbstractString)
    io = IOBuffer()
    print(io, '"')
    for c in s
        if c == '"' || c == '\\'
            print(io, '\\', c)
        elseif c < ' '
            print(io, "\\u", string(UInt16(c), base = 16, pad = 4))
        else
            print(io, c)
        end
    end
    print(io, '"')
    return String(take!(io))
end

function export_index(bin::AbstractString, dir::AbstractString)
    d = load_index(String(bin))
    mkpath(dir)
    fp8 = d.format == "indexreal.fp8.v1"
    N = 2 * d.n_frag
    rdim = 2 * d.m * 4^d.c
    open(joinpath(dir, "meta.json"), "w") do io
        print(io, "{\"format\": ", _json_str(d.format), ", \"source\": ", _json_str(d.source),
              ", \"k\": ", d.k, ", \"kstep\": ", d.kstep, ", \"s\": ", d.s, ", \"m\": ", d.m,
              ", \"c\": ", d.c, ", \"normalize\": ", d.normalize, ", \"n_frag\": ", d.n_frag,
              ", \"rdim\": ", rdim, ", \"N\": ", N, ", \"norms_shape\": [",
              size(d.norms, 1), ", ", size(d.norms, 2), "]}")
    end
    if fp8
        write(joinpath(dir, "embeds8.u8"), reinterpret(UInt8, vec(d.embeds8)))
    else
        write(joinpath(dir, "embeds16.f16"), vec(permutedims(Float16.(d.embeds))))
    end
    write(joinpath(dir, "norms.f32"), vec(Float32.(d.norms)))
    write(joinpath(dir, "starts.i64"), Int64.(d.starts))
    write(joinpath(dir, "strand.u8"), UInt8.(d.strand))
    open(joinpath(dir, "heads.txt"), "w") do io
        for h in d.heads
            println(io, h)
        end
    end
    return dir
end
