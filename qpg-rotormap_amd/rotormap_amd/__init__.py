"""rotormap_amd: RotorMap's mapping hot path on MI355X (gfx950).

encoder -> reads x index GEMM (FP8 e4m3 / FP16 MFMA, FP32 accumulate) ->
fused per-row top-k -> FP32 rescoring, in librotormap_hip.so (C ABI,
include/rotormap_hip.h); this package is the host-side mirror of the
reference's Julia layers (encode/, gemm/, search/, index/, reads/).
"""
from ._lib import (RM_E4M3, RM_F16, RM_F32, RM_MODE_FP8, RM_MODE_FP8_RESCORE, RM_MODE_FP16,
                   RotorMapArgumentError, RotorMapError, ctx, lib)
from .encode import (RopeEncoder, dim, encode_frag_real_batch, encode_frag_real_batch_alloc,
                     phase_table, rdim, revcomp_frag_words)
from .gemm import f8_to_f32, f32_to_f8, fp8_gemm_hip, launch_seg_rowtopk_merge
from .fasta import fasta_reads_gpu, iter_fasta_gpu
from .genome import Reference, Record, pack_codes, read_fasta, synth_reference, unpack_words, write_fasta
from .index import (IndexData, build_index, load_index, load_index_bin, load_index_dir, save_index,
                    save_index_bin, save_index_dir, window_table)
from .reads import ReadSet, even_alloc, parse_read_head, sample_reads
from .serve import Server, ServedRequest, poisson_trace, run_trace
from .search import MODES, ResidentIndex, TopKBatch, TopKEngine, count_mapped
from .stream import RopeRealBatch, rope_encode_real_stream, rope_topk_stream, topk_flow

__all__ = [n for n in dir() if not n.startswith("_")]
