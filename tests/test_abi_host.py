"""CPU-only checks of the C ABI and the host logic: every function the header
declares is exported and bound, argument errors come back as status codes
without touching a device, and the host-side sampler / mutation / trace
generators follow their stated closed forms."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import rotormap_amd as R
from rotormap_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rotormap_hip.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rm_\w+)\s*\(", text)))


def test_every_header_function_is_exported_and_bound():
    names = header_functions()
    assert len(names) >= 40, names
    lib = C.CDLL(L.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    unbound = [n for n in names if n not in L.SIGNATURES]
    assert not unbound, f"no ctypes signature: {unbound}"
    extra = [n for n in L.SIGNATURES if n not in names]
    assert not extra, f"bound but not declared: {extra}"


JULIA_GLUE = ["RotorMapHIP.jl", "RotorMap/common/hip.jl", "RotorMap/encode/kernel_hip.jl",
              "RotorMap/gemm/fp8_hip.jl", "RotorMap/search/engine_hip.jl"]


@pytest.mark.parametrize("path", JULIA_GLUE)
def test_julia_glue_calls_declared_functions(path):
    """Julia is not in this image, so the glue cannot run here: every symbol it
    ccalls must at least be declared in the header and exported, with as many
    argument types as the C declaration has parameters; blocks balance."""
    glue = os.path.join(ROOT, "qpg-rotormap_amd", "julia", path)
    text = open(glue).read()
    calls = re.findall(r"ccall\(\(:(rm_\w+),\s*\w+\),\s*\w+,\s*\(([^)]*)\)", text, re.S)
    assert calls, path
    declared = header_functions()
    lib = C.CDLL(L.LIB_PATH)
    for name, argt in calls:
        assert name in declared, name
        assert hasattr(lib, name), name
        nargs = len([a for a in argt.replace("\n", " ").split(",") if a.strip()])
        assert nargs == len(L.SIGNATURES[name][1]), (path, name, nargs)
    # every opening block keyword has its `end` (a crude balance check)
    code = re.sub(r'"""(?:.|\n)*?"""', "", text)
    code = re.sub(r"#.*", "", code)
    code = re.sub(r'"(?:[^"\\]|\\.)*"', '""', code)
    opens = len(re.findall(r"\b(function|module|mutable struct|struct|if|for|while|try|do|begin|let)\b", code))
    ends = len(re.findall(r"\bend\b", code)) - len(re.findall(r"\[[^\]\n]*\bend\b[^\]\n]*\]", code))
    assert opens == ends, (path, opens, ends)


def test_header_and_integration_example_compile_as_c(tmp_path):
    """The header is plain C (C callers and FFI generators read it), and the C
    example of the repository's integration guide compiles against it with no
    implicit declarations or pointer mismatches."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no gcc")
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```c\n(.*?)```", text, re.S)
    assert blocks, "the integration guide has no C example"
    body = "\n".join(l for l in blocks[0].splitlines() if not l.startswith("#include"))
    src = tmp_path / "example.c"
    src.write_text("#include <stddef.h>\n#include <stdint.h>\n#include \"rotormap_hip.h\"\n"
                   "void example(const uint8_t *host_B, int64_t N, const uint32_t *d_words,\n"
                   "             int64_t n, float *d_vals, int32_t *d_locs) {\n" + body + "\n}\n")
    r = subprocess.run([gcc, "-std=c11", "-fsyntax-only", "-Wall", "-Werror",
                        "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_version_and_status_codes():
    assert b"gfx950" in L.lib().rm_version()
    text = open(HEADER).read()
    for name, val in [("RM_OK", 0), ("RM_ERR_ARG", 1), ("RM_ERR_SHAPE", 2), ("RM_ERR_NOMEM", 3),
                      ("RM_ERR_HIP", 4), ("RM_ERR_STATE", 5),
                      ("RM_SCHED_FIFO", L.RM_SCHED_FIFO), ("RM_SCHED_SRPT", L.RM_SCHED_SRPT)]:
        assert re.search(rf"#define\s+{name}\s+{val}\b", text), name


def test_null_arguments_are_rejected_without_a_device():
    lib = L.lib()
    assert lib.rm_encode_batch(None, None, 1, 20000, 8, 4, 4, 0, None, 0, 0, 0, None, 0, 0, 0,
                               None, None, None) == L.RM_ERR_ARG
    assert "NULL" in L.last_error() or "argument" in L.last_error()
    assert lib.rm_map_batch(None, None, None, 1, 20000, 8, 4, 4, 20, 32, 1, None, None, 1, None,
                            None) == L.RM_ERR_ARG
    assert lib.rm_topk_batch(None, None, None, 2048, None, 2048, 1, 20, 32, 1, None, None, 1,
                             None) == L.RM_ERR_ARG
    assert lib.rm_server_wait(None, 0, 0, None, None, None, None, None) == L.RM_ERR_ARG
    assert lib.rm_server_set_policy(None, L.RM_SCHED_SRPT) == L.RM_ERR_ARG
    assert lib.rm_server_set_lowload(None, 4096, 16384) == L.RM_ERR_ARG
    assert lib.rm_index_set_rescore(None, L.RM_F16) == L.RM_ERR_ARG
    assert lib.rm_index_rescore(None) == -1
    assert lib.rm_rowtopk_merge(None, None, None, 1, None, L.RM_F32, 1, 1, 1, 0, 20,
                                None) == L.RM_ERR_ARG
    assert lib.rm_plan_set(None, 100, L.RM_MODE_FP8, 64, 0, 0) == L.RM_ERR_ARG
    assert lib.rm_server_destroy(None) == L.RM_OK
    assert lib.rm_index_cols(None) == 0
    assert lib.rm_device_free_bytes(None) == -1


def test_phase_table_bad_arguments():
    out = np.zeros(16, np.int32)
    assert L.lib().rm_phase_table(10, 17, 1, L.ptr(out)) != L.RM_OK


def test_batch_bytes_model():
    """Per-batch device bytes: packed words + encoded queries (+ fp32 copy
    for rescoring) + split partials at the 256-split bound and the splits'
    shared bounds + outputs."""
    M, k, rdim = 8192, 20000, 2048
    W = (k + 15) // 16
    for mode, qb, extra, kp in [(L.RM_MODE_FP8, 1, 0, 20), (L.RM_MODE_FP8_RESCORE, 1, 4, 32),
                                (L.RM_MODE_FP16, 2, 0, 20)]:
        want = M * W * 4 + M * rdim * qb + M * rdim * extra + M * 256 * kp * 8 + M * 8 + M * 20 * 8
        assert L.lib().rm_batch_bytes(M, k, rdim, 20, 32, mode) == want


def test_even_alloc_largest_remainder():
    w = np.array([7, 0, 13, 5, 1])
    for n in [0, 1, 2, 5, 26, 100, 1001]:
        a = R.even_alloc(w, n)
        assert a.sum() == n and a[1] == 0
        exact = n * w / w.sum()
        assert np.all(np.abs(a - exact) < 1.0)
        # largest remainders take the leftovers: no record with a larger
        # remainder is left below one with a smaller remainder
        fl = np.floor(exact).astype(int)
        got = a - fl
        rem = exact - fl
        for i in range(len(w)):
            for j in range(len(w)):
                if got[i] == 1 and got[j] == 0:
                    assert rem[i] >= rem[j] - 1e-12


def test_sample_reads_even_strata_midpoints():
    """:even puts read t of record j at rank floor((2t-1) W_j / (2 n_j)) of
    its N-free window pool (RotorMap/reads/sample.jl:563-580)."""
    ref = R.synth_reference([50_000, 30_000], gc=0.41, seed=9)
    k = 1000
    rs = R.sample_reads(ref, 40, k, draw="even")
    Wj = np.array([r.length - k + 1 for r in ref.records])
    alloc = R.even_alloc(Wj, 40)
    for j in range(2):
        st = rs.starts[rs.recs == j]
        t = np.arange(1, alloc[j] + 1)
        np.testing.assert_array_equal(st, (2 * t - 1) * Wj[j] // (2 * alloc[j]) + 1)
        codes = R.unpack_words(ref.records[j].words[:-1], ref.records[j].length)
        for i in np.where(rs.recs == j)[0][:5]:
            s = rs.starts[i]
            np.testing.assert_array_equal(R.unpack_words(rs.words[i], k), codes[s - 1:s - 1 + k])


def test_sample_reads_skip_n_runs():
    ref = R.synth_reference([200_000], gc=0.41, seed=3, n_frac=0.2, n_run=(5_000, 20_000))
    rec = ref.records[0]
    assert rec.n_runs
    rs = R.sample_reads(ref, 200, 2000, draw="random", seed=4)
    for s in rs.starts:
        for lo, hi in rec.n_runs:
            assert s - 1 + 2000 <= lo or s - 1 >= hi


@pytest.mark.parametrize("model", [0, 1])
def test_mutate_counts(model):
    """n_err = ceil(k err); model 0 substitutes only (each hit changes the
    base); model 1 splits n_err into n_ins = n_del = floor(n_err/3) and
    substitutions, keeping the length (RotorMap/common/dna.jl:120-183)."""
    k, n = 20000, 16
    ref = R.synth_reference([100_000], gc=0.5, seed=1)
    clean = R.sample_reads(ref, n, k, err=0.0, draw="random", seed=2)
    mut = R.sample_reads(ref, n, k, err=0.05, draw="random", seed=2, model=model)
    np.testing.assert_array_equal(clean.starts, mut.starts)
    n_err = math.ceil(k * 0.05)
    for i in range(n):
        a = R.unpack_words(clean.words[i], k)
        b = R.unpack_words(mut.words[i], k)
        diff = int(np.sum(a != b))
        if model == 0:
            assert 0.9 * n_err <= diff <= n_err
        else:
            assert diff > n_err  # indels shift the tail
    again = R.sample_reads(ref, n, k, err=0.05, draw="random", seed=2, model=model)
    np.testing.assert_array_equal(again.words, mut.words)


def test_poisson_trace_statistics():
    t, n = R.poisson_trace(20000, rate=200.0, seed=1)
    assert np.all(np.diff(t) >= 0) and t[0] == 0
    assert abs(len(t) / t[-1] - 200.0) < 10.0
    assert n.min() >= 1 and n.max() <= 100_000
    mean = 100_000 / math.log(100_001)  # E of floor(exp(U(0, ln(1e5+1)))) ~ 8,686
    assert abs(n.mean() - mean) / mean < 0.05


def test_extract_windows_host_matches_unpack():
    ref = R.synth_reference([5000], gc=0.41, seed=11)
    rec = ref.records[0]
    k = 100
    starts = np.array([0, 1, 15, 16, 17, 4899, 4900], dtype=np.int64)
    out = np.zeros((len(starts), (k + 15) // 16), np.uint32)
    L.check(L.lib().rm_extract_windows_host(L.ptr(out), L.ptr(rec.words), L.ptr(starts),
                                            len(starts), k))
    codes = R.unpack_words(rec.words[:-1], rec.length)
    for i, s in enumerate(starts):
        np.testing.assert_array_equal(R.unpack_words(out[i], k), codes[s:s + k])


def test_index_dir_round_trip(tmp_path):
    """save_index_dir / load_index_dir (the julia/export_index.jl layout)."""
    rng = np.random.default_rng(3)
    N, rd, m = 40, 2048, 1
    heads = [f">chr{1 + (i % 3)} synthetic record" for i in range(N)]
    d = R.IndexData(rng.integers(0, 256, (N, rd), dtype=np.uint8), rng.random((N, m), dtype=np.float32),
                    heads, np.arange(1, N + 1, dtype=np.int64) * 7, (np.arange(N) % 2).astype(np.uint8),
                    20000, 2000, 5, m, 5, 0, N // 2, "/data/ref.fa", "indexreal.fp8.v1")
    p = str(tmp_path / "idx.rmidx")
    R.save_index_dir(d, p)
    e = R.load_index_dir(p)
    np.testing.assert_array_equal(e.embeds8, d.embeds8)
    np.testing.assert_array_equal(e.norms, d.norms)
    np.testing.assert_array_equal(e.starts, d.starts)
    np.testing.assert_array_equal(e.strand, d.strand)
    assert e.heads == heads and e.rec_names == ["chr1", "chr2", "chr3"]
    assert e.rec_of_col.tolist() == [i % 3 for i in range(N)]
    assert (e.k, e.kstep, e.s, e.m, e.c, e.n_frag, e.source) == (20000, 2000, 5, 1, 5, 20, "/data/ref.fa")
    import json
    meta = json.load(open(tmp_path / "idx.rmidx" / "meta.json"))
    meta["normalize"] = 1
    json.dump(meta, open(tmp_path / "idx.rmidx" / "meta.json", "w"))
    with pytest.raises(R.RotorMapArgumentError):
        R.load_index_dir(p)
