"""Encoder parity on the GPU: bit-exact against the oracle's fixed-point
encoder (integer histograms, the fp32 normalisation, storage roundings),
within 4 ulp of 1.0 of the FP64 golden restatement of
RotorMap/encode/reference.jl."""
import numpy as np
import pytest

import oracle.oracle as O
import rotormap_amd as R

pytestmark = pytest.mark.gpu

# the reference's encoder configurations and fragment lengths
# (RotorMap/legacy/test/ropeflowreal.jl:684-712)
CONFIGS = [(8, 4, 4), (5, 1, 4), (16, 2, 4), (8, 1, 6), (5, 1, 5)]
KS = [17, 2000, 2001]


def random_words(n, k, seed):
    rng = np.random.default_rng(seed)
    W = (k + 15) // 16
    w = rng.integers(0, 2 ** 32, size=(n, W), dtype=np.uint64).astype(np.uint32)
    r = k & 15
    if r:
        w[:, -1] &= np.uint32((1 << (2 * r)) - 1)
    return w


def gpu_encode(torch, words, re, normalize, dtype, layout="row"):
    n = words.shape[0]
    dw = torch.from_numpy(words.view(np.int32).reshape(-1).copy()).cuda()
    rd = R.rdim(re)
    tdt = {"f32": torch.float32, "f16": torch.float16, "e4m3": torch.uint8}[dtype]
    if layout == "row":
        dest = torch.zeros(n, rd, dtype=tdt, device="cuda")
    else:  # the reference's column-major (n, 2D)
        dest = torch.zeros(rd, n, dtype=tdt, device="cuda").t()
    norms = torch.zeros(n, re.m, dtype=torch.float32, device="cuda").t()
    counts = torch.zeros(n, 4 ** re.c, dtype=torch.int32, device="cuda")
    R.encode_frag_real_batch(dest, norms, re, dw, normalize, counts=counts)
    torch.cuda.synchronize()
    out = dest.cpu().numpy()
    if dtype == "f16":
        out = out.view(np.float16)
    return np.ascontiguousarray(out), norms.cpu().numpy().T.copy(), counts.cpu().numpy()


@pytest.mark.parametrize("cfg", CONFIGS)
@pytest.mark.parametrize("k", KS)
@pytest.mark.parametrize("normalize", [0, 1, 2, 3])
def test_encoder_bitexact_fixed_point(torch_gpu, cfg, k, normalize):
    s, m, c = cfg
    if s > k:
        pytest.skip("s > k")
    re = R.RopeEncoder(k=k, s=s, m=m, c=c)
    words = random_words(6, k, seed=1000 + k + 7 * s + normalize)
    out, norms, counts = gpu_encode(torch_gpu, words, re, normalize, "f32")
    for f in range(words.shape[0]):
        codes = O.frag_codes(words[f], k)
        ref, rnorm = O.rope_fixed(codes, k, s, m, c, normalize)
        np.testing.assert_array_equal(out[f], ref.astype(np.float32))
        exp_n = rnorm.astype(np.float32)
        np.testing.assert_array_equal(norms[f], exp_n)
        np.testing.assert_array_equal(counts[f], O.smer_counts(codes, k, s, c))
        # tolerance pin against the FP64 golden encoder: 4 ulp of 1.0 (2^-22)
        # elementwise on the unit-energy outputs, 2^-21 relative on the norms
        # (the reference's own bounds for its fp32 kernel are rtol 1e-2, atol
        # 2e-4: RotorMap/encode/kernel.jl:366-369)
        gold, gnorm = O.ref_rope_frag_real(codes, k, s, m, c, normalize)
        np.testing.assert_allclose(out[f], gold, rtol=0, atol=2.0 ** -22)
        rows = [0] if normalize == 0 else range(m)
        for i in rows:
            np.testing.assert_allclose(norms[f][i], gnorm[i], rtol=2.0 ** -21)


@pytest.mark.parametrize("dtype", ["f16", "e4m3"])
@pytest.mark.parametrize("layout", ["row", "col"])
def test_encoder_storage_roundings(torch_gpu, dtype, layout):
    k, (s, m, c) = 20000, (8, 4, 4)
    re = R.RopeEncoder(k=k, s=s, m=m, c=c)
    words = random_words(4, k, seed=77)
    out, norms, _ = gpu_encode(torch_gpu, words, re, 0, dtype, layout)
    for f in range(4):
        ref, _ = O.rope_fixed(O.frag_codes(words[f], k), k, s, m, c, 0)
        exp = O.store_chain(ref, dtype)
        if dtype == "f16":
            np.testing.assert_array_equal(out[f].view(np.uint16), exp.view(np.uint16))
        else:
            np.testing.assert_array_equal(out[f], exp)


def test_encoder_k20000_both_index_configs(torch_gpu):
    for (s, m, c) in [(8, 4, 4), (5, 1, 5)]:
        k = 20000
        re = R.RopeEncoder(k=k, s=s, m=m, c=c)
        words = random_words(3, k, seed=s)
        out, norms, counts = gpu_encode(torch_gpu, words, re, 0, "f32")
        for f in range(3):
            codes = O.frag_codes(words[f], k)
            ref, rn = O.rope_fixed(codes, k, s, m, c, 0)
            np.testing.assert_array_equal(out[f], ref.astype(np.float32))
            assert counts[f].sum() == k - s + 1  # closed form (i)
            e = float(np.sum(out[f].astype(np.float64) ** 2))
            assert abs(e - 1.0) < 1e-5  # unit energy (RotorMap/encode/kernel.jl:378-383)


@pytest.mark.parametrize("dtype", ["f32", "f16"])
@pytest.mark.parametrize("normalize", [0, 1, 2, 3])
def test_encoder_unit_energy_every_row(torch_gpu, dtype, normalize):
    """The reference's streamed-batch check (RotorMap/encode/stream.jl:274-287):
    every row has unit energy over its [Re; Im] halves, the whole row for
    normalize 0, else the largest copy (stride m within each half), to rtol
    1e-3 for fp32 storage and 1e-2 for fp16."""
    k, (s, m, c) = 2001, (8, 4, 4)
    re = R.RopeEncoder(k=k, s=s, m=m, c=c)
    words = random_words(96, k, seed=31 + normalize)
    out, _, _ = gpu_encode(torch_gpu, words, re, normalize, dtype)
    x = out.astype(np.float64)
    if normalize == 0:
        e = np.sum(x * x, axis=1)
    else:
        e = np.max(np.stack([np.sum(x[:, i::m] ** 2, axis=1) for i in range(m)]), axis=0)
    np.testing.assert_allclose(e, 1.0, rtol=1e-3 if dtype == "f32" else 1e-2)


def test_encoder_deterministic_across_batches(torch_gpu):
    k, (s, m, c) = 2001, (8, 4, 4)
    re = R.RopeEncoder(k=k, s=s, m=m, c=c)
    words = random_words(64, k, seed=5)
    a, _, _ = gpu_encode(torch_gpu, words, re, 0, "f32")
    b, _, _ = gpu_encode(torch_gpu, words[17:18].copy(), re, 0, "f32")
    np.testing.assert_array_equal(a[17], b[0])


def test_revcomp_and_extract(torch_gpu):
    torch = torch_gpu
    for k in (17, 2000, 2001, 20000):
        words = random_words(5, k, seed=k)
        dw = torch.from_numpy(words.view(np.int32).reshape(-1).copy()).cuda()
        rc = torch.empty_like(dw)
        R.revcomp_frag_words(rc, dw, k, 5)
        got = rc.cpu().numpy().view(np.uint32).reshape(5, -1)
        for f in range(5):
            exp = O.forward_pack(O.revcomp_codes(O.frag_codes(words[f], k)))
            np.testing.assert_array_equal(got[f], exp)
    # window slicing from a packed record
    rng = np.random.default_rng(3)
    codes = rng.integers(0, 4, size=50_000).astype(np.uint8)
    rec = R.pack_codes(codes)
    starts = np.array([0, 1, 15, 16, 17, 33, 29_000], dtype=np.int64)
    for k in (17, 2000, 20000):
        st = torch.from_numpy(starts).cuda()
        out = torch.empty(len(starts) * ((k + 15) // 16), dtype=torch.int32, device="cuda")
        drec = torch.from_numpy(rec.view(np.int32)).cuda()
        R._lib.check(R.lib().rm_extract_windows(R.ctx(0), R._lib.ptr(out), R._lib.ptr(drec),
                                                R._lib.ptr(st), len(starts), k, None))
        torch.cuda.synchronize()
        got = out.cpu().numpy().view(np.uint32).reshape(len(starts), -1)
        for i, s0 in enumerate(starts):
            np.testing.assert_array_equal(got[i], O.forward_pack(codes[s0:s0 + k]))


def test_quantize_e4m3_exhaustive(torch_gpu):
    """Device e4m3 bits equal the host conversion, as the reference's
    _check_quantize_path asserts (RotorMap/search/engine_fp8.jl:600-617),
    here over every binary16 value and beyond the saturation point."""
    torch = torch_gpu
    # every binary16 value (as fp32) plus random fp32 values incl. saturation
    h = np.arange(65536, dtype=np.uint16).view(np.float16).astype(np.float32)
    rng = np.random.default_rng(0)
    x = np.concatenate([h, rng.standard_normal(100_000).astype(np.float32) * 50,
                        np.array([448, 449, 464, 465, 1e9, -1e9, 2 ** -10, 2 ** -9 * 1.5],
                                 dtype=np.float32)])
    x = x[np.isfinite(x)]
    src = torch.from_numpy(x).cuda()
    dst = torch.empty(len(x), dtype=torch.uint8, device="cuda")
    R.f32_to_f8(dst, src)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dst.cpu().numpy(), O.f32_to_e4m3(x))
    back = torch.empty(256, dtype=torch.float32, device="cuda")
    R.f8_to_f32(back, torch.arange(256, dtype=torch.uint8, device="cuda"))
    torch.cuda.synchronize()
    got, exp = back.cpu().numpy(), O.e4m3_table()
    nan = np.isnan(exp)  # codes 0x7F, 0xFF: NaN (any payload)
    assert np.all(np.isnan(got[nan]))
    np.testing.assert_array_equal(got[~nan].view(np.uint32), exp[~nan].view(np.uint32))


@pytest.mark.parametrize("cfg", [(5, 1, 5), (8, 4, 4), (16, 2, 4)])
def test_encoder_homopolymer_runs_bitexact(torch_gpu, cfg):
    """Fragments with long single-base runs (N runs read as G, poly-A ...):
    waves whose 64 windows share one key take the ballot + cross-lane sum
    path; bins, counts and floats stay bit-exact against the oracle."""
    s, m, c = cfg
    k = 20000
    re = R.RopeEncoder(k=k, s=s, m=m, c=c)
    rng = np.random.default_rng(55 + s)
    frags = []
    for f in range(6):
        codes = rng.integers(0, 4, k).astype(np.uint8)
        for _ in range(4):
            lo = int(rng.integers(0, k - 400))
            codes[lo:lo + int(rng.integers(100, 6000))] = int(rng.integers(0, 4))
        frags.append(codes)
    frags.append(np.full(k, 0, np.uint8))  # all A: one key everywhere
    frags.append(np.full(k, 2, np.uint8))  # all G (an N run)
    words = np.stack([O.forward_pack(cd) for cd in frags])
    for normalize in (0, 1, 2, 3):
        out, norms, counts = gpu_encode(torch_gpu, words, re, normalize, "f32")
        for f, codes in enumerate(frags):
            ref, rnorm = O.rope_fixed(codes, k, s, m, c, normalize)
            np.testing.assert_array_equal(out[f], ref.astype(np.float32), err_msg=f"frag {f} n{normalize}")
            np.testing.assert_array_equal(norms[f], rnorm.astype(np.float32))
            np.testing.assert_array_equal(counts[f], O.smer_counts(codes, k, s, c))
