"""Streaming flows on the GPU (FASTA file -> batches), following the
reference's stream checks (RotorMap/encode/stream.jl:274-299,
RotorMap/search/engine_fp8.jl:836-883): batch shapes, `first` indices and
headers in file order, and every streamed row equal to a one-shot
computation over the same reads -- the fused rope_topk_stream, the
reference's two-stage flow (fp16 embeddings -> e4m3 on the device ->
top-k), and rope_encode_real_stream's embeddings against the batch encoder."""
import numpy as np
import pytest

import rotormap_amd as R

pytestmark = pytest.mark.gpu

K = 20_000
NREADS = 1700
BATCH = 700


@pytest.fixture(scope="module")
def setup(torch_gpu, tmp_path_factory):
    ref = R.synth_reference([1_000_000, 700_000], gc=0.41, seed=61)
    re = R.RopeEncoder(k=K, s=5, m=1, c=5)
    idx = R.build_index(ref, re, kstep=1_000)
    reads = R.sample_reads(ref, NREADS, K, err=0.05, seed=62, draw="random", model=1)
    heads = reads.heads()
    path = tmp_path_factory.mktemp("fa") / "reads.fa"
    acgt = np.frombuffer(b"ACGT", np.uint8)
    with open(path, "wb") as f:
        for h, w in zip(heads, reads.words):
            seq = acgt[R.unpack_words(w, K)].tobytes()
            f.write(h.encode() + b"\n")
            for i in range(0, K, 70):
                f.write(seq[i:i + 70] + b"\n")
    return ref, re, idx, reads, heads, str(path)


def oneshot(torch, eng, re, words):
    dw = torch.from_numpy(np.ascontiguousarray(words).view(np.int32).reshape(-1).copy()).cuda()
    v, l = eng.map_words(re, dw, words.shape[0])
    torch.cuda.synchronize()
    return v.cpu().numpy(), l.cpu().numpy()


def check_batches(bats, heads):
    firsts = [b.first for b in bats]
    assert firsts == list(range(1, NREADS + 1, BATCH))
    assert [len(b.heads) for b in bats] == [BATCH, BATCH, NREADS - 2 * BATCH]
    assert sum((b.heads for b in bats), []) == heads  # file order, headers intact


@pytest.mark.parametrize("mode", ["fp8_rescore", "fp8", "fp16"])
def test_rope_topk_stream_fused_equals_oneshot(torch_gpu, setup, mode):
    ref, re, idx, reads, heads, path = setup
    eng = R.TopKEngine(R.ResidentIndex(idx.embeds8, fp16=(mode == "fp16")), k=20, kprime=32,
                       mode=mode, rows_cap=NREADS)
    v0, l0 = oneshot(torch_gpu, eng, re, reads.words)
    bats = list(R.rope_topk_stream(re, path, eng, batch_size=BATCH))
    check_batches(bats, heads)
    v = np.concatenate([b.vals for b in bats])
    l = np.concatenate([b.locs for b in bats])
    np.testing.assert_array_equal(l, l0)
    np.testing.assert_array_equal(v, v0)
    for b in bats:
        assert b.norms.shape == (1, len(b.heads)) and np.all(b.norms > 0)
        assert np.all(np.diff(b.vals, axis=1) <= 0)
    res = R.count_mapped(l, reads.recs, reads.starts, idx.rec_of_col, idx.starts, K)
    assert res["correct"] >= 0.99 * NREADS


@pytest.mark.parametrize("mode", ["fp8", "fp16"])
def test_two_stage_flow_equals_fused(torch_gpu, setup, mode):
    """rope_encode_real_stream(fp16=True) -> topk_flow quantises the fp16
    embeddings to e4m3 on the device: the same queries as the fused path,
    so the rows are identical (FP8 and FP16 modes)."""
    ref, re, idx, reads, heads, path = setup
    B = R.ResidentIndex(idx.embeds8, fp16=(mode == "fp16"))
    eng = R.TopKEngine(B, k=20, mode=mode, rows_cap=BATCH)
    a = list(R.rope_topk_stream(re, path, eng, batch_size=BATCH, fused=False))
    b = list(R.rope_topk_stream(re, path, eng, batch_size=BATCH, fused=True))
    check_batches(a, heads)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x.locs, y.locs)
        np.testing.assert_array_equal(x.vals, y.vals)
        np.testing.assert_array_equal(x.norms, y.norms)


@pytest.mark.parametrize("fp16", [False, True])
def test_rope_encode_real_stream_matches_batch_encoder(torch_gpu, setup, fp16):
    torch = torch_gpu
    ref, re, idx, reads, heads, path = setup
    bats = list(R.rope_encode_real_stream(re, path, batch_size=BATCH, fp16=fp16))
    check_batches(bats, heads)
    emb = np.concatenate([b.embeds for b in bats])
    assert emb.dtype == (np.float16 if fp16 else np.float32) and emb.shape == (NREADS, 2048)
    dw = torch.from_numpy(reads.words.view(np.int32).reshape(-1).copy()).cuda()
    dest = torch.zeros(NREADS, 2048, dtype=torch.float16 if fp16 else torch.float32, device="cuda")
    R.encode_frag_real_batch(dest, None, re, dw)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(emb, dest.cpu().numpy())
    # unit energy per row, the reference's streamed-batch check (stream.jl:274-287)
    e = np.sum(emb.astype(np.float64) ** 2, axis=1)
    np.testing.assert_allclose(e, 1.0, rtol=1e-2 if fp16 else 1e-3)


def test_stream_argument_errors(torch_gpu, setup):
    ref, re, idx, reads, heads, path = setup
    eng = R.TopKEngine(idx.embeds8, k=20, mode="fp8", rows_cap=100)
    with pytest.raises(R.RotorMapArgumentError):
        next(R.rope_topk_stream(re, path, eng, batch_size=BATCH))  # batch > rows_cap
    with pytest.raises(R.RotorMapArgumentError):
        next(R.rope_encode_real_stream(re, path, k=K - 1))  # k != re.k
    with pytest.raises(R.RotorMapArgumentError):  # fp32 embeddings into the e4m3 stage
        next(R.topk_flow(R.rope_encode_real_stream(re, path, batch_size=50, fp16=False), eng))
