"""Tile plans and uneven batches on the GPU (small index): the autotuner's
plans, pinned plans of every tile shape, and one launch over concatenated
uneven batches all give each row exactly the result of rm_map_batch on its
batch alone."""
import numpy as np
import pytest

import rotormap_amd as R

pytestmark = pytest.mark.gpu

K, KSTEP = 20_000, 1_000


@pytest.fixture(scope="module")
def setup(torch_gpu):
    ref = R.synth_reference([1_500_000, 900_000], gc=0.41, seed=77)
    re = R.RopeEncoder(k=K, s=5, m=1, c=5)
    idx = R.build_index(ref, re, kstep=KSTEP)
    reads = R.sample_reads(ref, 6000, K, err=0.05, seed=78, draw="random", model=1)
    return ref, re, idx, reads


def per_batch(torch, eng, re, words, sizes):
    W = words.shape[1]
    vs, ls = [], []
    for lo, n in zip(np.concatenate([[0], np.cumsum(sizes)[:-1]]), sizes):
        dw = torch.from_numpy(words[lo:lo + n].view(np.int32).reshape(-1).copy()).cuda()
        v, l = eng.map_words(re, dw, int(n))
        torch.cuda.synchronize()
        vs.append(v.cpu().numpy())
        ls.append(l.cpu().numpy())
    return np.concatenate(vs), np.concatenate(ls)


@pytest.mark.parametrize("mode", ["fp8_rescore", "fp8", "fp16"])
def test_autotune_then_map_batches_bitexact(torch_gpu, setup, mode):
    torch = torch_gpu
    ref, re, idx, reads = setup
    sizes = np.array([1, 37, 300, 2, 1200, 64, 65, 129, 700, 3])
    words = np.ascontiguousarray(reads.words[:int(sizes.sum())])
    eng = R.TopKEngine(R.ResidentIndex(idx.embeds8, fp16=(mode == "fp16")), k=20, kprime=32,
                       mode=mode)
    buckets = np.array([1, 2, 4, 64, 128, 256, 512, 1024, 2048, 4096])
    try:
        v_ref, l_ref = per_batch(torch, eng, re, words, sizes)  # default plans
        tile, splits, qg, ms = eng.autotune(buckets)
        assert set(tile.tolist()) <= {64, 128, 256} and np.all(splits >= 1) and np.all(ms > 0)
        for b, t in zip(buckets, tile):
            assert eng.plan(int(b)) == (int(t), int(splits[list(buckets).index(b)]),
                                        int(qg[list(buckets).index(b)]))
        v1, l1 = per_batch(torch, eng, re, words, sizes)  # tuned plans
        offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        dw = torch.from_numpy(words.view(np.int32).reshape(-1).copy()).cuda()
        v2, l2 = eng.map_batches(re, dw, offs)
        torch.cuda.synchronize()
        v2, l2 = v2.cpu().numpy(), l2.cpu().numpy()
        print(f"[plans] {mode}: tiles {tile.tolist()} splits {splits.tolist()} qg {qg.tolist()}")
    finally:
        for b in buckets:
            eng.set_plan(int(b), 0, 0, 0)
    for v, l, what in ((v1, l1, "tuned per-batch"), (v2, l2, "map_batches")):
        np.testing.assert_array_equal(l, l_ref, err_msg=what)
        np.testing.assert_array_equal(v, v_ref, err_msg=what)


def test_pinned_tiles_end_to_end(torch_gpu, setup):
    """Every tile shape, pinned for a 129-row batch (three partial 64-row
    tiles, two 128-row, one 256-row), returns the default plan's rows."""
    torch = torch_gpu
    ref, re, idx, reads = setup
    eng = R.TopKEngine(idx.embeds8, k=20, kprime=32, mode="fp8_rescore")
    words = np.ascontiguousarray(reads.words[:129])
    v0, l0 = per_batch(torch, eng, re, words, [129])
    try:
        for tile in (64, 128, 256):
            for S in (1, 5):
                eng.set_plan(129, tile, S, 0)
                v, l = per_batch(torch, eng, re, words, [129])
                np.testing.assert_array_equal(l, l0, err_msg=f"tile {tile} S {S}")
                np.testing.assert_array_equal(v, v0, err_msg=f"tile {tile} S {S}")
    finally:
        eng.set_plan(129, 0, 0, 0)
