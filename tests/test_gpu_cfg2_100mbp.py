"""The FP8 100 Mbp configuration on the GPU: a seeded i.i.d. reference of
4 x 25 Mbp (GC 0.41), s5m1c5, k = 20,000, kstep = 2,000 -> N = 99,928 e4m3
columns; one fixed batch of 65,536 reads sampled :even, substitution rate
per read from {0, 0.02, 0.05, 0.10}.

- Recall floors per error rate in all three modes (err 0: every read).
- FP64 spot checks of 16 reads over all columns, as the reference's
  _check_topk_stream (RotorMap/search/engine_fp8.jl:836-883): the golden
  FP64 encoder's e4m3 query against the stored columns, top-20 values
  within atol 5e-3 and each returned column's FP64 score within 5e-3 of its
  value; and, tighter, against the FP64 scores of the GPU's own query
  operands in each mode (fp32 accumulation error only: atol 1e-5).
- A read's row alone equals its row inside the full batch."""
import os

import numpy as np
import pytest

import oracle.oracle as O
import rotormap_amd as R

pytestmark = pytest.mark.gpu

K, KSTEP = 20_000, 2_000
M = 65_536
ERRS = (0.0, 0.02, 0.05, 0.10)


@pytest.fixture(scope="module")
def cfg2(torch_gpu):
    torch = torch_gpu
    ref = R.synth_reference([25_000_000] * 4, gc=0.41, seed=100)
    re = R.RopeEncoder(k=K, s=5, m=1, c=5)
    idx = R.build_index(ref, re, kstep=KSTEP, batch_size=8192)
    assert idx.embeds8.shape == (99_928, 2048)
    rng = np.random.default_rng(7)
    errs = np.array(ERRS)[rng.integers(0, len(ERRS), M)]
    reads = R.sample_reads(ref, M, K, err=errs, seed=11, draw="even", model=0)
    dw = torch.from_numpy(reads.words.view(np.int32).reshape(-1).copy()).cuda()
    B = R.ResidentIndex(idx.embeds8, fp16=True)
    out = {}
    for mode in ("fp8_rescore", "fp8", "fp16"):
        eng = R.TopKEngine(B, k=20, kprime=32, mode=mode, rows_cap=M)
        v, l = eng.map_words(re, dw, M)
        torch.cuda.synchronize()
        out[mode] = (v.cpu().numpy(), l.cpu().numpy())
    return ref, re, idx, reads, errs, dw, B, out


def test_100mbp_recall_per_error_rate(cfg2):
    ref, re, idx, reads, errs, dw, B, out = cfg2
    floors = {0.0: 1.0, 0.02: 0.999, 0.05: 0.995, 0.10: 0.99}
    for mode, (v, l) in out.items():
        res = R.count_mapped(l, reads.recs, reads.starts, idx.rec_of_col, idx.starts, K)
        for e in ERRS:
            sel = errs == e
            rec = res["hit"][sel].mean()
            print(f"[cfg2] {mode} err {e}: top-20 recall {rec:.5f} over {sel.sum()} reads")
            assert rec >= floors[e], (mode, e, rec)
        assert np.all(np.diff(v, axis=1) <= 0), mode  # rows sorted descending
        assert l.min() >= 1 and l.max() <= idx.embeds8.shape[0]


def test_100mbp_fp64_spot_checks(torch_gpu, cfg2):
    torch = torch_gpu
    ref, re, idx, reads, errs, dw, B, out = cfg2
    rng = np.random.default_rng(13)
    rows = np.sort(rng.choice(M, 16, replace=False))
    W = (K + 15) // 16
    threads = os.cpu_count() or 1
    # (1) the reference's model: golden FP64 encoder -> e4m3 query vs the stored columns
    q_gold = np.stack([O.golden_query(reads.words[r], K, 5, 1, 5, "e4m3") for r in rows])
    refv, _ = O.scan_topk_f64(q_gold, idx.embeds8, 20, threads)
    v8, l8 = out["fp8"]
    for i, r in enumerate(rows):
        np.testing.assert_allclose(v8[r], refv[i], atol=5e-3, rtol=0)
        s = O.dots_f64(q_gold[i], idx.embeds8, l8[r])
        assert np.all(np.abs(s - v8[r]) <= 5e-3), r
    # (2) the GPU's own operands: the fused kernel's e4m3 / fp16 queries and
    # the rescoring's fp32 queries from the same encoder
    sub = torch.cat([dw[r * W:(r + 1) * W] for r in rows.tolist()])
    e8 = torch.empty(len(rows), 2048, dtype=torch.uint8, device="cuda")
    e16 = torch.empty(len(rows), 2048, dtype=torch.float16, device="cuda")
    e32 = torch.empty(len(rows), 2048, dtype=torch.float32, device="cuda")
    R.encode_frag_real_batch(e8, None, re, sub, dest2=e32)
    R.encode_frag_real_batch(e16, None, re, sub)
    torch.cuda.synchronize()
    q8 = O.e4m3_table()[e8.cpu().numpy()].astype(np.float64)
    q16 = e16.cpu().numpy().astype(np.float64)
    q32 = e32.cpu().numpy().astype(np.float64)
    deq16 = O.e4m3_table()[idx.embeds8].astype(np.float16)  # the exact widening
    for mode, Q, Bx in (("fp8", q8, idx.embeds8), ("fp16", q16, deq16)):
        fv, fl = O.scan_topk_f64(Q, Bx, 21, threads)
        v, l = out[mode]
        for i, r in enumerate(rows):
            np.testing.assert_allclose(v[r], fv[i, :20], atol=1e-5, rtol=0, err_msg=mode)
            s = O.dots_f64(Q[i], Bx, l[r])
            assert np.all(np.abs(s - v[r]) <= 1e-5), (mode, r)
            if fv[i, 19] - fv[i, 20] > 2e-5:  # a clear boundary: the same column set
                assert set(l[r].tolist()) == set(fl[i, :20].tolist()), (mode, r)
    vr, lr = out["fp8_rescore"]
    for i, r in enumerate(rows):
        s = O.dots_f64(q32[i], idx.embeds8, lr[r])
        assert np.all(np.abs(s - vr[r]) <= 1e-5), r


def test_100mbp_alone_equals_in_batch(torch_gpu, cfg2):
    torch = torch_gpu
    ref, re, idx, reads, errs, dw, B, out = cfg2
    W = (K + 15) // 16
    for mode in ("fp8_rescore", "fp8", "fp16"):
        eng = R.TopKEngine(B, k=20, kprime=32, mode=mode)
        for lo, n in ((0, 1), (1234, 7), (M - 300, 300)):
            v, l = eng.map_words(re, dw[lo * W:(lo + n) * W], n)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(v.cpu().numpy(), out[mode][0][lo:lo + n], err_msg=mode)
            np.testing.assert_array_equal(l.cpu().numpy(), out[mode][1][lo:lo + n], err_msg=mode)
