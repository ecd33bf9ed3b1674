"""The human-scale configuration on the GPU: a 3.1 Gbp synthetic reference
(24 records of 50-250 Mbp, GC 0.41, 5 % of bases in N runs of 10 kbp-1 Mbp),
s5m1c5, k = 20,000, kstep = 2,000 -> N = 3,086,568 e4m3 columns, 6.3 GB
resident.

- Exact reads (err 0, :even over N-free windows) all map, in every mode.
- A read's row alone equals its row inside a 4,096-read batch.
- FP64 spot check of 4 reads over all 3.09 M columns against the GPU's own
  e4m3 queries (fp32 accumulation error only: atol 1e-5), and the
  reference's model (golden FP64 encoder, atol 5e-3, RotorMap/search/
  engine_fp8.jl:866-883).
- Uneven batches: after rm_autotune, one rm_map_batches launch over
  concatenated batches equals per-batch rm_map_batch, bit for bit; so does a
  concatenation large enough to run as several launches.
- Serving: a short Poisson trace; every request equals the one-shot rows."""
import os

import numpy as np
import pytest

import oracle.oracle as O
import rotormap_amd as R

pytestmark = pytest.mark.gpu

K, KSTEP = 20_000, 2_000
HUMAN_MBP = [248, 242, 198, 190, 181, 171, 159, 145, 138, 134, 135, 133, 114, 107, 102, 90, 83,
             80, 59, 64, 50, 51, 156, 57]
NB = 4096


@pytest.fixture(scope="module")
def human(torch_gpu):
    torch = torch_gpu
    ref = R.synth_reference([mb * 1_000_000 for mb in HUMAN_MBP], gc=0.41, seed=300,
                            n_frac=0.05, n_run=(10_000, 1_000_000))
    re = R.RopeEncoder(k=K, s=5, m=1, c=5)
    idx = R.build_index(ref, re, kstep=KSTEP, batch_size=8192)
    assert idx.embeds8.shape[0] > 3_000_000
    exact = R.sample_reads(ref, NB, K, err=0.0, seed=21, draw="even", model=1)
    mixed = R.sample_reads(ref, NB, K, err=np.random.default_rng(5).uniform(0, 0.10, NB), seed=22,
                           draw="random", model=1)
    B = R.ResidentIndex(idx.embeds8)
    dw = torch.from_numpy(exact.words.view(np.int32).reshape(-1).copy()).cuda()
    eng = R.TopKEngine(B, k=20, kprime=32, mode="fp8_rescore", rows_cap=NB)
    v, l = eng.map_words(re, dw, NB)
    torch.cuda.synchronize()
    return ref, re, idx, exact, mixed, B, dw, (v.cpu().numpy(), l.cpu().numpy())


def test_human_exact_reads_all_map(torch_gpu, human):
    torch = torch_gpu
    ref, re, idx, exact, mixed, B, dw, (v, l) = human
    res = R.count_mapped(l, exact.recs, exact.starts, idx.rec_of_col, idx.starts, K)
    print(f"[cfg3] fp8_rescore exact reads: {res['correct']}/{res['total']}, rank-1 "
          f"{res['rank_hist'][0]}")
    assert res["correct"] == NB
    eng8 = R.TopKEngine(B, k=20, mode="fp8", rows_cap=NB)
    v8, l8 = eng8.map_words(re, dw, NB)
    torch.cuda.synchronize()
    r8 = R.count_mapped(l8.cpu().numpy(), exact.recs, exact.starts, idx.rec_of_col, idx.starts, K)
    assert r8["correct"] == NB
    # FP16 over the exact fp16 widening (12.6 GB more): the first 1,024 reads
    B16 = R.ResidentIndex(idx.embeds8, fp16=True)
    e16 = R.TopKEngine(B16, k=20, mode="fp16")
    n16 = 1024
    v16, l16 = e16.map_words(re, dw[:n16 * ((K + 15) // 16)], n16)
    torch.cuda.synchronize()
    r16 = R.count_mapped(l16.cpu().numpy(), exact.recs[:n16], exact.starts[:n16], idx.rec_of_col,
                         idx.starts, K)
    assert r16["correct"] == n16
    del e16, B16
    torch.cuda.empty_cache()


def test_human_alone_equals_in_batch(torch_gpu, human):
    torch = torch_gpu
    ref, re, idx, exact, mixed, B, dw, (v, l) = human
    W = (K + 15) // 16
    eng = R.TopKEngine(B, k=20, kprime=32, mode="fp8_rescore")
    for lo, n in ((0, 1), (777, 5), (NB - 70, 70)):
        va, la = eng.map_words(re, dw[lo * W:(lo + n) * W], n)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(va.cpu().numpy(), v[lo:lo + n])
        np.testing.assert_array_equal(la.cpu().numpy(), l[lo:lo + n])


def test_human_fp64_spot_check(torch_gpu, human):
    torch = torch_gpu
    ref, re, idx, exact, mixed, B, dw, _ = human
    W = (K + 15) // 16
    rows = [3, 1500, 2900, 4000]
    mw = torch.from_numpy(mixed.words[rows].view(np.int32).reshape(-1).copy()).cuda()
    eng = R.TopKEngine(B, k=20, mode="fp8")
    v, l = eng.map_words(re, mw, len(rows))
    e8 = torch.empty(len(rows), 2048, dtype=torch.uint8, device="cuda")
    R.encode_frag_real_batch(e8, None, re, mw)
    torch.cuda.synchronize()
    v, l = v.cpu().numpy(), l.cpu().numpy()
    q8 = O.e4m3_table()[e8.cpu().numpy()].astype(np.float64)
    fv, fl = O.scan_topk_f64(q8, idx.embeds8, 21, os.cpu_count() or 1)
    for i in range(len(rows)):
        np.testing.assert_allclose(v[i], fv[i, :20], atol=1e-5, rtol=0)
        if fv[i, 19] - fv[i, 20] > 2e-5:
            assert set(l[i].tolist()) == set(fl[i, :20].tolist()), i
        qg = O.golden_query(mixed.words[rows[i]], K, 5, 1, 5, "e4m3")
        assert np.all(np.abs(O.dots_f64(qg, idx.embeds8, l[i]) - v[i]) <= 5e-3), i
    print(f"[cfg3] fp64 spot check: max |gpu - fp64| = {np.max(np.abs(v - fv[:, :20])):.2e}")


def test_human_uneven_batches_after_autotune(torch_gpu, human):
    torch = torch_gpu
    ref, re, idx, exact, mixed, B, dw, _ = human
    sizes = np.array([1, 9, 64, 200, 1, 1500, 130, 2191])
    assert sizes.sum() == NB
    words = np.ascontiguousarray(mixed.words)
    eng = R.TopKEngine(B, k=20, kprime=32, mode="fp8_rescore", rows_cap=NB)
    buckets = np.array([1, 16, 64, 256, 2048, 4096])
    W = (K + 15) // 16
    try:
        tile, splits, qg, ms = eng.autotune(buckets)
        print(f"[cfg4] tuned tiles {tile.tolist()} splits {splits.tolist()} qg {qg.tolist()} "
              f"ms {np.round(ms, 2).tolist()}")
        offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        dwm = torch.from_numpy(words.view(np.int32).reshape(-1).copy()).cuda()
        vb, lb = eng.map_batches(re, dwm, offs)
        parts = []
        for b in range(len(sizes)):
            parts.append(eng.map_words(re, dwm[offs[b] * W:offs[b + 1] * W], int(sizes[b])))
        torch.cuda.synchronize()
    finally:
        for b in buckets:
            eng.set_plan(int(b), 0, 0, 0)
    vb, lb = vb.cpu().numpy(), lb.cpu().numpy()
    for b, (v, l) in enumerate(parts):
        np.testing.assert_array_equal(v.cpu().numpy(), vb[offs[b]:offs[b + 1]], err_msg=f"batch {b}")
        np.testing.assert_array_equal(l.cpu().numpy(), lb[offs[b]:offs[b + 1]], err_msg=f"batch {b}")
    res = R.count_mapped(lb, mixed.recs, mixed.starts, idx.rec_of_col, idx.starts, K)
    print(f"[cfg4] recall {res['correct']}/{res['total']}")
    assert res["correct"] >= 0.99 * NB


def test_human_uneven_batches_cut_into_launches(torch_gpu, human):
    # 20,480 rows in uneven batches: above one launch's 64 x CUs rows, so the
    # concatenation runs as the launch rule's launches; every batch must equal
    # rm_map_batch on it alone (plans differ, results may not)
    torch = torch_gpu
    ref, re, idx, exact, mixed, B, dw, _ = human
    reps = 5
    n = reps * NB
    words = np.ascontiguousarray(np.concatenate([mixed.words] * reps))
    sizes = np.array([3, 16381, 1, n - 16385])
    eng = R.TopKEngine(B, k=20, kprime=32, mode="fp8_rescore", rows_cap=n)
    lr = eng.launch_rows(n)
    print(f"[cfg4] {n} rows as launches of {lr}")
    assert 0 < lr < n, lr
    W = (K + 15) // 16
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    dwm = torch.from_numpy(words.view(np.int32).reshape(-1).copy()).cuda()
    vb, lb = eng.map_batches(re, dwm, offs)
    parts = [eng.map_words(re, dwm[offs[b] * W:offs[b + 1] * W], int(sizes[b]))
             for b in range(len(sizes))]
    torch.cuda.synchronize()
    vb, lb = vb.cpu().numpy(), lb.cpu().numpy()
    for b, (v, l) in enumerate(parts):
        np.testing.assert_array_equal(v.cpu().numpy(), vb[offs[b]:offs[b + 1]], err_msg=f"batch {b}")
        np.testing.assert_array_equal(l.cpu().numpy(), lb[offs[b]:offs[b + 1]], err_msg=f"batch {b}")
    # the repeated reads map as they did in the 4,096-read batch
    for r in range(reps):
        np.testing.assert_array_equal(lb[r * NB:(r + 1) * NB], lb[:NB], err_msg=f"copy {r}")


def test_human_serving_trace(torch_gpu, human):
    torch = torch_gpu
    ref, re, idx, exact, mixed, B, dw, _ = human
    pool = np.ascontiguousarray(mixed.words)
    eng = R.TopKEngine(B, k=20, kprime=32, mode="fp8_rescore", rows_cap=NB)
    dwm = torch.from_numpy(pool.view(np.int32).reshape(-1).copy()).cuda()
    v0, l0 = eng.map_words(re, dwm, NB)
    torch.cuda.synchronize()
    v0, l0 = v0.cpu().numpy(), l0.cpu().numpy()
    t, n = R.poisson_trace(16, rate=50.0, seed=31, max_reads=NB)
    with R.Server(B, re, max_rows=2048, n_streams=2, policy="srpt") as srv:
        res = R.run_trace(srv, pool, t, n, seed=32, keep_results=True)
    print(f"[cfg5] {res['requests']} requests, {res['reads']} reads: p50 {res['p50_ms']:.1f} ms "
          f"p99 {res['p99_ms']:.1f} ms")
    for r, o, m in zip(res["results"], res["offsets"], np.minimum(n, NB)):
        np.testing.assert_array_equal(r.vals, v0[o:o + m])
        np.testing.assert_array_equal(r.locs, l0[o:o + m])
