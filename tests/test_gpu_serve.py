"""Serving loop on the GPU: every request's rows come back exactly as a
one-shot rm_map_batch of the same reads gives them (a row's top-k does not
depend on the launch it rides in), whether the request is alone, split over
launches or packed with others; admission, cancellation and error paths;
a short Poisson trace with latency stamps in order."""
import numpy as np
import pytest

import rotormap_amd as R
from rotormap_amd import _lib as L

pytestmark = pytest.mark.gpu

K, KSTEP = 20_000, 2_000


@pytest.fixture(scope="module")
def setup(torch_gpu):
    ref = R.synth_reference([600_000, 400_000], gc=0.5, seed=2024)
    re = R.RopeEncoder(k=K, s=8, m=4, c=4)
    idx = R.build_index(ref, re, kstep=KSTEP)
    reads = R.sample_reads(ref, 3000, K, err=0.05, seed=5, draw="random", model=0)
    return ref, re, idx, reads


def oneshot(torch, idx, re, words, mode):
    eng = R.TopKEngine(idx.embeds8, k=20, kprime=32, mode=mode)
    n = words.shape[0]
    dw = torch.from_numpy(words.view(np.int32).reshape(-1).copy()).cuda()
    v, l = eng.map_words(re, dw, n)
    torch.cuda.synchronize()
    return v.cpu().numpy(), l.cpu().numpy()


@pytest.mark.parametrize("mode", ["fp8_rescore", "fp8", "fp16"])
def test_server_matches_oneshot(torch_gpu, setup, mode):
    ref, re, idx, reads = setup
    v0, l0 = oneshot(torch_gpu, idx, re, reads.words, mode)
    # 700-row launches: requests split over launches and packed together
    sizes = [1, 0, 650, 1400, 3, 700, 246]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    assert offs[-1] == 3000
    with R.Server(R.ResidentIndex(idx.embeds8, fp16=(mode == "fp16")), re, mode=mode,
                  max_rows=700, n_streams=2) as srv:
        ids = [srv.submit(reads.words[offs[i]:offs[i + 1]], copy=(i % 2 == 0))
               for i in range(len(sizes))]
        out = [srv.wait(rid, n) for rid, n in ids]
        st = srv.stats()
    for i, r in enumerate(out):
        np.testing.assert_array_equal(r.vals, v0[offs[i]:offs[i + 1]])
        np.testing.assert_array_equal(r.locs, l0[offs[i]:offs[i + 1]])
        assert r.t_submit <= r.t_admit <= r.t_finish
    assert st["rows_done"] == 3000 and st["launches"] >= 5
    print(f"[serve] {mode}: {st}")


def test_server_alone_equals_mixed(torch_gpu, setup):
    ref, re, idx, reads = setup
    w = reads.words[100:400]
    with R.Server(idx.embeds8, re, max_rows=4096, n_streams=1) as srv:
        a = srv.wait(*srv.submit(w))
    with R.Server(idx.embeds8, re, max_rows=512, n_streams=3) as srv:
        ids = [srv.submit(reads.words[:100]), srv.submit(w), srv.submit(reads.words[400:1500])]
        b = [srv.wait(*x) for x in ids][1]
    np.testing.assert_array_equal(a.vals, b.vals)
    np.testing.assert_array_equal(a.locs, b.locs)


@pytest.mark.parametrize("policy", ["srpt", "fifo"])
def test_server_policy_order(torch_gpu, setup, policy):
    """One launch slot of 256 rows; a 1024-row request, then a 10-row one.
    Shortest-remaining-first puts the small request in the first or second
    launch, so it finishes strictly before the large one (which needs four).
    In arrival order its rows follow all of the large request's.  Both
    orders return the one-shot rows."""
    ref, re, idx, reads = setup
    v0, l0 = oneshot(torch_gpu, idx, re, reads.words[:1034], "fp8_rescore")
    with R.Server(idx.embeds8, re, max_rows=256, n_streams=1, policy=policy) as srv:
        big = srv.submit(reads.words[:1024])
        small = srv.submit(reads.words[1024:1034])
        b, s = srv.wait(*big), srv.wait(*small)
        st = srv.stats()
    np.testing.assert_array_equal(b.vals, v0[:1024])
    np.testing.assert_array_equal(b.locs, l0[:1024])
    np.testing.assert_array_equal(s.vals, v0[1024:])
    np.testing.assert_array_equal(s.locs, l0[1024:])
    if policy == "srpt":
        assert s.t_finish < b.t_finish
    else:
        assert s.t_finish >= b.t_finish
    assert st["rows_done"] == 1034 and st["launches"] == 5


def test_server_srpt_matches_oneshot(torch_gpu, setup):
    ref, re, idx, reads = setup
    v0, l0 = oneshot(torch_gpu, idx, re, reads.words, "fp8_rescore")
    sizes = [1, 0, 650, 1400, 3, 700, 246]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    with R.Server(idx.embeds8, re, max_rows=700, n_streams=2, policy="srpt") as srv:
        ids = [srv.submit(reads.words[offs[i]:offs[i + 1]]) for i in range(len(sizes))]
        out = [srv.wait(rid, n) for rid, n in ids]
        with pytest.raises(ValueError):
            srv.set_policy("lifo")
    for i, r in enumerate(out):
        np.testing.assert_array_equal(r.vals, v0[offs[i]:offs[i + 1]])
        np.testing.assert_array_equal(r.locs, l0[offs[i]:offs[i + 1]])
        assert r.t_submit <= r.t_admit <= r.t_finish


def test_server_lowload_launch_cap(torch_gpu, setup):
    """Low-load mode: while at most backlog_rows rows wait unlaunched a launch
    holds at most launch_rows rows, one at a time: a 3,000-row request on
    2,048-row slots runs as one full launch (backlog 3,000) and then 256-row
    slices (backlog <= 1,000): 5 launches.  Without it, 2.  Rows equal the
    one-shot rows either way."""
    ref, re, idx, reads = setup
    v0, l0 = oneshot(torch_gpu, idx, re, reads.words, "fp8_rescore")
    with R.Server(idx.embeds8, re, max_rows=2048, n_streams=2, policy="srpt",
                  lowload=(256, 1000)) as srv:
        out = srv.wait(*srv.submit(reads.words))
        assert srv.stats()["launches"] == 5
        srv.set_lowload(0, 0)
        out2 = srv.wait(*srv.submit(reads.words))
        assert srv.stats()["launches"] == 7
        with pytest.raises(R.RotorMapArgumentError):
            srv.set_lowload(-1, 0)
    for r in (out, out2):
        np.testing.assert_array_equal(r.vals, v0)
        np.testing.assert_array_equal(r.locs, l0)


def test_server_lowload_whole_requests(torch_gpu, setup):
    """Low-load mode packs whole requests: small requests submitted behind a
    large one ride launches of their own and finish before it, with one
    launch in flight at a time; every row equals its one-shot row."""
    ref, re, idx, reads = setup
    v0, l0 = oneshot(torch_gpu, idx, re, reads.words, "fp8_rescore")
    spans = [(0, 2500), (2500, 2800), (2800, 2810), (2810, 3000)]
    with R.Server(idx.embeds8, re, max_rows=1024, n_streams=2, policy="srpt",
                  lowload=(512, 1 << 20)) as srv:
        ids = [srv.submit(reads.words[a:b]) for a, b in spans]
        out = [srv.wait(rid, n) for rid, n in ids]
        st = srv.stats()
    assert st["max_inflight"] == 1
    assert st["launches"] >= 5 + 1  # the large one alone needs 5 slices of 512
    big_done = out[0].t_finish
    for r in out[1:]:
        assert r.t_finish < big_done
    for (a, b), r in zip(spans, out):
        np.testing.assert_array_equal(r.vals, v0[a:b])
        np.testing.assert_array_equal(r.locs, l0[a:b])


def test_server_admission_and_errors(torch_gpu, setup):
    ref, re, idx, reads = setup
    B = R.ResidentIndex(idx.embeds8)
    with pytest.raises(R.RotorMapError) as e:
        R.Server(B, re, max_rows=65536, n_streams=2, hbm_budget=1 << 20)
    assert e.value.code == L.RM_ERR_NOMEM and "bytes available" in str(e.value)
    with pytest.raises(R.RotorMapArgumentError):
        R.Server(B, R.RopeEncoder(k=K, s=5, m=1, c=4), max_rows=64)  # rdim 512 != 2048
    with pytest.raises(R.RotorMapArgumentError):
        R.Server(B, re, max_rows=0)
    with R.Server(B, re, max_rows=256, n_streams=1) as srv:
        with pytest.raises(R.RotorMapArgumentError):
            srv.wait(12345, 1)
        with pytest.raises(R.RotorMapArgumentError):
            srv.submit(reads.words[:3].reshape(-1)[:-1])  # not a multiple of W
        # a long queue behind one slot: the last request is still queued
        ids = [srv.submit(reads.words[:2000]) for _ in range(4)]
        last = srv.submit(reads.words[:10])
        srv.cancel(last[0])
        with pytest.raises(R.RotorMapError) as e:
            srv.wait(*last)
        assert "cancelled" in str(e.value)
        for x in ids:
            assert srv.wait(*x).locs.shape == (2000, 20)


def test_server_trace(torch_gpu, setup):
    ref, re, idx, reads = setup
    pool = np.ascontiguousarray(reads.words)
    t, n = R.poisson_trace(40, rate=200.0, seed=3, max_reads=2000)
    with R.Server(idx.embeds8, re, max_rows=4096, n_streams=2) as srv:
        res = R.run_trace(srv, pool, t, n, seed=4, keep_results=True)
    assert res["requests"] == 40 and res["reads"] == int(np.minimum(n, 3000).sum())
    assert 0 < res["p50_ms"] <= res["p99_ms"] <= res["p999_ms"]
    v0, l0 = oneshot(torch_gpu, idx, re, reads.words, "fp8_rescore")
    for r, o, m in zip(res["results"], res["offsets"], np.minimum(n, 3000)):
        np.testing.assert_array_equal(r.locs, l0[o:o + m])
    print(f"[serve] trace p50 {res['p50_ms']:.2f} ms p99 {res['p99_ms']:.2f} ms")


def test_server_after_plan_change(torch_gpu, setup):
    """A server sizes its slots' split partials from the plans of its
    creation; a plan set later (rm_plan_set / rm_autotune) with more splits
    than a slot holds runs with the split count that fits, and the rows
    still equal one-shot (the split count never changes a result)."""
    ref, re, idx, reads = setup
    B = R.ResidentIndex(idx.embeds8)
    eng = R.TopKEngine(B, k=20, kprime=32, mode="fp8_rescore")
    v0, l0 = oneshot(torch_gpu, idx, re, reads.words[:1500], "fp8_rescore")
    try:
        with R.Server(B, re, max_rows=700, n_streams=2) as srv:
            for M in (700, 100):  # the buckets of a full slot and of the tail launch
                eng.set_plan(M, 64, 256, 0)
            out = srv.wait(*srv.submit(reads.words[:1500]))
    finally:
        for M in (700, 100):
            eng.set_plan(M, 0, 0, 0)
    np.testing.assert_array_equal(out.vals, v0)
    np.testing.assert_array_equal(out.locs, l0)
