"""GEMM + fused top-k parity on the GPU.

The GEMM is checked against the FP64 product of the stored operands within
the reference's tolerances (RotorMap/gemm/fp8_ptx.jl:248-268).  The fused
top-k is checked BIT-EXACTLY: the fused kernel's accumulators are the same
MFMA sequence as the C-writing chunk kernel, so the oracle's total-order
top-k over the GPU's own C must equal the fused result value for value and
column for column, for every split count.  Shapes are the reference's top-k
test list (RotorMap/gemm/topk_kernels.jl:521-523)."""
import numpy as np
import pytest

import oracle.oracle as O
import rotormap_amd as R
from rotormap_amd import _lib as L

pytestmark = pytest.mark.gpu

SHAPES = [(64, 100), (257, 128), (1000, 333), (1, 4096), (511, 2048), (2048, 2 ** 14), (33, 37),
          (64, 25), (64, 20)]


def rand_e4m3(shape, scale, seed):
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal(shape) * scale).astype(np.float32)
    return O.f32_to_e4m3(x).reshape(shape)


def unit_rows_e4m3(n, K, seed):
    """Unit-norm rows quantised like the pipeline (fp32 -> fp16 -> e4m3)."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, K))
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    return O.store_chain(x.ravel(), "e4m3").reshape(n, K), x.astype(np.float32)


def gemm_c(torch, Q, B, out="f32"):
    M, K = Q.shape
    N = B.shape[0]
    dq = torch.from_numpy(Q).cuda()
    db = torch.from_numpy(B).cuda()
    C = torch.zeros(N, M, dtype=torch.float32 if out == "f32" else torch.float16, device="cuda").t()
    R.fp8_gemm_hip(C, dq, db, N)
    torch.cuda.synchronize()
    return C.cpu().numpy().astype(np.float32)


def test_gemm_chunk_fp8_vs_fp64(torch_gpu):
    torch = torch_gpu
    Q = rand_e4m3((256, 256), 0.05, 1)
    B = rand_e4m3((256, 256), 0.02, 2)
    ref = O.gemm_e4m3_f64(Q, B)
    C32 = gemm_c(torch, Q, B)
    rel = np.linalg.norm(C32 - ref) / np.linalg.norm(ref)
    assert rel < 1e-6, rel
    C16 = gemm_c(torch, Q, B, out="f16")
    assert np.linalg.norm(C16 - ref) / np.linalg.norm(ref) < 5e-3


def test_gemm_chunk_ragged_and_layouts(torch_gpu):
    torch = torch_gpu
    M, N, K = 300, 700, 2048
    Q = rand_e4m3((M, K), 0.03, 3)
    B = rand_e4m3((N, K), 0.03, 4)
    ref = O.gemm_e4m3_f64(Q, B)
    C = gemm_c(torch, Q, B)
    assert np.max(np.abs(C - ref)) < 1e-5
    # the reference's column-major A (a8: element (r, k) at r + k*M) gives the same bits
    dq_col = torch.from_numpy(np.ascontiguousarray(Q.T)).cuda().t()
    db = torch.from_numpy(B).cuda()
    C2 = torch.zeros(N, M, dtype=torch.float32, device="cuda").t()
    R.fp8_gemm_hip(C2, dq_col, db, N)
    # chunked B columns (the chunk loop's base-pointer offsets) == whole, bitwise
    C3 = torch.zeros(N, M, dtype=torch.float32, device="cuda").t()
    for lo in range(0, N, 256):
        wi = min(N, lo + 256) - lo
        R.fp8_gemm_hip(C3[:, lo:lo + wi], torch.from_numpy(Q).cuda(), db, wi, b_col0=lo)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(C2.cpu().numpy(), C)
    np.testing.assert_array_equal(C3.cpu().numpy(), C)


def test_gemm_chunk_fp16_vs_fp64(torch_gpu):
    torch = torch_gpu
    rng = np.random.default_rng(9)
    Q = (rng.standard_normal((200, 512)) * 0.05).astype(np.float16)
    B = (rng.standard_normal((333, 512)) * 0.05).astype(np.float16)
    ref = O.gemm_f16_f64(Q, B)
    C = torch.zeros(333, 200, dtype=torch.float32, device="cuda").t()
    R.fp8_gemm_hip(C, torch.from_numpy(Q).cuda(), torch.from_numpy(B).cuda(), 333)
    torch.cuda.synchronize()
    got = C.cpu().numpy()
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-6


TILES = (256, 128, 64)  # query rows per tile of the ring kernels (k' <= 32)


def fused(torch, eng, Q, Q32=None, splits=None, qg=0, tile=0):
    """One fused launch under a pinned plan (tile 0 = the default tile)."""
    M = Q.shape[0]
    pinned = splits is not None or tile
    if pinned:
        eng.set_plan(M, tile, splits or 0, qg)
        if tile:
            assert eng.plan(M)[0] == tile
    dq = torch.from_numpy(Q).cuda()
    d32 = torch.from_numpy(Q32).cuda() if Q32 is not None else None
    v, l = eng.batch_gemm_topk(dq, d32)
    torch.cuda.synchronize()
    if pinned:
        eng.set_plan(M, 0, 0, 0)
    return v.cpu().numpy(), l.cpu().numpy()


@pytest.mark.parametrize("M,N", SHAPES)
def test_fused_topk_bitexact(torch_gpu, M, N):
    torch = torch_gpu
    K = 512
    k = 20
    Q = rand_e4m3((M, K), 0.05, M)
    B = rand_e4m3((N, K), 0.05, N)
    C = gemm_c(torch, Q, B)
    ev, el = O.topk_rows(C, k)
    eng = R.TopKEngine(B, k=k, mode="fp8")
    tiles = -(-N // 256)
    for tile in TILES:
        for S in sorted({1, 2, 3, tiles}):
            if S > tiles:
                continue
            for qg in (0, 2, 3):  # query tiles grouped per XCD (0 = all)
                v, l = fused(torch, eng, Q, splits=S, qg=qg, tile=tile)
                msg = f"tile={tile}, splits={S}, qg={qg}"
                np.testing.assert_array_equal(l, el, err_msg="locs, " + msg)
                np.testing.assert_array_equal(v, ev, err_msg="vals, " + msg)


def test_fused_topk_ties_and_k64(torch_gpu):
    torch = torch_gpu
    # duplicate columns force exact ties: order must be (value desc, column asc)
    K, M = 512, 130
    base = rand_e4m3((40, K), 0.05, 11)
    B = np.concatenate([base] * 30)  # 1200 columns, each value 30 times
    Q = rand_e4m3((M, K), 0.05, 12)
    C = gemm_c(torch, Q, B)
    # k <= 32: the ring tiles of 256, 128 and 64 queries; 33..64: the
    # 128-query register-staged tile
    for k in (20, 32, 40, 48, 64):
        ev, el = O.topk_rows(C, k)
        eng = R.TopKEngine(B, k=k, mode="fp8")
        for tile in (TILES if k <= 32 else (0,)):
            for S in (1, 4):
                v, l = fused(torch, eng, Q, splits=S, tile=tile)
                np.testing.assert_array_equal(l, el, err_msg=f"k={k} tile={tile} S={S}")
                np.testing.assert_array_equal(v, ev, err_msg=f"k={k} tile={tile} S={S}")


def test_fused_topk_zero_score_ties(torch_gpu):
    """Scores of exactly +0 tie across hundreds of columns: every third index
    column is all zeros, the other columns score below zero (negative index
    codes against positive queries) except a few positive ones.  The top-k
    is the few positive columns, then the zero columns in column order, so
    the lists fill with equal values and every merge places candidates by
    the column rule alone."""
    torch = torch_gpu
    K, M, N = 256, 200, 3000
    B = rand_e4m3((N, K), 0.05, 31) | np.uint8(0x80)  # negative or -0 codes
    B[::3] = 0  # +0 columns
    B[[5, 700, 2999]] &= np.uint8(0x7F)  # a few positive columns
    Q = rand_e4m3((M, K), 0.05, 32) & np.uint8(0x7F)  # positive queries
    assert not np.isin(B, (0x7F, 0xFF)).any() and not np.isin(Q, (0x7F, 0xFF)).any()  # no NaN codes
    C = gemm_c(torch, Q, B)
    assert (C[:, ::3] == 0).all() and (np.signbit(C[:, ::3]) == 0).all()
    for k in (20, 32):
        ev, el = O.topk_rows(C, k)
        assert (ev[:, 3:] == 0).all()  # the ties on +0 fill every list past 3 entries
        eng = R.TopKEngine(B, k=k, mode="fp8")
        for tile in TILES:
            for S in (1, 4):
                v, l = fused(torch, eng, Q, splits=S, tile=tile)
                msg = f"k={k} tile={tile} S={S}"
                np.testing.assert_array_equal(l, el, err_msg="locs, " + msg)
                np.testing.assert_array_equal(v, ev, err_msg="vals, " + msg)
                assert (np.signbit(v) == np.signbit(ev)).all(), "sign of zero, " + msg


def test_fused_topk_fp16_bitexact(torch_gpu):
    torch = torch_gpu
    rng = np.random.default_rng(21)
    M, N, K = 300, 1500, 512
    Q = (rng.standard_normal((M, K)) * 0.05).astype(np.float16)
    B8 = rand_e4m3((N, K), 0.05, 22)
    B16 = O.e4m3_table()[B8].astype(np.float16)  # exact widening
    C = torch.zeros(N, M, dtype=torch.float32, device="cuda").t()
    R.fp8_gemm_hip(C, torch.from_numpy(Q).cuda(), torch.from_numpy(B16).cuda(), N)
    torch.cuda.synchronize()
    ev, el = O.topk_rows(C.cpu().numpy(), 20)
    eng = R.TopKEngine(B8, k=20, mode="fp16")
    for tile in TILES:
        for S in (1, 3):
            v, l = fused(torch, eng, Q, splits=S, tile=tile)
            np.testing.assert_array_equal(l, el, err_msg=f"tile={tile} S={S}")
            np.testing.assert_array_equal(v, ev, err_msg=f"tile={tile} S={S}")


@pytest.mark.parametrize("K", [64, 128, 192, 256, 320])
def test_fused_topk_short_k_bitexact(torch_gpu, K):
    """K of 1-5 64-byte stages (FP8) and 2-10 (FP16): stage counts that are
    not a multiple of 4 take the fused kernel's per-tile K loop, the others
    the loop that streams across tiles; both over 6 tiles with merges and
    split counts 1 and 3."""
    torch = torch_gpu
    M, N, k = 300, 1500, 20
    Q = rand_e4m3((M, K), 0.05, 50 + K)
    B = rand_e4m3((N, K), 0.05, 60 + K)
    ev, el = O.topk_rows(gemm_c(torch, Q, B), k)
    eng = R.TopKEngine(B, k=k, mode="fp8")
    for tile in TILES:
        for S in (1, 3):
            v, l = fused(torch, eng, Q, splits=S, tile=tile)
            np.testing.assert_array_equal(l, el, err_msg=f"fp8 K={K} tile={tile} splits={S}")
            np.testing.assert_array_equal(v, ev, err_msg=f"fp8 K={K} tile={tile} splits={S}")
    Q16 = (np.random.default_rng(70 + K).standard_normal((M, K)) * 0.05).astype(np.float16)
    B16 = O.e4m3_table()[B].astype(np.float16)
    C = torch.zeros(N, M, dtype=torch.float32, device="cuda").t()
    R.fp8_gemm_hip(C, torch.from_numpy(Q16).cuda(), torch.from_numpy(B16).cuda(), N)
    torch.cuda.synchronize()
    ev, el = O.topk_rows(C.cpu().numpy(), k)
    eng16 = R.TopKEngine(B, k=k, mode="fp16")
    for tile in TILES:
        for S in (1, 3):
            v, l = fused(torch, eng16, Q16, splits=S, tile=tile)
            np.testing.assert_array_equal(l, el, err_msg=f"fp16 K={K} tile={tile} splits={S}")
            np.testing.assert_array_equal(v, ev, err_msg=f"fp16 K={K} tile={tile} splits={S}")


def test_fused_topk_random_plans_bitexact(torch_gpu):
    """Seeded random shapes, K, k, column splits and query grouping in both
    precisions: the fused top-k equals the oracle's total-order top-k of the
    GPU's own C, value for value and column for column."""
    torch = torch_gpu
    rng = np.random.default_rng(2024)
    for case in range(12):
        M = int(rng.integers(1, 700))
        N = int(rng.integers(1, 6000))
        K = 64 * int(rng.integers(1, 17))
        k = int(rng.choice([1, 5, 20, 32]))
        fp16 = bool(case % 3 == 2)
        tiles = -(-N // 256)
        S = int(rng.integers(1, min(tiles, 9) + 1))
        qg = int(rng.choice([0, 1, 2, 3]))
        tile = int(rng.choice(TILES))
        B = rand_e4m3((N, K), 0.05, 100 + case)
        if fp16:
            Q = (rng.standard_normal((M, K)) * 0.05).astype(np.float16)
            C = torch.zeros(N, M, dtype=torch.float32, device="cuda").t()
            B16 = O.e4m3_table()[B].astype(np.float16)
            R.fp8_gemm_hip(C, torch.from_numpy(Q).cuda(), torch.from_numpy(B16).cuda(), N)
            torch.cuda.synchronize()
            C = C.cpu().numpy()
        else:
            Q = rand_e4m3((M, K), 0.05, 200 + case)
            C = gemm_c(torch, Q, B)
        ev, el = O.topk_rows(C, k)
        eng = R.TopKEngine(B, k=k, mode="fp16" if fp16 else "fp8")
        v, l = fused(torch, eng, Q, splits=S, qg=qg, tile=tile)
        msg = f"case {case}: M={M} N={N} K={K} k={k} S={S} qg={qg} tile={tile} fp16={fp16}"
        np.testing.assert_array_equal(l, el, err_msg=msg)
        np.testing.assert_array_equal(v, ev, err_msg=msg)


def test_fused_rescore_bitexact(torch_gpu):
    torch = torch_gpu
    M, N, K, k, kp = 200, 3000, 2048, 20, 32
    Q8, Q32 = unit_rows_e4m3(M, K, 31)
    B, _ = unit_rows_e4m3(N, K, 32)
    C = gemm_c(torch, Q8, B)
    _, cand = O.topk_rows(C, kp)
    eng = R.TopKEngine(B, k=k, kprime=kp, mode="fp8_rescore")
    for tile in TILES:
        v, l = fused(torch, eng, Q8, Q32, tile=tile)
        for r in range(M):
            ev, el = O.rescore_rank(Q32[r], B, cand[r], k)
            np.testing.assert_array_equal(l[r], el, err_msg=f"tile={tile} row {r}")
            np.testing.assert_array_equal(v[r], ev, err_msg=f"tile={tile} row {r}")
    # rescored values vs the FP64 dot of the fp32 query with the stored columns
    deq = O.e4m3_table()[B].astype(np.float64)
    ref = Q32.astype(np.float64) @ deq.T
    got = ref[np.arange(M)[:, None], l - 1]
    assert np.max(np.abs(got - v)) < 1e-5


def test_rowtopk_merge_dropin(torch_gpu):
    torch = torch_gpu
    M, N, K, w, k = 257, 1000, 512, 300, 20
    Q = rand_e4m3((M, K), 0.05, 41)
    B = rand_e4m3((N, K), 0.05, 42)
    dq, db = torch.from_numpy(Q).cuda(), torch.from_numpy(B).cuda()
    for cdt in (torch.float32, torch.float16):
        Dv = torch.full((k, M), float("-inf"), device="cuda").t()
        Dl = torch.zeros(k, M, dtype=torch.int32, device="cuda").t()
        Call = np.zeros((M, N), dtype=np.float32)
        for lo in range(0, N, w):
            wi = min(N, lo + w) - lo
            buf = torch.zeros(w, M, dtype=cdt, device="cuda").t()  # engine-style parent
            R.fp8_gemm_hip(buf[:, :wi], dq, db, wi, b_col0=lo)
            R.launch_seg_rowtopk_merge(Dv, Dl, buf[:, :wi], lo, k)
            torch.cuda.synchronize()
            Call[:, lo:lo + wi] = buf[:, :wi].float().cpu().numpy()
        ev, el = O.topk_rows(Call, k)
        np.testing.assert_array_equal(Dl.cpu().numpy(), el)
        np.testing.assert_array_equal(Dv.cpu().numpy(), ev)
    # uneven chunks: a first chunk narrower than k leaves the running list
    # part-full, so the next chunk's admission bound comes from its own
    # columns while earlier entries exist; chunks narrower than 2k take none
    Dv = torch.full((k, M), float("-inf"), device="cuda").t()
    Dl = torch.zeros(k, M, dtype=torch.int32, device="cuda").t()
    Call = np.zeros((M, 1000), dtype=np.float32)
    lo = 0
    for wi in (15, 500, 3, 39, 443):
        buf = torch.zeros(wi, M, dtype=torch.float32, device="cuda").t()
        R.fp8_gemm_hip(buf, dq, db, wi, b_col0=lo)
        R.launch_seg_rowtopk_merge(Dv, Dl, buf, lo, k)
        torch.cuda.synchronize()
        Call[:, lo:lo + wi] = buf.cpu().numpy()
        lo += wi
    ev, el = O.topk_rows(Call, k)
    np.testing.assert_array_equal(Dl.cpu().numpy(), el)
    np.testing.assert_array_equal(Dv.cpu().numpy(), ev)


@pytest.mark.parametrize("M,k", [(100, 64), (700, 20), (64, 1)])
def test_rowtopk_merge_segments_and_ties(torch_gpu, M, k):
    """Chunks wide enough that the row top-k splits each chunk into many
    column segments, merged afterwards with the running list: coarse e4m3
    operands give many equal scores, so the column-ascending tie order across
    segments and chunks is exercised; a ragged last chunk included."""
    torch = torch_gpu
    N, K, w = 11_000, 64, 4096
    rng = np.random.default_rng(M + k)
    qc = O.f32_to_e4m3([-0.5, 0.0, 0.5, 1.0])
    bc = O.f32_to_e4m3([0.0, 0.25, 0.5])
    Q = qc[rng.integers(0, 4, (M, K))]
    B = bc[rng.integers(0, 3, (N, K))]
    dq, db = torch.from_numpy(Q).cuda(), torch.from_numpy(B).cuda()
    for cdt in (torch.float32, torch.float16):
        Dv = torch.full((k, M), float("-inf"), device="cuda").t()
        Dl = torch.zeros(k, M, dtype=torch.int32, device="cuda").t()
        Call = np.zeros((M, N), dtype=np.float32)
        buf = torch.zeros(w, M, dtype=cdt, device="cuda").t()
        for lo in range(0, N, w):
            wi = min(N, lo + w) - lo
            R.fp8_gemm_hip(buf[:, :wi], dq, db, wi, b_col0=lo)
            R.launch_seg_rowtopk_merge(Dv, Dl, buf[:, :wi], lo, k)
            torch.cuda.synchronize()
            Call[:, lo:lo + wi] = buf[:, :wi].float().cpu().numpy()
        assert len(np.unique(Call[0])) < N // 10  # heavy ties
        ev, el = O.topk_rows(Call, k)
        np.testing.assert_array_equal(Dl.cpu().numpy(), el)
        np.testing.assert_array_equal(Dv.cpu().numpy(), ev)
    # e4m3 C chunks (random finite codes, both signs)
    codes = rng.integers(0, 0x7f, (M, 2 * w + 5)).astype(np.uint8)
    codes |= (rng.random(codes.shape) < 0.5).astype(np.uint8) << 7
    Dv = torch.full((k, M), float("-inf"), device="cuda").t()
    Dl = torch.zeros(k, M, dtype=torch.int32, device="cuda").t()
    for lo in range(0, codes.shape[1], w):
        ch = codes[:, lo:lo + w]
        R.launch_seg_rowtopk_merge(Dv, Dl, torch.from_numpy(ch.T.copy()).cuda().t(), lo, k)
    torch.cuda.synchronize()
    ev, el = O.topk_rows(O.e4m3_table()[codes], k)
    np.testing.assert_array_equal(Dl.cpu().numpy(), el)
    np.testing.assert_array_equal(Dv.cpu().numpy(), ev)


def test_batch_cut_into_launches_bitexact(torch_gpu):
    """A batch above 64 x CUs rows on an index long enough for the launches'
    splits (4,297 column tiles, 4 splits of 1,075) runs as equal launches of
    at most that many rows.  Its rows equal one launch of the whole batch
    (a pinned plan, which turns the cut off) and separate calls on the
    launches' row ranges, bit for bit, in plain FP8 and with rescoring."""
    torch = torch_gpu
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    K, N, k, kp = 64, 1_100_000, 20, 32
    M = 3 * 64 * cus - 5 * 256 + 13  # three launches of 63 query tiles (4 splits on 256 CUs)
    B = rand_e4m3((N, K), 0.05, 41)
    Q = rand_e4m3((M, K), 0.05, 42)
    rng = np.random.default_rng(43)
    Q32 = (O.e4m3_table()[Q] + rng.standard_normal((M, K)) * 1e-3).astype(np.float32)
    dq, d32 = torch.from_numpy(Q).cuda(), torch.from_numpy(Q32).cuda()
    for mode in ("fp8", "fp8_rescore"):
        eng = R.TopKEngine(B, k=k, kprime=kp, mode=mode)
        rows = eng.launch_rows(M)
        assert rows > 0 and rows % 256 == 0 and rows <= 64 * cus and -(-M // rows) == 3, rows
        q32 = d32 if mode == "fp8_rescore" else None
        v, l = eng.batch_gemm_topk(dq, q32)
        torch.cuda.synchronize()
        v, l = v.cpu().numpy(), l.cpu().numpy()
        # one launch: a pinned plan for M's bucket (the launches' own plan)
        t_, s_, g_ = eng.plan(rows)
        eng.set_plan(M, t_, s_, g_)
        assert eng.launch_rows(M) == 0
        v1, l1 = eng.batch_gemm_topk(dq, q32)
        torch.cuda.synchronize()
        eng.set_plan(M, 0, 0, 0)
        np.testing.assert_array_equal(l1.cpu().numpy(), l, err_msg=mode)
        np.testing.assert_array_equal(v1.cpu().numpy(), v, err_msg=mode)
        # the launches' row ranges as separate calls
        for r0 in range(0, M, rows):
            r1 = min(M, r0 + rows)
            vc, lc = eng.batch_gemm_topk(dq[r0:r1], q32[r0:r1] if q32 is not None else None)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(lc.cpu().numpy(), l[r0:r1], err_msg=f"{mode} rows {r0}")
            np.testing.assert_array_equal(vc.cpu().numpy(), v[r0:r1], err_msg=f"{mode} rows {r0}")
        # the autotuner keeps the cut for this bucket: it reports the launches'
        # plan and stores none, and the rows stay the same
        tile, best, qg, ms = eng.autotune([M])
        assert eng.launch_rows(M) == rows
        assert (int(tile[0]), int(best[0]), int(qg[0])) == eng.plan(rows) and ms[0] > 0
        v2, l2 = eng.batch_gemm_topk(dq, q32)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(l2.cpu().numpy(), l, err_msg=mode + " after autotune")
        np.testing.assert_array_equal(v2.cpu().numpy(), v, err_msg=mode + " after autotune")
        eng.index.close()


def test_persistent_grid_bitexact(torch_gpu):
    """One workgroup per CU walking the (row tile, split) items and one
    workgroup per item give identical top-k, also when each workgroup runs
    many items (16 row tiles x 128 splits = 2048 items)."""
    torch = torch_gpu
    K, k, M, N = 512, 20, 4096, 2 ** 15
    Q = rand_e4m3((M, K), 0.05, 5)
    B = rand_e4m3((N, K), 0.05, 6)
    ev, el = O.topk_rows(gemm_c(torch, Q, B), k)
    eng = R.TopKEngine(B, k=k, mode="fp8")
    ctx = L.ctx(0)
    try:
        for on in (1, 0):
            L.check(R.lib().rm_set_persistent(ctx, on))
            for tile, S in ((256, 1), (256, 16), (256, 128), (64, 8), (128, 32)):
                v, l = fused(torch, eng, Q, splits=S, qg=0, tile=tile)
                msg = f"persistent={on}, tile={tile}, S={S}"
                np.testing.assert_array_equal(l, el, err_msg="locs, " + msg)
                np.testing.assert_array_equal(v, ev, err_msg="vals, " + msg)
    finally:
        L.check(R.lib().rm_set_persistent(ctx, 1))


def test_index_past_2_32_elements(torch_gpu):
    """2.1 M columns x 2048 = 4.3 G elements, past the 32-bit work-item count
    of one dispatch: the FP16 copy is widened over the whole range and the
    last columns are found by both paths (a random column's self product
    beats every cross product by ~10x)."""
    torch = torch_gpu
    N, K = 2_100_000, 2048
    B = torch.empty(N, K, dtype=torch.uint8, device="cuda")
    for r0 in range(0, N, 1 << 18):  # random signed e4m3 codes of magnitude < 0x30
        x = torch.randint(0, 96, (min(1 << 18, N - r0), K), dtype=torch.uint8, device="cuda",
                          generator=torch.Generator(device="cuda").manual_seed(r0))
        B[r0:r0 + x.shape[0]] = (x % 48) | ((x >= 48).to(torch.uint8) << 7)
    rows = torch.tensor([0, 1, 2_097_151, 2_097_152, 2_099_998, 2_099_999], device="cuda")
    want = rows.cpu().numpy() + 1
    Q8 = B[rows].contiguous()
    Q32 = torch.empty(len(rows), K, dtype=torch.float32, device="cuda")
    R.f8_to_f32(Q32, Q8)
    torch.cuda.synchronize()
    for mode, q in (("fp8", Q8), ("fp16", Q32.half())):
        eng = R.TopKEngine(R.ResidentIndex(B, fp16=(mode == "fp16")), k=4, mode=mode)
        v, l = eng.batch_gemm_topk(q)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(l.cpu().numpy()[:, 0], want, err_msg=mode)
        del eng
        torch.cuda.empty_cache()


def test_default_plan_keeps_query_tiles_in_mall(torch_gpu):
    """With no tuned plan, a batch takes the smallest ring tile that holds it
    (64 rows up to 64, 128 up to 128, else 256) and the smallest split count
    (from 'one item per CU' up, doubling) for which the query tiles worked on
    at once, min(R, CUs / S) tiles of that many rows, fit in 128 MB: a
    65,536-read e4m3 batch (2 KB rows, 128 MB) stays unsplit on 256 CUs; the
    same batch in fp16 (4 KB rows, 256 MB) takes two splits, with its query
    tiles grouped CUs / 8 / 2 per XCD; a 4,096-read batch takes CUs / 16.
    While a split would sweep fewer than 32 of the index's column tiles the
    tile halves (down to 64 rows): on this 512-tile index 100 and 200 reads
    run 64-row tiles.  The index is too short for launches to be cut."""
    torch = torch_gpu
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    idx = R.ResidentIndex(np.zeros((512 * 256, 2048), np.uint8), fp16=True)
    ctx = L.ctx(0)
    lib = R.lib()
    try:
        for mode in (L.RM_MODE_FP8, L.RM_MODE_FP8_RESCORE, L.RM_MODE_FP16):
            for M in (65_536, 4_096, 100, 30):
                L.check(lib.rm_plan_set(ctx, M, mode, 0, 0, 0))

        tiles = 512  # column tiles of the index

        def splits(M, bq, row_bytes):
            rt = -(-M // bq)
            S = max(1, min(cus // rt, 256))
            while S < 256 and min(rt, -(-cus // S)) * bq * row_bytes > 128 << 20:
                S = min(2 * S, 256)
            return S

        def expect(M, row_bytes, kp=20):
            bq = 128 if kp > 32 else (64 if M <= 64 else 128 if M <= 128 else 256)
            S = splits(M, bq, row_bytes)
            while kp <= 32 and bq > 64 and -(-tiles // S) < 32:
                bq //= 2
                S = splits(M, bq, row_bytes)
            # query tiles grouped CUs / 8 / S per XCD once there are CUs / 8 of them
            xcd = max(1, cus // 8)
            qg = max(1, xcd // S) if S > 1 and -(-M // bq) >= xcd else 0
            return bq, S, qg

        def plan(M, mode, kp=20):
            t, s_, g = L.C.c_int32(), L.C.c_int32(), L.C.c_int32()
            L.check(lib.rm_plan_get(ctx, idx.handle, M, mode, kp, L.C.byref(t), L.C.byref(s_),
                                    L.C.byref(g)))
            return t.value, s_.value, g.value

        p8, p16 = plan(65_536, L.RM_MODE_FP8), plan(65_536, L.RM_MODE_FP16)
        assert (p8, p16) == (expect(65_536, 2048), expect(65_536, 4096))
        if cus == 256:
            assert (p8[1], p16[1]) == (1, 2) and p16[2] == 16
        eng = R.TopKEngine(idx, k=20, mode="fp8")
        assert eng.launch_rows(65_536) == 0  # 128 column tiles per split: one launch
        for M in (4_096, 200, 100, 30):
            assert plan(M, L.RM_MODE_FP8) == expect(M, 2048), M
        if cus == 256:
            assert plan(200, L.RM_MODE_FP8)[0] == 64 and plan(4_096, L.RM_MODE_FP8)[0] == 256
        assert plan(30, L.RM_MODE_FP8, kp=48) == expect(30, 2048, kp=48)
        # a pinned plan; a tile the list length cannot run falls back
        L.check(lib.rm_plan_set(ctx, 4_096, L.RM_MODE_FP8, 64, 8, 2))
        assert plan(4_096, L.RM_MODE_FP8) == (64, 8, 2)
        assert plan(4_096, L.RM_MODE_FP8, kp=48)[0] == 128
        L.check(lib.rm_plan_set(ctx, 4_096, L.RM_MODE_FP8, 0, 0, 0))
        assert plan(4_096, L.RM_MODE_FP8) == expect(4_096, 2048)
    finally:
        idx.close()


def test_fp16_verbatim_index(torch_gpu):
    """rm_index_create_f16 keeps the fp16 columns verbatim for the FP16 mode
    (index_to_f16) and quantises them to e4m3(fp32(fp16)) on the device for
    the FP8 modes (index_to_f8): the device e4m3 copy equals the host
    quantisation, and the fused FP16 top-k over the verbatim columns equals
    the oracle's top-k of the GPU's own fp16 C, for every tile shape."""
    torch = torch_gpu
    rng = np.random.default_rng(123)
    M, N, K = 300, 1500, 512
    Q = (rng.standard_normal((M, K)) * 0.05).astype(np.float16)
    B16 = (rng.standard_normal((N, K)) * 0.05).astype(np.float16)
    idx = R.ResidentIndex(B16)
    assert idx.verbatim_f16 and idx.fp16
    deq = torch.empty(N * K, dtype=torch.float32, device="cuda")
    L.check(R.lib().rm_dequantize_e4m3(L.ctx(0), L.ptr(deq), R.lib().rm_index_e4m3(idx.handle),
                                       N * K, None))
    torch.cuda.synchronize()
    exp8 = O.f32_to_e4m3(B16.astype(np.float32).ravel())
    np.testing.assert_array_equal(deq.cpu().numpy(), O.e4m3_table()[exp8])
    C = torch.zeros(N, M, dtype=torch.float32, device="cuda").t()
    R.fp8_gemm_hip(C, torch.from_numpy(Q).cuda(), torch.from_numpy(B16).cuda(), N)
    torch.cuda.synchronize()
    ev, el = O.topk_rows(C.cpu().numpy(), 20)
    eng = R.TopKEngine(idx, k=20, mode="fp16")
    for tile in TILES:
        v, l = fused(torch, eng, Q, splits=2, tile=tile)
        np.testing.assert_array_equal(l, el, err_msg=f"tile={tile}")
        np.testing.assert_array_equal(v, ev, err_msg=f"tile={tile}")
    # the FP8 modes run on the quantised copy
    eng8 = R.TopKEngine(idx, k=20, mode="fp8")
    eng8h = R.TopKEngine(exp8.reshape(N, K), k=20, mode="fp8")
    Q8 = O.f32_to_e4m3(Q.astype(np.float32).ravel()).reshape(M, K)
    a, b = fused(torch, eng8, Q8), fused(torch, eng8h, Q8)
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_array_equal(a[0], b[0])


def test_rescore_on_verbatim_fp16_columns(torch_gpu):
    """FP8 + rescore on an fp16 index with rm_index_set_rescore(RM_F16): the
    k' candidates come from the FP8 GEMM over the e4m3 copy, and the final
    scores are the fixed-order FP32 dots against the verbatim fp16 columns
    (the oracle's rescore_rank over float16 columns), bit for bit, for every
    tile shape; switching back to e4m3 restores the e4m3 rescoring."""
    torch = torch_gpu
    M, N, K, k, kp = 160, 2500, 2048, 20, 32
    rng = np.random.default_rng(77)
    x = rng.standard_normal((N, K))
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    B16 = x.astype(np.float16)
    idx = R.ResidentIndex(B16)
    assert idx.rescore == "e4m3"
    B8 = O.f32_to_e4m3(B16.astype(np.float32).ravel()).reshape(N, K)
    Q8, Q32 = unit_rows_e4m3(M, K, 78)
    _, cand = O.topk_rows(gemm_c(torch, Q8, B8), kp)
    eng = R.TopKEngine(idx, k=k, kprime=kp, mode="fp8_rescore")
    idx.set_rescore("fp16")
    assert idx.rescore == "fp16"
    for tile in TILES:
        v, l = fused(torch, eng, Q8, Q32, tile=tile)
        for r in range(M):
            ev, el = O.rescore_rank(Q32[r], B16, cand[r], k)
            np.testing.assert_array_equal(l[r], el, err_msg=f"tile={tile} row {r}")
            np.testing.assert_array_equal(v[r], ev, err_msg=f"tile={tile} row {r}")
    ref = Q32.astype(np.float64) @ B16.astype(np.float64).T
    assert np.max(np.abs(ref[np.arange(M)[:, None], l - 1] - v)) < 1e-5
    idx.set_rescore("e4m3")
    v, l = fused(torch, eng, Q8, Q32)
    for r in range(0, M, 7):
        ev, el = O.rescore_rank(Q32[r], B8, cand[r], k)
        np.testing.assert_array_equal(l[r], el)
        np.testing.assert_array_equal(v[r], ev)
    # an e4m3-only index has no fp16 columns to rescore against
    with pytest.raises(L.RotorMapError):
        R.ResidentIndex(B8).set_rescore("fp16")


@pytest.mark.parametrize("K,kp", [(512, 32), (2048, 48), (320, 20)])
def test_rescore_general_path_bitexact(torch_gpu, K, kp):
    """The rescoring kernel's general path (k' > 32, or K not a multiple of
    1,024: per-candidate gathers through LDS) against the oracle's fixed
    order, on e4m3 columns and on an fp16 index's verbatim columns."""
    torch = torch_gpu
    M, N, k = 96, 1800, 20
    Q8, Q32 = unit_rows_e4m3(M, K, 90 + K)
    rng = np.random.default_rng(91 + kp)
    x = rng.standard_normal((N, K))
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    B16 = x.astype(np.float16)
    B8 = O.f32_to_e4m3(B16.astype(np.float32).ravel()).reshape(N, K)
    _, cand = O.topk_rows(gemm_c(torch, Q8, B8), kp)
    for cols in ("e4m3", "fp16"):
        idx = R.ResidentIndex(B16)
        idx.set_rescore(cols)
        eng = R.TopKEngine(idx, k=k, kprime=kp, mode="fp8_rescore")
        v, l = fused(torch, eng, Q8, Q32)
        Bc = B16 if cols == "fp16" else B8
        for r in range(M):
            ev, el = O.rescore_rank(Q32[r], Bc, cand[r], k)
            np.testing.assert_array_equal(l[r], el, err_msg=f"{cols} row {r}")
            np.testing.assert_array_equal(v[r], ev, err_msg=f"{cols} row {r}")
        idx.close()
