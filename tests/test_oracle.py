"""Pins of the CPU oracle (no GPU): every oracle function is checked against
something it did not compute itself -- IEEE / OCP format definitions,
closed forms of the RoPE histogram, the reference's committed sweep anchors,
brute force and numpy."""
import json
import math
import os

import numpy as np
import pytest

import oracle.oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

# The fixed-point encoder (the GPU kernel's specification) normalises in fp32:
# against the FP64 golden its unit-energy outputs agree within 4 ulp of 1.0
# elementwise (2^-22; about 1.7 ulp measured), its norms within 2^-21
# relative, beyond the fixed-point table's own bound of nw * 2^-30 per bin
# component (before normalisation).
FP32_ATOL = 2.0 ** -22
FP32_NORM_RTOL = 2.0 ** -21


# ---- storage formats ------------------------------------------------------

def test_f16_rounding_matches_ieee_numpy():
    # every binary16 value, the midpoints between neighbours, and random floats
    h = np.arange(65536, dtype=np.uint16).view(np.float16)
    fin = np.isfinite(h)
    x = h[fin].astype(np.float32)
    pos = np.sort(np.unique(np.abs(h[fin].astype(np.float64))))
    mids = ((pos[:-1] + pos[1:]) / 2).astype(np.float32)
    rng = np.random.default_rng(0)
    rnd = (rng.standard_normal(20000) * np.exp(rng.uniform(-20, 12, 20000))).astype(np.float32)
    for arr in (x, mids, -mids, rnd, np.array([65504, 65519.99, 65520, 1e-8, 6e-8], np.float32)):
        got = O.f32_to_f16_bits(arr)
        with np.errstate(over="ignore"):  # values past 65520 round to inf on purpose
            exp = arr.astype(np.float16).view(np.uint16)
        np.testing.assert_array_equal(got, exp)


def e4m3_value_spec(code):
    """OCP E4M3FN value of a code, from the format definition."""
    s = -1.0 if code & 0x80 else 1.0
    e, m = (code >> 3) & 15, code & 7
    if e == 15 and m == 7:
        return math.nan
    if e == 0:
        return s * m * 2.0 ** -9
    return s * (1 + m / 8) * 2.0 ** (e - 7)


def test_e4m3_decode_table():
    t = O.e4m3_table()
    for c in range(256):
        v = e4m3_value_spec(c)
        if math.isnan(v):
            assert math.isnan(t[c])
        else:
            assert t[c] == v, c
    assert t[0x7E] == 448.0 and t[0x08] == 2.0 ** -6 and t[0x01] == 2.0 ** -9


def test_e4m3_encode_roundtrip_ties_and_saturation():
    t = O.e4m3_table()
    codes = [c for c in range(256) if not math.isnan(t[c]) and c != 0x80]
    np.testing.assert_array_equal(O.f32_to_e4m3(t[codes]), np.array(codes, dtype=np.uint8))
    # midpoints round to the even code; just above / below go to the neighbour
    pos = [c for c in range(0, 0x7F)]
    for a, b in zip(pos[:-1], pos[1:]):
        mid = np.float32((np.float64(t[a]) + np.float64(t[b])) / 2)
        even = a if a % 2 == 0 else b
        assert O.f32_to_e4m3(mid)[0] == even
        assert O.f32_to_e4m3(np.nextafter(mid, np.float32(0)))[0] == a
        assert O.f32_to_e4m3(np.nextafter(mid, np.float32(1e9)))[0] == b
    # satfinite: everything beyond 448 clamps, sign kept
    big = np.array([448.0, 449.0, 464.0, 480.0, 1e30, -1e30], np.float32)
    np.testing.assert_array_equal(O.f32_to_e4m3(big), [0x7E, 0x7E, 0x7E, 0x7E, 0x7E, 0xFE])


# ---- encoder closed forms ---------------------------------------------------

def geometric(theta, nw):
    """sum_{p<nw} exp(i theta p) = (1 - e^{i theta nw}) / (1 - e^{i theta})."""
    z = np.exp(1j * theta)
    return (1 - z ** nw) / (1 - z)


@pytest.mark.parametrize("s,m,c", [(8, 4, 4), (5, 1, 5), (5, 1, 4), (16, 2, 4)])
@pytest.mark.parametrize("k", [17, 2000, 2001])
def test_all_A_fragment_single_bin(s, m, c, k):
    if s > k:
        pytest.skip()
    codes = np.zeros(k, np.uint8)
    nw = k - s + 1
    D = m * 4 ** c
    for enc in (O.ref_rope_frag_real, O.rope_fixed):
        out, nr = enc(codes, k, s, m, c, 0)
        h = np.array([geometric((1.0 if m == 1 else 2 * i / (m - 1) + 1) * 2 * np.pi / k, nw)
                      for i in range(m)])
        nrm = np.sqrt(np.sum(np.abs(h) ** 2)) + 1e-16
        exp = np.zeros(2 * D)
        exp[:m] = h.real / nrm          # key 0 -> bins 0..m-1
        exp[D:D + m] = h.imag / nrm
        # fixed point: every phasor is rounded to 2^-30, so the bin sums are
        # within nw * 2^-30 per component of the exact series
        fixed = enc is O.rope_fixed
        bound = 2 * nw * 2.0 ** -30 if fixed else 1e-9
        atol = max(1e-9, bound / nrm) + (FP32_ATOL if fixed else 0.0)
        np.testing.assert_allclose(out, exp, rtol=0, atol=atol)
        assert abs(nr[0] - nrm) <= max(1e-9 * nrm, bound) + (FP32_NORM_RTOL * nrm if fixed else 0)


def test_all_G_fragment():
    k = 2000
    g = np.full(k, 2, np.uint8)
    # c == s: the only key is the zeroed G bin -> zero vector (norm 1e-16)
    for enc in (O.ref_rope_frag_real, O.rope_fixed):
        out, nr = enc(g, k, 5, 1, 5, 0)
        assert np.all(out == 0.0)
    # c < s (s8m4c4): all-G key is (0xAA ^ ((0xAA*0x9E3779B1) & 0xFF)) = 0x20,
    # not the zeroed 0xAA bin -- the reference keeps this quirk
    key = 0xAA ^ ((0xAA * 0x9E3779B1) & 0xFF)
    assert key == 0x20
    for enc in (O.ref_rope_frag_real, O.rope_fixed):
        out, nr = enc(g, k, 8, 4, 4, 0)
        nz = np.nonzero(out)[0]
        D = 4 * 256
        assert set(nz.tolist()) <= set(range(key * 4, key * 4 + 4)) | set(range(D + key * 4, D + key * 4 + 4))
        assert abs(np.sum(out ** 2) - 1) < (1e-6 if enc is O.rope_fixed else 1e-12)


@pytest.mark.parametrize("s,m,c", [(8, 4, 4), (5, 1, 5)])
def test_counts_sum_and_unit_energy(s, m, c):
    rng = np.random.default_rng(1)
    k = 20000
    codes = rng.integers(0, 4, k).astype(np.uint8)
    cnt = O.smer_counts(codes, k, s, c)
    assert cnt.sum() == k - s + 1
    for enc in (O.ref_rope_frag_real, O.rope_fixed):
        out, _ = enc(codes, k, s, m, c, 0)
        assert abs(np.sum(out ** 2) - 1.0) < (1e-6 if enc is O.rope_fixed else 1e-12)


def test_histogram_phase_sum_closed_form():
    """Before G zeroing, every copy's bins sum to the geometric series; after,
    the sum is that minus the G bin's share.  Checked through the fixed-point
    oracle's integer bin sums (exact integers of the phasor table)."""
    rng = np.random.default_rng(2)
    k, s, m, c = 3001, 8, 4, 4
    codes = rng.integers(0, 4, k).astype(np.uint8)
    codes[100:140] = 2  # make sure the G key occurs
    _, _, acc = O.rope_fixed(codes, k, s, m, c, 0, want_acc=True)
    tab = O.phase_table(k, s, m)
    nw = k - s + 1
    for i in range(m):
        tot = acc[i::m].sum(axis=0)
        f = 1.0 if m == 1 else 2 * i / (m - 1) + 1
        g = geometric(f * 2 * np.pi / k, nw) * 2 ** 30
        gkey = 2 * sum(4 ** j for j in range(c))
        # add back the zeroed G bin: the windows whose key is gkey
        keys = []
        for p in range(nw):
            smer = sum(int(codes[p + j]) << (2 * j) for j in range(s))
            lower, upper = smer & 255, smer >> 8
            keys.append(lower ^ ((upper * 0x9E3779B1) & 255))
        gsum = tab[i][np.array(keys) == gkey].sum(axis=0)
        assert abs(tot[0] + gsum[0] - g.real) < nw and abs(tot[1] + gsum[1] - g.imag) < nw


def test_shift_and_mutation_sweep_anchors():
    """RotorMap/experiments/out/rope_shift_mutation_sweep.csv:2: a 500-base
    shift of a random k = 20,000 fragment (s5m1c5) rotates the inner product
    by cos(2*pi*500/k): Re/|.| = 0.98745 / 0.99979 = 0.98765.  (The row's
    |.| itself depends on the one random draw and is not pinned.)
    rope_mutation_sweep.csv:2-3: the self product is 1 and ten edits
    (err 0.0005 under the sub/ins/del model) cost ~0.26 % (0.99736)."""
    rng = np.random.default_rng(42)
    k, d = 20000, 500
    frag = rng.integers(0, 4, k).astype(np.uint8)
    shifted = np.concatenate([frag[d:], rng.integers(0, 4, d).astype(np.uint8)])
    for enc in (O.ref_rope_frag_real, O.rope_fixed):
        a, _ = enc(frag, k, 5, 1, 5, 0)
        b, _ = enc(shifted, k, 5, 1, 5, 0)
        D = len(a) // 2
        za, zb = a[:D] + 1j * a[D:], b[:D] + 1j * b[D:]
        ip = np.sum(np.conj(za) * zb)
        assert abs(np.dot(a, b) - ip.real) < 1e-12
        assert abs(ip.real / abs(ip) - 0.98745 / 0.99979) < 2e-3
        assert abs(ip.real / abs(ip) - math.cos(2 * math.pi * d / k)) < 2e-3
        assert abs(np.dot(a, a) - 1.0) < (1e-6 if enc is O.rope_fixed else 1e-12)
    # ten edits on the pristine fragment: sub/ins/del model of the library
    import rotormap_amd as R
    from rotormap_amd import _lib as L
    a, _ = O.ref_rope_frag_real(frag, k, 5, 1, 5, 0)
    ips = []
    for seed in range(8):
        w = R.pack_codes(frag, guard=False).reshape(1, -1).copy()
        err, ids = np.array([0.0005]), np.array([0], np.int64)
        L.check(R.lib().rm_mutate_words(L.ptr(w), 1, k, L.ptr(err), seed, L.ptr(ids), 1))
        b, _ = O.ref_rope_frag_real(O.frag_codes(w[0], k), k, 5, 1, 5, 0)
        ips.append(np.dot(a, b))
    assert abs(np.mean(ips) - 0.99736) < 1.5e-3


@pytest.mark.parametrize("s,m,c", [(8, 4, 4), (5, 1, 4), (16, 2, 4), (8, 1, 6), (5, 1, 5)])
@pytest.mark.parametrize("normalize", [0, 1, 2, 3])
def test_fixed_point_within_tolerance_of_golden(s, m, c, normalize):
    rng = np.random.default_rng(s * 10 + normalize)
    for k in (17, 2001):
        if s > k:
            continue
        codes = rng.integers(0, 4, k).astype(np.uint8)
        a, na = O.ref_rope_frag_real(codes, k, s, m, c, normalize)
        b, nb = O.rope_fixed(codes, k, s, m, c, normalize)
        np.testing.assert_allclose(b, a, rtol=0, atol=FP32_ATOL)
        rows = [0] if normalize == 0 else range(m)
        for i in rows:
            assert abs(nb[i] - na[i]) <= FP32_NORM_RTOL * max(1e-6, na[i])


def test_golden_encoder_vs_bruteforce_python():
    """The FP64 golden against a direct Python transcription of the complex
    histogram (tiny k, all normalize modes)."""
    rng = np.random.default_rng(3)
    for (k, s, m, c) in [(17, 5, 1, 4), (23, 8, 4, 2), (40, 16, 2, 4)]:
        codes = rng.integers(0, 4, k).astype(np.uint8)
        D = m * 4 ** c
        hist = np.zeros(D, complex)
        for p in range(k - s + 1):
            smer = sum(int(codes[p + j]) << (2 * j) for j in range(s))
            mask = 4 ** c - 1
            key = (smer & mask) ^ (((smer >> (2 * c)) * 0x9E3779B1 % 2 ** 32) & mask)
            for i in range(m):
                f = 1.0 if m == 1 else 2 * i / (m - 1) + 1
                hist[key * m + i] += np.exp(1j * f * 2 * np.pi / k * p)
        g = 2 * sum(4 ** j for j in range(c))
        hist[g * m:g * m + m] = 0
        nrm = np.sqrt(np.sum(np.abs(hist) ** 2)) + 1e-16
        h = hist / nrm
        out, _ = O.ref_rope_frag_real(codes, k, s, m, c, 0)
        np.testing.assert_allclose(out, np.concatenate([h.real, h.imag]), atol=1e-12)


def test_golden_fixture_file():
    """tests/golden/encoder_closed_forms.json: expected outputs derived from
    the geometric series alone (make_golden.py), checked against both
    oracle encoders."""
    with open(os.path.join(GOLDEN, "encoder_closed_forms.json")) as f:
        cases = json.load(f)
    assert len(cases) >= 4
    for cs in cases:
        k, s, m, c = cs["k"], cs["s"], cs["m"], cs["c"]
        codes = np.array(cs["codes"], np.uint8) if "codes" in cs else np.full(k, cs["fill"], np.uint8)
        for enc in (O.ref_rope_frag_real, O.rope_fixed):
            out, nr = enc(codes, k, s, m, c, 0)
            tol = 1e-9 if enc is O.ref_rope_frag_real else \
                max(1e-9, 2 * k * 2.0 ** -30 / nr[0]) + FP32_ATOL
            for idx, val in cs["nonzero"]:
                assert abs(out[idx] - val) < tol, (cs["name"], idx)
            mask = np.ones(len(out), bool)
            mask[[i for i, _ in cs["nonzero"]]] = False
            assert np.all(np.abs(out[mask]) < 1e-12)


# ---- reverse complement, top-k, GEMM, rescoring -------------------------------

def test_revcomp_involution_and_packing():
    rng = np.random.default_rng(4)
    for k in (1, 15, 16, 17, 2001):
        codes = rng.integers(0, 4, k).astype(np.uint8)
        rc = O.revcomp_codes(codes)
        np.testing.assert_array_equal(rc, 3 - codes[::-1])
        np.testing.assert_array_equal(O.revcomp_codes(rc), codes)
        w = O.forward_pack(codes)
        np.testing.assert_array_equal(O.frag_codes(w, k), codes)
        if k % 16:
            assert w[-1] >> (2 * (k % 16)) == 0  # zero tail


def test_topk_total_order_vs_lexsort():
    rng = np.random.default_rng(5)
    C = rng.integers(-5, 5, size=(50, 300)).astype(np.float32)  # many ties
    for k in (1, 20, 64):
        v, l = O.topk_rows(C, k)
        for r in range(50):
            order = np.lexsort((np.arange(300), -C[r]))[:k]
            np.testing.assert_array_equal(l[r], order + 1)
            np.testing.assert_array_equal(v[r], C[r][order])
    v, l = O.topk_rows(C[:, :5], 8)  # fewer columns than k: empty slots
    assert np.all(l[:, 5:] == 0) and np.all(np.isneginf(v[:, 5:]))


def test_gemm_f64_vs_numpy():
    rng = np.random.default_rng(6)
    Q = rng.integers(0, 256, (7, 128)).astype(np.uint8)
    B = rng.integers(0, 256, (9, 128)).astype(np.uint8)
    Q[(Q & 0x7F) == 0x7F] = 0
    B[(B & 0x7F) == 0x7F] = 0
    t = np.array([e4m3_value_spec(c) if not math.isnan(e4m3_value_spec(c)) else 0 for c in range(256)])
    np.testing.assert_allclose(O.gemm_e4m3_f64(Q, B), t[Q] @ t[B].T, rtol=1e-12, atol=1e-12)


def test_rescore_dot_fixed_order():
    rng = np.random.default_rng(7)
    K = 2048
    q = rng.standard_normal(K).astype(np.float32)
    b = O.f32_to_e4m3(rng.standard_normal(K).astype(np.float32) * 0.1)
    d = O.e4m3_table()[b]
    # the lane / butterfly order restated with numpy float32 scalars
    C_ = K // 64
    s = np.zeros(64, np.float32)
    for l in range(64):
        acc = np.float32(0)
        for e in range(l * C_, (l + 1) * C_):
            acc = np.float32(acc + np.float32(q[e] * d[e]))
        s[l] = acc
    for off in (32, 16, 8, 4, 2, 1):
        s = np.array([np.float32(s[l] + s[l ^ off]) for l in range(64)], np.float32)
    got = O.rescore_dot(q, b)
    assert got == s[0]
    assert abs(got - float(np.dot(q.astype(np.float64), d.astype(np.float64)))) < 1e-4
    # fp16 columns: the same order over binary16 elements; the exact fp16
    # widening of an e4m3 column gives the e4m3 column's bits
    assert O.rescore_dot(q, d.astype(np.float16)) == got
    h = (rng.standard_normal(K) * 0.03).astype(np.float16)
    s = np.zeros(64, np.float32)
    for l in range(64):
        acc = np.float32(0)
        for e in range(l * C_, (l + 1) * C_):
            acc = np.float32(acc + np.float32(q[e] * np.float32(h[e])))
        s[l] = acc
    for off in (32, 16, 8, 4, 2, 1):
        s = np.array([np.float32(s[l] + s[l ^ off]) for l in range(64)], np.float32)
    assert O.rescore_dot(q, h) == s[0]


def test_rescore_rank_fp16_columns_vs_numpy():
    """Ranking of rescored fp16 columns: (score desc, column asc) with
    duplicate columns forcing ties and an empty candidate slot."""
    rng = np.random.default_rng(17)
    K, N = 128, 40
    B = (rng.standard_normal((N, K)) * 0.05).astype(np.float16)
    B[7] = B[3]  # equal scores: column 4 ranks before column 8
    q = rng.standard_normal(K).astype(np.float32)
    cand = np.array([8, 4, 0, 12, 1, 30, 22, 40], np.int32)
    v, l = O.rescore_rank(q, B, cand, 5)
    live = cand[cand > 0]
    sc = np.array([O.rescore_dot(q, B[c - 1]) for c in live], np.float32)
    order = np.lexsort((live, -sc))[:5]
    np.testing.assert_array_equal(l, live[order])
    np.testing.assert_array_equal(v, sc[order])
    assert list(l).index(4) < list(l).index(8) if 8 in l and 4 in l else True


# ---- the CPU golden mapping path and the modes' score models --------------

def _tiny_map_case(seed=5, k=300, s=4, m=1, c=4, n=6, N=90):
    """Random reads, a random e4m3 database plus the reads' own encodings as
    columns (so each read has a clear best hit)."""
    rng = np.random.default_rng(seed)
    W = (k + 15) // 16
    codes = rng.integers(0, 4, (n, k)).astype(np.uint8)
    words = np.stack([O.forward_pack(cd) for cd in codes])
    K = 2 * m * 4 ** c
    hits = [O.store_chain(O.ref_rope_frag_real(cd, k, s, m, c, 0)[0], "e4m3") for cd in codes]
    B = O.f32_to_e4m3((rng.standard_normal(N * K) * 0.04).astype(np.float32)).reshape(N, K)
    B[rng.choice(N, n, replace=False)] = np.stack(hits)
    return words, codes, B, (k, s, m, c)


def _numpy_topk(scores, ktop):
    order = np.lexsort((np.arange(len(scores)), -scores.astype(np.float32)))[:ktop]
    return scores[order], order + 1


def test_cpu_map_vs_numpy_bruteforce():
    """rmo_cpu_map (the CPU baseline) = golden FP64 encoder -> e4m3 chain ->
    FP64 dots over the dequantised columns -> total-order top-k, against the
    same model written with numpy: the golden encoder of the oracle's
    python-checked restatement, numpy's e4m3 table and float64 matmul."""
    words, codes, B, (k, s, m, c) = _tiny_map_case()
    vals, locs = O.cpu_map(words, k, s, m, c, B, 10, threads=2)
    deq = O.e4m3_table()[B].astype(np.float64)
    for r in range(len(codes)):
        h, _ = O.ref_rope_frag_real(codes[r], k, s, m, c, 0)
        q = O.e4m3_table()[O.store_chain(h, "e4m3")].astype(np.float64)
        ev, el = _numpy_topk(deq @ q, 10)
        np.testing.assert_array_equal(locs[r], el)
        np.testing.assert_allclose(vals[r], ev.astype(np.float32), rtol=0, atol=1e-6)
        # the read's own column is its best hit (self product = unit energy
        # up to the e4m3 rounding)
        assert abs(vals[r, 0] - 1.0) < 0.05


def test_golden_query_chains_and_dots():
    words, codes, B, (k, s, m, c) = _tiny_map_case(seed=8)
    h, _ = O.ref_rope_frag_real(codes[0], k, s, m, c, 0)
    np.testing.assert_array_equal(O.golden_query(words[0], k, s, m, c, "f64"), h)
    np.testing.assert_array_equal(O.golden_query(words[0], k, s, m, c, "f32"),
                                  h.astype(np.float32).astype(np.float64))
    np.testing.assert_array_equal(O.golden_query(words[0], k, s, m, c, "f16"),
                                  h.astype(np.float32).astype(np.float16).astype(np.float64))
    q8 = O.golden_query(words[0], k, s, m, c, "e4m3")
    np.testing.assert_array_equal(q8, O.e4m3_table()[O.store_chain(h, "e4m3")].astype(np.float64))
    cols = np.array([1, 0, 5, len(B)], np.int32)
    got = O.dots_f64(q8, B, cols)
    deq = O.e4m3_table()[B].astype(np.float64)
    np.testing.assert_allclose(got[[0, 2, 3]], deq[[0, 4, len(B) - 1]] @ q8, rtol=0, atol=1e-12)
    assert np.isnan(got[1])
    B16 = deq.astype(np.float16)
    np.testing.assert_allclose(O.dots_f64(q8, B16, cols[[0, 2]]),
                               B16.astype(np.float64)[[0, 4]] @ q8, rtol=0, atol=1e-12)


def test_scan_topk_f64_vs_numpy():
    words, codes, B, (k, s, m, c) = _tiny_map_case(seed=9)
    Q = np.stack([O.golden_query(w, k, s, m, c, "f32") for w in words])
    deq = O.e4m3_table()[B].astype(np.float64)
    for Bx in (B, deq.astype(np.float16)):
        v, l = O.scan_topk_f64(Q, Bx, 7, threads=2)
        full = Q @ deq.T  # the e4m3 -> fp16 widening is exact
        for r in range(len(Q)):
            ev, el = _numpy_topk(full[r], 7)
            np.testing.assert_array_equal(l[r], el)
            np.testing.assert_allclose(v[r], ev, rtol=0, atol=1e-12)
