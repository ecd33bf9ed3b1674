"""ctypes wrapper of the CPU oracle (oracle/_build/librm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker; never by the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "_build", "librm_oracle.so")
_o = None


def build():
    r = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return SO


def lib():
    global _o
    if _o is None:
        if not os.path.exists(SO):
            build()
        o = C.CDLL(SO)
        vp, i32, i64 = C.c_void_p, C.c_int, C.c_int64
        sig = {
            "rmo_frag_codes": (None, [vp, i32, vp]),
            "rmo_forward_pack": (None, [vp, i32, vp]),
            "rmo_byte_code": (C.c_uint8, [C.c_uint8]),
            "rmo_revcomp_codes": (None, [vp, i32, vp]),
            "rmo_smer_counts": (None, [vp, i32, i32, i32, vp]),
            "rmo_ref_rope_frag_real": (None, [vp, i32, i32, i32, i32, i32, vp, vp]),
            "rmo_phase_table": (None, [i32, i32, i32, vp]),
            "rmo_rope_fixed": (None, [vp, i32, i32, i32, i32, i32, vp, vp, vp, vp]),
            "rmo_f32_to_f16": (C.c_uint16, [C.c_float]),
            "rmo_f16_to_f32": (C.c_float, [C.c_uint16]),
            "rmo_f32_to_e4m3": (C.c_uint8, [C.c_float]),
            "rmo_e4m3_to_f32": (C.c_float, [C.c_uint8]),
            "rmo_f64_to_e4m3_chain": (C.c_uint8, [C.c_double]),
            "rmo_gemm_e4m3_f64": (None, [vp, vp, i64, i64, i64, vp]),
            "rmo_gemm_f16_f64": (None, [vp, vp, i64, i64, i64, vp]),
            "rmo_topk_rows_f32": (None, [vp, i64, i64, i64, i32, vp, vp]),
            "rmo_rescore_dot": (C.c_float, [vp, vp, i32]),
            "rmo_rescore_rank": (None, [vp, vp, i32, vp, i32, i32, vp, vp]),
            "rmo_rescore_dot_t": (C.c_float, [vp, vp, i32, i32]),
            "rmo_rescore_rank_t": (None, [vp, vp, i32, i32, vp, i32, i32, vp, vp]),
            "rmo_fasta_reads": (i64, [vp, i64, i32, vp, i64, vp, vp]),
            "rmo_cpu_map": (None, [vp, i64, i32, i32, i32, i32, vp, i64, i32, i32, vp, vp, i32]),
            "rmo_golden_query": (None, [vp, i32, i32, i32, i32, i32, vp]),
            "rmo_dots_f64": (None, [vp, vp, i32, i64, vp, i32, vp]),
            "rmo_scan_topk_f64": (None, [vp, i64, vp, i32, i64, i64, i32, vp, vp, i32]),
        }
        for n, (r, a) in sig.items():
            f = getattr(o, n)
            f.restype = r
            f.argtypes = a
        _o = o
    return _o


def _p(a):
    return C.c_void_p(a.ctypes.data)


def frag_codes(words, k):
    w = np.ascontiguousarray(words, dtype=np.uint32)
    out = np.empty(k, dtype=np.uint8)
    lib().rmo_frag_codes(_p(w), k, _p(out))
    return out


def forward_pack(codes):
    c = np.ascontiguousarray(codes, dtype=np.uint8)
    out = np.empty((len(c) + 15) // 16, dtype=np.uint32)
    lib().rmo_forward_pack(_p(c), len(c), _p(out))
    return out


def revcomp_codes(codes):
    c = np.ascontiguousarray(codes, dtype=np.uint8)
    out = np.empty_like(c)
    lib().rmo_revcomp_codes(_p(c), len(c), _p(out))
    return out


def smer_counts(codes, k, s, c):
    cd = np.ascontiguousarray(codes, dtype=np.uint8)
    out = np.empty(4 ** c, dtype=np.int32)
    lib().rmo_smer_counts(_p(cd), k, s, c, _p(out))
    return out


def ref_rope_frag_real(codes, k, s, m, c, normalize=0):
    """FP64 golden encoder: (out (2D,), norms (m,))."""
    cd = np.ascontiguousarray(codes, dtype=np.uint8)
    D = m * 4 ** c
    out = np.empty(2 * D)
    nr = np.empty(m)
    lib().rmo_ref_rope_frag_real(_p(cd), k, s, m, c, normalize, _p(out), _p(nr))
    return out, nr


_tabs = {}


def phase_table(k, s, m):
    key = (k, s, m)
    if key not in _tabs:
        t = np.empty((m, k - s + 1, 2), dtype=np.int32)
        lib().rmo_phase_table(k, s, m, _p(t))
        _tabs[key] = t
    return _tabs[key]


def rope_fixed(codes, k, s, m, c, normalize=0, want_acc=False):
    """Fixed-point spec encoder: (out f64 (2D,), norms f64 (m,) [, acc int64 (D,2)])."""
    cd = np.ascontiguousarray(codes, dtype=np.uint8)
    D = m * 4 ** c
    out = np.empty(2 * D)
    nr = np.empty(m)
    acc = np.empty((D, 2), dtype=np.int64) if want_acc else None
    lib().rmo_rope_fixed(_p(cd), k, s, m, c, normalize, _p(phase_table(k, s, m)), _p(out), _p(nr),
                         _p(acc) if want_acc else None)
    return (out, nr, acc) if want_acc else (out, nr)


def f32_to_f16_bits(x):
    x = np.asarray(x, dtype=np.float32).ravel()
    return np.array([lib().rmo_f32_to_f16(float(v)) for v in x], dtype=np.uint16)


def f32_to_e4m3(x):
    x = np.asarray(x, dtype=np.float32).ravel()
    return np.array([lib().rmo_f32_to_e4m3(float(v)) for v in x], dtype=np.uint8)


def e4m3_table():
    return np.array([lib().rmo_e4m3_to_f32(i) for i in range(256)], dtype=np.float32)


def store_chain(v, dtype):
    """The encoder's storage rounding of FP64 values: 'f32' | 'f16' | 'e4m3'
    (f16 = binary16 of the fp32 value; e4m3 = e4m3(fp16(fp32)))."""
    f = np.asarray(v, dtype=np.float64).astype(np.float32)
    if dtype == "f32":
        return f
    h = f.astype(np.float16)  # numpy: correctly rounded fp32 -> binary16 (RNE)
    if dtype == "f16":
        return h
    return f32_to_e4m3(h.astype(np.float32))


def gemm_e4m3_f64(Q, B):
    Q = np.ascontiguousarray(Q, dtype=np.uint8)
    B = np.ascontiguousarray(B, dtype=np.uint8)
    M, K = Q.shape
    N = B.shape[0]
    C_ = np.empty((M, N))
    lib().rmo_gemm_e4m3_f64(_p(Q), _p(B), M, N, K, _p(C_))
    return C_


def gemm_f16_f64(Q, B):
    Q = np.ascontiguousarray(Q, dtype=np.float16).view(np.uint16)
    B = np.ascontiguousarray(B, dtype=np.float16).view(np.uint16)
    M, K = Q.shape
    N = B.shape[0]
    C_ = np.empty((M, N))
    lib().rmo_gemm_f16_f64(_p(Q), _p(B), M, N, K, _p(C_))
    return C_


def topk_rows(Cm, k):
    """Total-order top-k per row of a row-major fp32 matrix: (vals, locs 1-based)."""
    Cm = np.ascontiguousarray(Cm, dtype=np.float32)
    M, N = Cm.shape
    v = np.empty((M, k), dtype=np.float32)
    l = np.empty((M, k), dtype=np.int32)
    lib().rmo_topk_rows_f32(_p(Cm), M, N, N, k, _p(v), _p(l))
    return v, l


def rescore_rank(q32, B, cand, ktop):
    """Fixed-order FP32 rescoring of 1-based candidate columns -> (vals, locs).
    B: (N, K) uint8 e4m3 columns, or float16 columns (an fp16 index's verbatim
    columns)."""
    q = np.ascontiguousarray(q32, dtype=np.float32)
    elem = 1 if B.dtype == np.float16 else 0
    B = np.ascontiguousarray(B, dtype=np.float16 if elem else np.uint8)
    c = np.ascontiguousarray(cand, dtype=np.int32)
    v = np.empty(ktop, dtype=np.float32)
    l = np.empty(ktop, dtype=np.int32)
    lib().rmo_rescore_rank_t(_p(q), _p(B), elem, B.shape[1], _p(c), len(c), ktop, _p(v), _p(l))
    return v, l


def rescore_dot(q32, col):
    """The rescoring dot of one column: uint8 e4m3 or float16 elements."""
    q = np.ascontiguousarray(q32, dtype=np.float32)
    elem = 1 if col.dtype == np.float16 else 0
    b = np.ascontiguousarray(col, dtype=np.float16 if elem else np.uint8)
    return lib().rmo_rescore_dot_t(_p(q), _p(b), elem, len(q))


def fasta_reads(raw, k):
    """(words (n, ceil(k/16)) uint32, headers list[bytes]) of a FASTA buffer."""
    raw = np.ascontiguousarray(np.frombuffer(raw, np.uint8) if isinstance(raw, (bytes, bytearray))
                               else raw, dtype=np.uint8)
    cap = int(np.count_nonzero(raw == ord(">"))) + 1
    W = (k + 15) // 16
    words = np.zeros((cap, W), np.uint32)
    hoff = np.zeros(cap, np.int64)
    hlen = np.zeros(cap, np.int32)
    n = lib().rmo_fasta_reads(_p(raw) if raw.size else None, raw.size, k, _p(words), cap,
                              _p(hoff), _p(hlen))
    heads = [raw[o:o + l].tobytes() for o, l in zip(hoff[:n], hlen[:n])]
    return words[:n], heads


CHAINS = {"e4m3": 0, "f32": 1, "f16": 2, "f64": 3}


def golden_query(words, k, s, m, c, chain="e4m3"):
    """The golden FP64 encoder's output of one read through a query storage
    chain ('e4m3' | 'f32' | 'f16' | 'f64'), as float64."""
    w = np.ascontiguousarray(words, dtype=np.uint32)
    q = np.empty(2 * m * 4 ** c)
    lib().rmo_golden_query(_p(w), k, s, m, c, CHAINS[chain], _p(q))
    return q


def dots_f64(q, B, cols):
    """FP64 dots of q with the 1-based columns `cols` of B (uint8 e4m3 or
    float16 rows); column 0 gives NaN."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    elem = 1 if B.dtype == np.float16 else 0
    Bv = B.view(np.uint16) if elem else B
    c = np.ascontiguousarray(cols, dtype=np.int32)
    out = np.empty(len(c))
    lib().rmo_dots_f64(_p(q), _p(Bv), elem, len(q), _p(c), len(c), _p(out))
    return out


def scan_topk_f64(Q, B, ktop, threads=None):
    """Full-scan FP64 top-k of queries Q (n, K) over all rows of B (uint8
    e4m3 or float16): (FP64 values, 1-based locs)."""
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    n, K = Q.shape
    elem = 1 if B.dtype == np.float16 else 0
    Bv = np.ascontiguousarray(B).view(np.uint16) if elem else np.ascontiguousarray(B)
    v = np.empty((n, ktop))
    l = np.empty((n, ktop), dtype=np.int32)
    lib().rmo_scan_topk_f64(_p(Q), n, _p(Bv), elem, Bv.shape[0], K, ktop, _p(v), _p(l),
                            threads or os.cpu_count() or 1)
    return v, l


def cpu_map(words, k, s, m, c, B, ktop, threads=None):
    """The CPU golden mapping path: (vals float32 (n, ktop), locs 1-based)."""
    w = np.ascontiguousarray(words, dtype=np.uint32)
    n = w.shape[0]
    B = np.ascontiguousarray(B, dtype=np.uint8)
    v = np.empty((n, ktop), dtype=np.float32)
    l = np.empty((n, ktop), dtype=np.int32)
    lib().rmo_cpu_map(_p(w), n, k, s, m, c, _p(B), B.shape[0], B.shape[1], ktop, _p(v), _p(l),
                      threads or os.cpu_count() or 1)
    return v, l
