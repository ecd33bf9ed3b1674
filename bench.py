"""Mapping throughput on one MI355X: reads/s of the fused FP8 + FP32-rescore
path, with top-20 recall (err = 0.05) beside the FP8 and FP16 paths, the
fused kernel's roofline, request latency through the serving loop, and the
CPU golden path timed on the host cores.

Workloads (synthetic; one step = one batch of reads resident in HBM):
  100mbp            100 Mbp, 4 records x 25 Mbp, i.i.d. bases, GC 0.41;
                    index k = 20,000, kstep = 2,000, s5m1c5 (rdim 2048), fwd +
                    rc windows -> N = 99,928 e4m3 columns; batches of 65,536
                    reads of 20 kbp sampled :even, substitution rate per read
                    from {0, 0.02, 0.05, 0.10}; recall over the err = 0.05 reads
  human (default)   3.1 Gbp, 24 records of 50-250 Mbp (human chromosome
                    sizes), GC 0.41, 5 % of bases in N runs of 10 kbp-1 Mbp ->
                    N ~ 3.09 M columns (6.3 GB e4m3 resident); batches of
                    65,536 reads :even (sub/ins/del), the first 16,384 at
                    err = 0.05 (the recall rows), the rest with err uniform
                    in [0, 0.15]; also uneven
                    batches (log-uniform sizes in [1, 131072], err in
                    [0, 0.10]) one launch per batch vs one launch in all, after
                    the autotuner
  step              encode -> fused e4m3 GEMM + top-32 -> FP32 rescore -> top-20
  fasta             the batch written as FASTA text (80-column lines), parsed and
                    2-bit packed on the GPU; checked equal to the packed reads
  serving           Poisson request traces at 50 / 200 / 1000 req/s (human: also
                    10 req/s, below saturation; reads per request log-uniform
                    on [1, 100000]) through the continuous-batching server
                    (queued rows fill a launch fewest-remaining-first;
                    --serve-policy fifo for arrival order): p50 / p99 / p999
                    latency, offered vs achieved reads/s

python bench.py [--workload human|100mbp] [--steps K] [--warmup W] [--gpus 1]
                [--reads M] [--no-cpu] [--no-serve] [--no-uneven] [--no-fasta]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "qpg-rotormap_amd"))

K_FRAG, KSTEP, S, M_COPIES, C = 20_000, 2_000, 5, 1, 5
GC = 0.41
PEAK_FP8_TFLOPS = 5000.0   # MI355X dense FP8 MFMA
PEAK_FP16_TFLOPS = 2500.0  # MI355X dense FP16 MFMA
PEAK_HBM_GBS = 8000.0
# human chromosome sizes (Mbp, chr1..22, X, Y) clamped to [50, 250]
HUMAN_MBP = [248, 242, 198, 190, 181, 171, 159, 145, 138, 134, 135, 133, 114, 107, 102, 90, 83,
             80, 59, 64, 50, 51, 156, 57]


def lscpu_summary():
    info = {}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except Exception:
        pass
    return {"model": info.get("Model name"), "sockets": info.get("Socket(s)"),
            "cores_per_socket": info.get("Core(s) per socket"),
            "threads_per_core": info.get("Thread(s) per core"), "nproc": os.cpu_count()}


def traffic_from_profile(name):
    """HBM bytes per launch of the fused kernel from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor)."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    except Exception:
        return None


def fasta_text(words, k, width=80):
    """Packed reads -> FASTA text: '>read_%08d' headers, sequence lines of
    `width` bases (built in chunks of 8192 reads)."""
    n = words.shape[0]
    nl = -(-k // width)
    rows = []
    shifts = 2 * np.arange(16, dtype=np.uint32)
    for c0 in range(0, n, 8192):
        w = words[c0:c0 + 8192]
        m = w.shape[0]
        codes = ((w[:, :, None] >> shifts) & 3).astype(np.uint8).reshape(m, -1)[:, :k]
        pad = np.full((m, nl * width), ord("\n"), np.uint8)
        pad[:, :k] = np.frombuffer(b"ACGT", np.uint8)[codes]
        body = np.concatenate([pad.reshape(m, nl, width), np.full((m, nl, 1), ord("\n"), np.uint8)],
                              axis=2).reshape(m, -1)[:, :k + nl]
        idx = np.arange(c0, c0 + m)
        head = np.zeros((m, 15), np.uint8)
        head[:, :6] = np.frombuffer(b">read_", np.uint8)
        for d in range(8):
            head[:, 6 + d] = ord("0") + (idx // 10 ** (7 - d)) % 10
        head[:, 14] = ord("\n")
        rows.append(np.concatenate([head, body], axis=1).reshape(-1))
    return np.concatenate(rows)


def rnd(x, n=3):
    return round(float(x), n)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=["human", "100mbp"], default="human")
    ap.add_argument("--steps", type=int, default=None, help="timed steps (20; human: 5)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (3; human: 1)")
    ap.add_argument("--reads", type=int, default=65_536)
    ap.add_argument("--fp16-steps", type=int, default=1)
    ap.add_argument("--cpu-reads", type=int, default=0, help="0 = size the CPU sample to ~15 s")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--no-fasta", action="store_true")
    ap.add_argument("--no-vendor", action="store_true",
                    help="skip timing hipBLASLt's GEMM at the same M and K")
    ap.add_argument("--serve-rates", default=None,
                    help="Poisson arrival rates, req/s (100mbp: 50,200,1000; human: 10,50,200,1000)")
    ap.add_argument("--serve-requests", default=None,
                    help="requests per rate (100mbp: 200,400,600; human: 40 each)")
    ap.add_argument("--serve-rows", type=int, default=16384, help="rows per serving launch")
    ap.add_argument("--serve-streams", type=int, default=2, help="serving launch slots")
    ap.add_argument("--serve-policy", choices=["fifo", "srpt"], default="srpt",
                    help="order in which queued rows fill a launch")
    ap.add_argument("--uneven-batches", type=int, default=48,
                    help="first batches of the 4096-size uneven sequence (human)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one 1 GiB device copy (the HBM-counter calibration kernel)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    human = args.workload == "human"
    steps = args.steps if args.steps is not None else (5 if human else 20)
    warmup = args.warmup if args.warmup is not None else (1 if human else 3)
    serve_rates = args.serve_rates or ("10,50,200,1000" if human else "50,200,1000")
    serve_req = args.serve_requests or ("40,40,40,40" if human else "200,400,600")

    import torch
    import rotormap_amd as R
    from rotormap_amd import _lib as L

    if args.calibrate_copy:
        a = torch.empty(1 << 30, dtype=torch.uint8, device="cuda:0")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    t0 = time.time()
    log = lambda *a: print(f"[bench {time.time() - t0:7.1f}s]", *a, file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    re = R.RopeEncoder(k=K_FRAG, s=S, m=M_COPIES, c=C)
    M = args.reads
    POOL = max(M, 100_000)  # serving draws requests (up to 100,000 reads) from this pool

    # ---- reference, index, reads ----------------------------------------
    if human:
        ref = R.synth_reference([mb * 1_000_000 for mb in HUMAN_MBP], gc=GC, seed=300,
                                n_frac=0.05, n_run=(10_000, 1_000_000))
    else:
        ref = R.synth_reference([25_000_000] * 4, gc=GC, seed=100)
    log(f"reference synthesised: {len(ref.records)} records, {ref.total / 1e6:.0f} Mbp")
    torch.cuda.synchronize()
    tb = time.perf_counter()
    idx = R.build_index(ref, re, kstep=KSTEP, batch_size=8192)
    t_index = time.perf_counter() - tb
    N = idx.embeds8.shape[0]
    log(f"index: {N} columns x {idx.embeds8.shape[1]} e4m3 built in {t_index:.1f} s")
    rng = np.random.default_rng(7)
    if human:
        # the recall rows ride in every batch: rows [0, 16384) at err 0.05
        errs = rng.uniform(0.0, 0.15, POOL)
        errs[:min(16_384, M)] = 0.05
        pool = R.sample_reads(ref, POOL, K_FRAG, err=errs, seed=11, draw="even", model=1)
    else:
        ERRS = (0.0, 0.02, 0.05, 0.10)
        errs = np.array(ERRS)[rng.integers(0, len(ERRS), POOL)]
        pool = R.sample_reads(ref, POOL, K_FRAG, err=errs, seed=11, draw="even", model=0)
    log(f"{POOL} reads sampled")
    words_all = torch.from_numpy(pool.words.view(np.int32).copy()).to(dev)  # (POOL, W)
    words = words_all[:M].reshape(-1)

    eng8 = R.TopKEngine(idx.embeds8, k=20, kprime=32, mode="fp8_rescore", rows_cap=M)
    ctx = L.ctx(0)

    def run(eng, w, n, steps_, warmup_, name):
        for _ in range(warmup_):
            eng.map_words(re, w, n)
        torch.cuda.synchronize()
        L.check(L.lib().rm_prof_read(ctx, None, None, None))  # clear
        L.check(L.lib().rm_prof_enable(ctx, 1))
        t = time.perf_counter()
        out = None
        for _ in range(steps_):
            out = eng.map_words(re, w, n)
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        kms = (L.C.c_double(), L.C.c_int64(), L.C.c_double())
        L.check(L.lib().rm_prof_read(ctx, L.C.byref(kms[0]), L.C.byref(kms[1]), L.C.byref(kms[2])))
        L.check(L.lib().rm_prof_enable(ctx, 0))
        log(f"{name}: {steps_} steps of {n} reads in {el:.3f}s; fused kernel "
            f"{kms[0].value / max(1, kms[1].value):.2f} ms/launch")
        return el, kms[0].value, kms[1].value, kms[2].value, out

    # ---- timed FP8 + rescore steps (the headline) ---------------------------
    el8, kms8, nl8, flop8, (v8, l8) = run(eng8, words, M, steps, warmup, "fp8+rescore")
    value = M * steps / el8
    k_avg_ms = kms8 / max(1, nl8)
    achieved = (flop8 / max(1, nl8)) / (k_avg_ms * 1e-3) / 1e12

    # ---- recall and the other modes ----------------------------------------
    eng8p = R.TopKEngine(eng8.index, k=20, mode="fp8", rows_cap=M)
    eng16 = R.TopKEngine(R.ResidentIndex(idx.embeds8, fp16=True), k=20, mode="fp16", rows_cap=M)
    el8p, _, _, _, (_, l8p) = run(eng8p, words, M, 1, 1, "fp8")
    el16, kms16, nl16, flop16, (_, l16) = run(eng16, words, M, args.fp16_steps, 1, "fp16")

    def recall(locs, rs, sel=None):
        if sel is not None:
            locs, recs, starts = locs[sel], rs.recs[sel], rs.starts[sel]
        else:
            recs, starts = rs.recs, rs.starts
        r = R.count_mapped(locs, recs, starts, idx.rec_of_col, idx.starts, K_FRAG)
        return r["correct"] / max(1, r["total"]), r["total"]

    sel = np.where(np.isclose(errs[:M], 0.05))[0]
    rec = {"fp8_rescore": recall(l8.cpu().numpy(), pool, sel)[0],
           "fp8": recall(l8p.cpu().numpy(), pool, sel)[0],
           "fp16": recall(l16.cpu().numpy(), pool, sel)[0],
           "reads": int(len(sel))}
    rec_all = recall(l8.cpu().numpy(), pool, np.arange(M))[0]
    log(f"recall at err 0.05: {rec}")

    if human:
        wl = ("3.1 Gbp synthetic human-scale reference (24 records of 50-250 Mbp, GC 0.41, 5 %% N "
              "in 10 kbp-1 Mbp runs), s5m1c5, k=20000 kstep=2000, N=%d e4m3 columns (%.2f GB "
              "resident); one batch of %d reads (:even, sub/ins/del; 16,384 at err 0.05 for "
              "recall, the rest err uniform in [0, 0.15]); "
              "top-20 after FP32 rescoring of the fused top-32" % (N, N * 2048 / 1e9, M))
    else:
        wl = ("100 Mbp synthetic reference (4 x 25 Mbp, GC 0.41), s5m1c5, k=20000 kstep=2000, "
              "N=%d e4m3 columns; one batch of %d reads (:even, substitutions at "
              "{0,0.02,0.05,0.10}); top-20 after FP32 rescoring of the fused top-32" % (N, M))
    result = {
        "metric": "mapped reads/sec at top-20 recall (err=0.05) FP8+rescore vs FP16; p50/p99 latency",
        "value": round(value, 1),
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(el8 / steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "e4m3 (fp32 accumulate, fp32 rescore)",
        "data": "synthetic",
        "config": {"workload": wl, "reads_per_step": M, "index_columns": int(N), "rdim": 2048,
                   "ktop": 20, "kprime": 32, "parallelism": "single GPU"},
        "recall_err005": rec,
        "recall_timed_batch_fp8_rescore": rnd(rec_all, 4),
        "reads_per_s": {"fp8_rescore": round(value, 1), "fp8": round(M / el8p, 1),
                        "fp16": round(M * args.fp16_steps / el16, 1)},
        "index_build_s": rnd(t_index, 2),
        "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK_FP8_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP8_TFLOPS, 4),
                     "traffic": traffic_from_profile("gemm_topk_fp8_%s_pmc.json" % args.workload),
                     "kernel": "gemm_topk_kernel<e4m3>", "kernel_ms": round(k_avg_ms, 3),
                     "flop_per_launch": flop8 / max(1, nl8)},
        "roofline_fp16": {"achieved": round((flop16 / max(1, nl16)) / (kms16 / max(1, nl16) * 1e-3) / 1e12, 2),
                          "peak": PEAK_FP16_TFLOPS, "unit": "TFLOP/s"},
    }
    del eng16
    torch.cuda.empty_cache()

    # ---- the vendor GEMM at the same M and K: hipBLASLt through torch, plain C write
    if not args.no_vendor:
        nv = int(min(N, 98_304))
        ga = (torch.randn(M, 2048, device=dev) / 8).to(torch.float8_e4m3fn)
        gb = (torch.randn(nv, 2048, device=dev) / 8).to(torch.float8_e4m3fn)
        one = torch.ones((), device=dev)

        def vt(fn, reps):
            fn()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                fn()
            e1.record()
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) / reps

        fl = 2.0 * M * nv * 2048
        ms8 = vt(lambda: torch._scaled_mm(ga, gb.t(), scale_a=one, scale_b=one,
                                          out_dtype=torch.bfloat16), 10)
        ga16, gb16 = ga.to(torch.float16), gb.to(torch.float16)
        ms16 = vt(lambda: ga16 @ gb16.t(), 5)
        v8, v16 = fl / ms8 / 1e9, fl / ms16 / 1e9
        result["vendor_gemm"] = {
            "library": "hipBLASLt via torch._scaled_mm (e4m3fn, bf16 C) and torch.matmul (fp16 C)",
            "shape_mnk": [M, nv, 2048], "fp8_tflops": round(v8, 1), "fp16_tflops": round(v16, 1),
            "fused_fp8_over_vendor_fp8": round(achieved / v8, 3),
            "fused_fp16_over_vendor_fp16": round(result["roofline_fp16"]["achieved"] / v16, 3)}
        log(f"vendor GEMM at M={M} N={nv} K=2048: fp8 {v8:.0f} TFLOP/s, fp16 {v16:.0f} TFLOP/s")
        del ga, gb, ga16, gb16
        torch.cuda.empty_cache()

    # ---- FASTA ingestion: the batch as FASTA text, parsed and packed on the GPU
    if not args.no_fasta:
        W = (K_FRAG + 15) // 16
        nfa = M
        width = 80
        text = fasta_text(pool.words[:nfa], K_FRAG, width)
        pin = torch.from_numpy(text).pin_memory()
        draw = torch.empty(text.size, dtype=torch.uint8, device=dev)
        torch.cuda.synchronize()
        th = time.perf_counter()
        draw.copy_(pin, non_blocking=True)
        torch.cuda.synchronize()
        t_h2d = time.perf_counter() - th
        fw = torch.empty(nfa, W, dtype=torch.int32, device=dev)
        nread = L.C.c_int64()
        st = L.stream_of(0)
        L.check(L.lib().rm_fasta_reads(ctx, L.ptr(draw), text.size, K_FRAG, L.ptr(fw), nfa, None,
                                       None, L.C.byref(nread), st))  # warm-up
        torch.cuda.synchronize()
        reps = 5
        tp = time.perf_counter()
        for _ in range(reps):
            L.check(L.lib().rm_fasta_reads(ctx, L.ptr(draw), text.size, K_FRAG, L.ptr(fw), nfa, None,
                                           None, L.C.byref(nread), st))
        torch.cuda.synchronize()
        t_parse = (time.perf_counter() - tp) / reps
        same = bool(nread.value == nfa and torch.equal(fw, words_all[:nfa]))
        result["fasta_ingest"] = {
            "bytes": int(text.size), "reads": int(nread.value), "line_width": width,
            "gpu_parse_s": rnd(t_parse, 5), "gpu_parse_gb_s": rnd(text.size / t_parse / 1e9, 1),
            "h2d_gb_s": rnd(text.size / t_h2d / 1e9, 1),
            "words_equal_packed_reads": same}
        log(f"fasta ingest: {text.size / 1e9:.2f} GB parsed in {t_parse * 1e3:.2f} ms "
            f"({text.size / t_parse / 1e9:.0f} GB/s on the GPU, H2D {text.size / t_h2d / 1e9:.1f} GB/s); "
            f"round trip {'ok' if same else 'MISMATCH'}")
        del pin, draw, fw, text

    # ---- uneven batches: one launch per batch vs one launch for all ---------
    if human and not args.no_uneven:
        urng = np.random.default_rng(404)
        sizes = np.floor(np.exp(urng.uniform(0, np.log(131_073), 4096))).astype(np.int64)
        sizes = np.clip(sizes, 1, 131_072)[:args.uneven_batches]
        total = int(sizes.sum())
        uerr = urng.uniform(0.0, 0.10, total)
        ureads = R.sample_reads(ref, total, K_FRAG, err=uerr, seed=13, draw="random", model=1)
        uw = torch.from_numpy(ureads.words.view(np.int32).reshape(-1).copy()).to(dev)
        W = (K_FRAG + 15) // 16
        offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        engu = R.TopKEngine(eng8.index, k=20, kprime=32, mode="fp8_rescore", rows_cap=total)
        tt = time.perf_counter()
        buckets = 2 ** np.arange(0, 18)
        tile, best, qg, ms = engu.autotune(buckets)
        t_tune = time.perf_counter() - tt
        log(f"autotune: {len(buckets)} buckets in {t_tune:.1f} s; tiles {tile.tolist()} "
            f"splits {best.tolist()}")
        torch.cuda.synchronize()
        t = time.perf_counter()
        for b in range(len(sizes)):
            engu.map_words(re, uw[offs[b] * W:offs[b + 1] * W], int(sizes[b]))
        torch.cuda.synchronize()
        t_each = time.perf_counter() - t
        t = time.perf_counter()
        vu, lu = engu.map_batches(re, uw, offs)
        torch.cuda.synchronize()
        t_one = time.perf_counter() - t
        ru = recall(lu.cpu().numpy(), ureads)[0]
        result["uneven"] = {
            "batches": int(len(sizes)), "reads": total, "sizes": "first %d of 4096 log-uniform "
            "in [1, 131072] (seed 404)" % len(sizes), "err": "uniform [0, 0.10], sub/ins/del",
            "one_launch_per_batch_reads_s": rnd(total / t_each, 1),
            "single_persistent_launch_reads_s": rnd(total / t_one, 1),
            "autotune_s": rnd(t_tune, 2),
            "autotuned_plans": {int(b): [int(t), int(s), int(g)]
                                for b, t, s, g in zip(buckets, tile, best, qg)},
            "recall_top20": rnd(ru, 4)}
        log(f"uneven: per-batch {total / t_each:.0f} reads/s, one launch {total / t_one:.0f} reads/s")
        del engu, uw

    # ---- serving -----------------------------------------------------------
    if not args.no_serve:
        pinned = torch.from_numpy(pool.words.view(np.int32)).pin_memory()
        rates = [float(x) for x in serve_rates.split(",")]
        nreq = [int(x) for x in serve_req.split(",")]
        serving = {"mode": "fp8_rescore", "rows_per_launch": args.serve_rows,
                   "launch_slots": args.serve_streams, "policy": args.serve_policy,
                   "request_reads": "log-uniform on [1, 100000]", "pool_reads": POOL, "rates": []}
        with R.Server(eng8.index, re, k=20, kprime=32, mode="fp8_rescore",
                      max_rows=args.serve_rows, n_streams=args.serve_streams,
                      policy=args.serve_policy) as srv:
            for i, (rate, nr) in enumerate(zip(rates, nreq)):
                t, n = R.poisson_trace(nr, rate, seed=100 + i)
                res = R.run_trace(srv, pinned, t, n, seed=200 + i)
                res.pop("results"), res.pop("offsets")
                res = {k: (rnd(v) if isinstance(v, float) else v) for k, v in res.items()}
                serving["rates"].append({"rate_req_s": rate, **res})
                log(f"serve {rate:g} req/s: p50 {res['p50_ms']} ms p99 {res['p99_ms']} ms "
                    f"p999 {res['p999_ms']} ms; offered {res['offered_reads_s']:.0f} "
                    f"achieved {res['achieved_reads_s']:.0f} reads/s")
            serving["server_stats"] = srv.stats()
        del pinned
        result["serving"] = serving
        # summary at the lowest stated arrival rate of the configuration (50 req/s)
        mid = next((r for r in serving["rates"] if r["rate_req_s"] >= 50), serving["rates"][-1])
        result["latency"] = {"rate_req_s": mid["rate_req_s"], "p50_ms": mid["p50_ms"],
                             "p99_ms": mid["p99_ms"], "p999_ms": mid["p999_ms"]}

    # ---- CPU golden path on the host cores ---------------------------------
    if not args.no_cpu:
        sys.path.insert(0, ROOT)
        import oracle.oracle as O
        cores = min(os.cpu_count() or 1, 16)
        B = np.ascontiguousarray(idx.embeds8)
        fn = O.lib().rmo_cpu_map
        fn.argtypes = [L.C.c_void_p, L.C.c_int64, L.C.c_int, L.C.c_int, L.C.c_int, L.C.c_int,
                       L.C.c_void_p, L.C.c_int64, L.C.c_int, L.C.c_int, L.C.c_void_p,
                       L.C.c_void_p, L.C.c_int]
        fn.restype = None
        locs8 = l8.cpu().numpy()

        def cpu(n_cpu):
            wsub = np.ascontiguousarray(pool.words[:n_cpu])
            vals = np.empty((n_cpu, 20), np.float32)
            locs = np.empty((n_cpu, 20), np.int32)
            t = time.perf_counter()
            fn(L.ptr(wsub), n_cpu, K_FRAG, S, M_COPIES, C, L.ptr(B), N, 2048, 20, L.ptr(vals),
               L.ptr(locs), cores)
            return time.perf_counter() - t, locs

        t1, _ = cpu(cores)  # one read per thread sizes the ~15 s sample
        n_cpu = args.cpu_reads or int(min(4096, cores * max(1, round(15.0 / max(t1, 1e-3)))))
        log(f"cpu baseline: {n_cpu} reads on {cores} threads ({t1:.2f} s per {cores})")
        cel, locs = cpu(n_cpu)
        agree = float(np.mean([len(set(locs[i]) & set(locs8[i])) / 20 for i in range(n_cpu)]))
        result["cpu_baseline"] = {"value": round(n_cpu / cel, 3), "unit": "reads/s",
                                  "cores": cores, "kind": "port",
                                  "sample": f"first {n_cpu} reads of the batch, FP64 golden encoder "
                                            f"+ FP64 dots over all {N} columns + top-20 "
                                            f"({cel:.1f} s)",
                                  "top20_overlap_with_gpu": round(agree, 4),
                                  "host": lscpu_summary()}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
