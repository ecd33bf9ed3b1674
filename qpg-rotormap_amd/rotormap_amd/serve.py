"""Serving on one GPU: continuous batching over a resident index, and a
Poisson load generator that reports p50 / p99 / p999 request latency.

The loop itself is native (rm_server_* in librotormap_hip.so): a queue of
requests, a dispatcher that packs queued rows into the next free launch slot
in arrival order ("fifo") or fewest remaining rows first ("srpt"; a request
larger than a slot is spread over launches, a launch carries several
requests), slots with their own streams and buffers
allocated once, and a completer that scatters finished rows back to their
requests.  The reference has no serving loop (its "server" mode is a file
watcher, RotorMap/legacy/src/RotorMap.jl:233); the per-launch work is
exactly its batch_gemm_topk! flow (RotorMap/search/engine_fp8.jl:399-437)
and the admission rule is its resident-index byte model
(RotorMap/search/engine_compact.jl:195-242) applied to the slots.
"""
import ctypes as C
import threading
import time
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .encode import RopeEncoder
from .search import ResidentIndex, _mode


@dataclass
class ServedRequest:
    """A finished request: vals/locs (n, k) row-major (locs 1-based global
    column ids, 0 = empty) and its submit / admit / finish times in seconds
    on the server clock."""
    vals: np.ndarray
    locs: np.ndarray
    t_submit: float
    t_admit: float
    t_finish: float

    @property
    def latency(self):
        return self.t_finish - self.t_submit


class Server:
    """rm_server over `index` for encoder `re`.  max_rows: rows per launch;
    n_streams: launch slots in flight; hbm_budget: byte cap on the slots'
    buffers (0 = the free HBM less 96 MiB); policy: "fifo" or "srpt", the
    order in which queued rows fill a launch (set_policy changes it);
    lowload: (launch_rows, backlog_rows) for set_lowload, or None."""

    def __init__(self, index, re: RopeEncoder, k=20, kprime=32, mode="fp8_rescore",
                 max_rows=65536, n_streams=2, hbm_budget=0, device=0, policy="fifo",
                 lowload=None):
        if not isinstance(index, ResidentIndex):
            index = ResidentIndex(index, fp16=(mode == "fp16"), device=device)
        self.index, self.re, self.k = index, re, int(k)
        self.words_per_read = (re.k + 15) // 16
        h = C.c_void_p()
        L.check(L.lib().rm_server_create(L.ctx(device), index.handle, re.k, re.s, re.m, re.c,
                                         int(k), int(kprime), _mode(mode), int(max_rows),
                                         int(n_streams), int(hbm_budget), C.byref(h)))
        self.handle = h
        self._refs = {}
        self._mu = threading.Lock()
        self.set_policy(policy)
        self.lowload = (0, 0)
        if lowload:
            self.set_lowload(*lowload)

    def set_lowload(self, launch_rows, backlog_rows):
        """Low-load mode (0 = off): while at most backlog_rows rows wait
        unlaunched, one launch is in flight at a time and it holds whole
        requests only (a request larger than launch_rows gets a slice), so a
        small request that arrives during a large one waits at most for the
        running launch; full, padded launches return once a backlog builds."""
        L.check(L.lib().rm_server_set_lowload(self.handle, int(launch_rows), int(backlog_rows)))
        self.lowload = (int(launch_rows), int(backlog_rows))

    def set_policy(self, policy):
        """"fifo" (arrival order) or "srpt" (fewest remaining rows first)."""
        codes = {"fifo": L.RM_SCHED_FIFO, "srpt": L.RM_SCHED_SRPT}
        if policy not in codes:
            raise ValueError(f"policy must be one of {sorted(codes)}, got {policy!r}")
        L.check(L.lib().rm_server_set_policy(self.handle, codes[policy]))
        self.policy = policy

    def submit(self, words, copy=True):
        """Enqueue n fragments: words (n, W) uint32 (or flat n*W).  copy=False
        keeps a reference to `words` until wait() returns (no host copy)."""
        w = np.asarray(words)
        if w.dtype not in (np.uint32, np.int32):
            raise L.RotorMapArgumentError(L.RM_ERR_ARG, "words must be uint32")
        if not w.flags.c_contiguous:
            if not copy:
                raise L.RotorMapArgumentError(L.RM_ERR_ARG, "copy=False needs contiguous words")
            w = np.ascontiguousarray(w)
        if w.size % self.words_per_read:
            raise L.RotorMapArgumentError(
                L.RM_ERR_SHAPE, f"words hold {w.size} words, not a multiple of W = {self.words_per_read}")
        n = w.size // self.words_per_read
        rid = C.c_int64()
        fn = L.lib().rm_server_submit if copy else L.lib().rm_server_submit_ref
        L.check(fn(self.handle, L.ptr(w), n, C.byref(rid)))
        if not copy:
            with self._mu:
                self._refs[rid.value] = w
        return rid.value, n

    def submit_ptr(self, addr, n):
        """Enqueue n fragments at host address `addr` without copying (the
        caller keeps the memory valid until wait() returns)."""
        rid = C.c_int64()
        L.check(L.lib().rm_server_submit_ref(self.handle, C.c_void_p(addr), int(n), C.byref(rid)))
        return rid.value

    def wait(self, rid, n, timeout_ms=-1):
        vals = np.empty((n, self.k), np.float32)
        locs = np.empty((n, self.k), np.int32)
        ts = [C.c_double(), C.c_double(), C.c_double()]
        try:
            L.check(L.lib().rm_server_wait(self.handle, rid, int(timeout_ms), L.ptr(vals),
                                           L.ptr(locs), *[C.byref(t) for t in ts]))
        finally:
            with self._mu:
                self._refs.pop(rid, None)
        return ServedRequest(vals, locs, ts[0].value, ts[1].value, ts[2].value)

    def cancel(self, rid):
        L.check(L.lib().rm_server_cancel(self.handle, rid))

    def stats(self):
        a = [C.c_int64(), C.c_int64(), C.c_int64()]
        L.check(L.lib().rm_server_stats(self.handle, *[C.byref(x) for x in a]))
        return {"launches": a[0].value, "rows_done": a[1].value, "max_inflight": a[2].value}

    def now(self):
        return L.lib().rm_server_now(self.handle)

    def close(self):
        if self.handle:
            L.lib().rm_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_trace(n_requests, rate, seed, min_reads=1, max_reads=100_000):
    """Arrival times (s) of a Poisson process at `rate` req/s and per-request
    read counts log-uniform on [min_reads, max_reads] (mean ~8,686 for
    [1, 100,000])."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, n_requests)
    t = np.cumsum(gaps) - gaps[0]
    lo, hi = np.log(min_reads), np.log(max_reads + 1)
    sizes = np.floor(np.exp(rng.uniform(lo, hi, n_requests))).astype(np.int64)
    return t, np.clip(sizes, min_reads, max_reads)


def run_trace(server, pool, arrivals, sizes, seed=0, keep_results=False):
    """Replay a trace against `server`.  pool: (P, W) uint32 host words
    (pinned torch tensor or numpy array); request i takes sizes[i] rows of
    the pool from a seeded offset (wrapping rows are avoided by clamping the
    offset).  Submission keeps the arrival schedule on the server clock;
    returns latency percentiles and offered / achieved reads/s (and the
    finished requests with their offsets into the pool if keep_results)."""
    P, W = pool.shape
    addr0 = pool.data_ptr() if hasattr(pool, "data_ptr") else pool.ctypes.data
    rng = np.random.default_rng(seed)
    sizes = np.minimum(sizes, P)
    offs = (rng.random(len(sizes)) * (P - sizes + 1)).astype(np.int64)
    ids = [None] * len(sizes)
    done = []
    lag = []

    errors = []

    def waiter():
        try:
            for i in range(len(sizes)):
                while ids[i] is None:
                    time.sleep(0.0002)
                r = server.wait(ids[i], int(sizes[i]))
                if not keep_results:
                    r.vals = r.locs = None
                done.append(r)
        except Exception as e:  # reported after the join
            errors.append(e)

    th = threading.Thread(target=waiter, daemon=True)
    th.start()
    t0 = server.now()
    for i, (ta, n, o) in enumerate(zip(arrivals, sizes, offs)):
        while True:
            dt = ta - (server.now() - t0)
            if dt <= 0:
                break
            time.sleep(min(dt, 0.002))
        lag.append(server.now() - t0 - ta)
        ids[i] = server.submit_ptr(addr0 + int(o) * W * 4, int(n))
    th.join()
    if errors:
        raise errors[0]
    lat = np.array([d.latency for d in done])
    t_first = min(d.t_submit for d in done)
    t_last = max(d.t_finish for d in done)
    total = int(np.sum(sizes))
    span = float(arrivals[-1] - arrivals[0]) if len(arrivals) > 1 else 1.0
    return {"requests": len(sizes), "reads": total,
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "p999_ms": float(np.percentile(lat, 99.9) * 1e3),
            "mean_ms": float(lat.mean() * 1e3),
            "offered_reads_s": total / max(span, 1e-9),
            "achieved_reads_s": total / max(t_last - t_first, 1e-9),
            "max_submit_lag_ms": float(max(lag) * 1e3),
            "results": done if keep_results else None, "offsets": offs}
