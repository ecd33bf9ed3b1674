"""Read sampling with provenance (RotorMap/reads/sample.jl:474-611,
RotorMap/reads/provenance.jl:46-78, RotorMap/common/harness.jl:130-164).

Selection is the reference's: `:even` is the closed-form largest-remainder
allocation over each record's N-free k-window pool with the read at the
midpoint of each of n_j equal strata; `:random` draws iid uniform windows.
Mutations come from the library's counter-hash RNG (the reference's Xoshiro
streams are not reproduced; the error models are).
"""
import re
from dataclasses import dataclass

import numpy as np

from . import _lib as L


@dataclass
class ReadSet:
    words: np.ndarray     # (n, ceil(k/16)) uint32 packed reads
    recs: np.ndarray      # (n,) 0-based record index
    starts: np.ndarray    # (n,) 1-based window start in the record
    err: np.ndarray       # (n,) error rate applied
    k: int
    names: list           # record names (first header token sans '>')

    def heads(self):
        """Provenance headers: >read_<i> start=<s> len=<k> src=<name>."""
        return [f">read_{i + 1} start={s} len={self.k} src={self.names[r]}"
                for i, (r, s) in enumerate(zip(self.recs.tolist(), self.starts.tolist()))]


def even_alloc(w, n):
    """Largest-remainder quotas proportional to w (sample.jl:474-485);
    remainders tie-broken by record order."""
    w = np.asarray(w, dtype=np.int64)
    total = int(w.sum())
    alloc = (n * w) // total
    fracs = n * w / total - alloc
    order = sorted(range(len(w)), key=lambda j: (-fracs[j], j))
    for t in range(n - int(alloc.sum())):
        alloc[order[t]] += 1
    assert int(alloc.sum()) == n
    return alloc


def nfree_pool(rec, k):
    """Sorted 0-based start ranges [(a, b)) of the record's k-windows that
    touch no N run (windows fully inside the record)."""
    last = rec.length - k  # last valid 0-based start
    if last < 0:
        return []
    ranges, cur = [], 0
    for lo, hi in sorted(rec.n_runs):
        # windows [s, s+k) with s+k <= lo are clean; next clean start is hi
        a, b = cur, min(lo - k, last) + 1
        if b > a:
            ranges.append((a, b))
        cur = max(cur, hi)
    if last + 1 > cur:
        ranges.append((cur, last + 1))
    return ranges


def _rank_to_start(ranges, rank0):
    for a, b in ranges:
        if rank0 < b - a:
            return a + rank0
        rank0 -= b - a
    raise IndexError("rank beyond pool")


def sample_reads(ref, n, k, err=0.0, seed=1, draw="even", nfree=True, model=0):
    """Sample n k-base reads.  err: scalar or per-read array.  model 0 =
    substitutions only, 1 = the reference's sub/ins/del `mutate`."""
    recs_all = ref.records
    pools = [nfree_pool(r, k) if nfree else ([(0, r.length - k + 1)] if r.length >= k else [])
             for r in recs_all]
    w = np.array([sum(b - a for a, b in p) for p in pools], dtype=np.int64)
    if w.sum() <= 0:
        raise L.RotorMapArgumentError(L.RM_ERR_ARG, f"no k = {k} window in the reference pool")
    recs = np.empty(n, dtype=np.int64)
    ranks = np.empty(n, dtype=np.int64)  # 0-based pool rank
    if draw == "even":
        alloc = even_alloc(w, n)
        g = 0
        for j, nj in enumerate(alloc.tolist()):
            if nj == 0:
                continue
            Wj = int(w[j])
            t = np.arange(1, nj + 1, dtype=np.int64)
            rk = ((2 * t - 1) * Wj) // (2 * nj) if nj <= Wj else (t - 1) % Wj
            recs[g:g + nj] = j
            ranks[g:g + nj] = rk
            g += nj
    elif draw == "random":
        rng = np.random.default_rng(seed)
        r = rng.integers(0, int(w.sum()), size=n)
        cum = np.cumsum(w)
        recs[:] = np.searchsorted(cum, r, side="right")
        ranks[:] = r - np.concatenate([[0], cum])[recs]
    else:
        raise L.RotorMapArgumentError(L.RM_ERR_ARG, "draw must be 'even' or 'random'")
    starts0 = np.empty(n, dtype=np.int64)
    for j in np.unique(recs):
        sel = np.where(recs == j)[0]
        a_ = np.array([a for a, _ in pools[j]], dtype=np.int64)
        ln = np.array([b - a for a, b in pools[j]], dtype=np.int64)
        cum = np.cumsum(ln)
        r = np.searchsorted(cum, ranks[sel], side="right")  # range holding each rank
        starts0[sel] = a_[r] + ranks[sel] - (cum[r] - ln[r])
    W = (k + 15) // 16
    words = np.zeros((n, W), dtype=np.uint32)
    for j in np.unique(recs):
        sel = np.where(recs == j)[0]
        st = np.ascontiguousarray(starts0[sel])
        out = np.zeros((len(sel), W), dtype=np.uint32)
        L.check(L.lib().rm_extract_windows_host(L.ptr(out), L.ptr(recs_all[j].words), L.ptr(st),
                                                len(sel), k))
        words[sel] = out
    errs = np.broadcast_to(np.asarray(err, dtype=np.float64), (n,)).copy()
    ids = np.arange(n, dtype=np.int64)
    L.check(L.lib().rm_mutate_words(L.ptr(words), n, k, L.ptr(errs), int(seed), L.ptr(ids), model))
    return ReadSet(words, recs, starts0 + 1, errs, k, [r.name for r in recs_all])


_SAMPLE = re.compile(r"^>?(\S+)\s+start=(\d+)\s+len=(\d+)\s+src=(.*)$")
_COLON = re.compile(r"^>?(\S+?):(\d+)(?::(rc|\+|-|0|1))?(?:\s|$)")


def parse_read_head(h):
    """Both provenance formats (RotorMap/reads/provenance.jl:46-78):
    '>read_i start=S len=K src=NAME' and '>acc:start[:rc]'; returns
    dict(id, start (1-based), len or None, src) or None."""
    m = _SAMPLE.match(h.strip())
    if m:
        return dict(id=m.group(1), start=int(m.group(2)), len=int(m.group(3)),
                    src=m.group(4).split()[0] if m.group(4).strip() else "")
    m = _COLON.match(h.strip())
    if m:
        st = int(m.group(2))
        return dict(id=h.strip().lstrip(">").split()[0], start=max(st, 1), len=None, src=m.group(1))
    return None
