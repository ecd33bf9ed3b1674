"""Reference sequences: synthetic generators, 2-bit packing, FASTA I/O.

Packing follows the reader's word convention (RotorMap/fasta/reader.jl:
21-28,166-197): codes A=0 C=1 G=2 T=3, every other byte -> 2 (kept in place),
base j in word j>>4 at bit 2*(j&15); packed records carry one zero guard word.
"""
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L

_LUT = np.full(256, 2, dtype=np.uint8)
for _ch, _v in (("A", 0), ("C", 1), ("G", 2), ("T", 3)):
    _LUT[ord(_ch)] = _v
    _LUT[ord(_ch.lower())] = _v


def bytes_to_codes(b):
    """Sequence bytes (newlines already removed) -> codes 0..3."""
    return _LUT[np.frombuffer(b, dtype=np.uint8)]


def pack_codes(codes, guard=True):
    """codes (L,) uint8 -> forward words, zero tail (+ one zero guard word)."""
    codes = np.asarray(codes, dtype=np.uint32)
    Lb = len(codes)
    W = (Lb + 15) // 16
    padded = np.zeros(W * 16, dtype=np.uint32)
    padded[:Lb] = codes & 3
    sh = (2 * np.arange(16, dtype=np.uint32))[None, :]
    words = np.bitwise_or.reduce(padded.reshape(W, 16) << sh, axis=1).astype(np.uint32)
    if guard:
        words = np.concatenate([words, np.zeros(1, dtype=np.uint32)])
    return words


def unpack_words(words, L_):
    """Forward words -> codes (L_,) uint8."""
    w = np.asarray(words, dtype=np.uint32)
    sh = (2 * np.arange(16, dtype=np.uint32))[None, :]
    codes = ((w[:, None] >> sh) & 3).astype(np.uint8).reshape(-1)
    return codes[:L_]


@dataclass
class Record:
    header: str          # verbatim header line, with the leading '>'
    length: int          # sequence characters
    words: np.ndarray    # packed, ceil(length/16) + 1 words
    n_runs: list = field(default_factory=list)  # [(lo, hi)) 0-based N intervals

    @property
    def name(self):
        """First token of the header sans '>' (RotorMap/reads/provenance.jl:44)."""
        return self.header.split()[0][1:]


@dataclass
class Reference:
    records: list

    @property
    def total(self):
        return sum(r.length for r in self.records)


def synth_reference(lengths, gc=0.5, seed=1, n_frac=0.0, n_run=(10_000, 1_000_000), names=None):
    """I.i.d. bases with P(G)=P(C)=gc/2 per record (counter-hash RNG, host
    C++), optionally with N runs covering ~n_frac of each record (lengths
    log-uniform in n_run) stored as code 2, as the reader maps N."""
    recs = []
    rng = np.random.default_rng(seed)
    for r, Lr in enumerate(lengths):
        Lr = int(Lr)
        w = np.zeros((Lr + 15) // 16 + 1, dtype=np.uint32)
        L.check(L.lib().rm_synth_record(seed, r, Lr, float(gc), L.ptr(w)))
        runs = []
        if n_frac > 0:
            target = int(n_frac * Lr)
            covered = 0
            lo_l, hi_l = np.log(n_run[0]), np.log(n_run[1])
            while covered < target:
                ln = int(np.exp(rng.uniform(lo_l, hi_l)))
                ln = min(ln, target - covered + n_run[0])
                s = int(rng.integers(0, max(1, Lr - ln)))
                runs.append((s, min(Lr, s + ln)))
                covered += ln
            runs.sort()
            merged = []
            for a, b in runs:
                if merged and a <= merged[-1][1]:
                    merged[-1] = (merged[-1][0], max(merged[-1][1], b))
                else:
                    merged.append((a, b))
            runs = merged
            for a, b in runs:
                L.check(L.lib().rm_mark_n(L.ptr(w), a, b))
        nm = names[r] if names else f"chr{r + 1}"
        recs.append(Record(f">{nm} synthetic record {r + 1}", Lr, w, runs))
    return Reference(recs)


def record_codes(rec):
    return unpack_words(rec.words[:-1], rec.length)


def write_fasta(ref, path, width=70):
    """FASTA text; N runs are written as 'N' (they read back as code 2)."""
    with open(path, "w") as f:
        for r in ref.records:
            codes = record_codes(r)
            seq = np.frombuffer(b"ACGT", dtype=np.uint8)[codes].copy()
            for a, b in r.n_runs:
                seq[a:b] = ord("N")
            f.write(r.header + "\n")
            s = seq.tobytes().decode()
            for i in range(0, len(s), width):
                f.write(s[i:i + width] + "\n")


def read_fasta(path):
    """Records of a FASTA file: header line verbatim, junk before the first
    '>' discarded, '\\n'/'\\r' skipped (RotorMap/common/harness.jl:68-100)."""
    raw = open(path, "rb").read()
    recs = []
    pos = 0
    n = len(raw)
    # record starts: '>' at a line start
    starts = [i for i in range(n) if raw[i] == 0x3E and (i == 0 or raw[i - 1] in (0x0A, 0x0D))]
    for si, s in enumerate(starts):
        e = starts[si + 1] if si + 1 < len(starts) else n
        h = raw.find(b"\n", s, e)
        h = e if h < 0 else h
        header = raw[s:h].rstrip(b"\r").decode(errors="replace")
        seq = raw[h + 1:e].replace(b"\n", b"").replace(b"\r", b"")
        arr = np.frombuffer(seq, dtype=np.uint8)
        isn = (_LUT[arr] == 2) & (arr != ord("G")) & (arr != ord("g"))
        runs = []
        if isn.any():
            d = np.diff(np.concatenate([[0], isn.astype(np.int8), [0]]))
            runs = list(zip(np.where(d == 1)[0].tolist(), np.where(d == -1)[0].tolist()))
        recs.append(Record(header, len(seq), pack_codes(bytes_to_codes(seq)), runs))
    return Reference(recs)
