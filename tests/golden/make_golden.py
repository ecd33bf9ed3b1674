"""Writes encoder_closed_forms.json: encoder outputs of periodic fragments
derived from the geometric series alone (no oracle code involved).

A fragment of period P (P = 1: homopolymer; P = 2: alternating) has one
compact key per residue class r of the window start; the bin of class r
collects sum_{p = r mod P, p < nw} exp(i f theta p), a geometric series in
exp(i f theta P).  The G-homopolymer key 2*sum_{j<c}4^j is zeroed; normalize
= 0 divides by the root of the total energy (+1e-16)
(RotorMap/encode/reference.jl:37-105).
Run: python tests/golden/make_golden.py
"""
import cmath
import json
import math
import os


def key(codes, p, s, c):
    smer = 0
    for j in range(s):
        smer |= codes[p + j] << (2 * j)
    mask = 4 ** c - 1
    return (smer & mask) ^ ((((smer >> (2 * c)) * 0x9E3779B1) % 2 ** 32) & mask)


def case(name, pattern, k, s, m, c):
    P = len(pattern)
    codes = [pattern[j % P] for j in range(k)]
    nw = k - s + 1
    D = m * 4 ** c
    gkey = 2 * sum(4 ** j for j in range(c))
    hist = {}
    for r in range(min(P, nw)):
        kr = key(codes, r, s, c)
        n_r = len(range(r, nw, P))
        for i in range(m):
            f = 1.0 if m == 1 else 2 * i / (m - 1) + 1
            th = f * 2 * math.pi / k
            z = cmath.exp(1j * th * P)
            series = cmath.exp(1j * th * r) * ((1 - z ** n_r) / (1 - z) if abs(1 - z) > 0 else n_r)
            b = kr * m + i
            hist[b] = hist.get(b, 0) + series
    for i in range(m):
        hist.pop(gkey * m + i, None)
    nrm = math.sqrt(sum(abs(v) ** 2 for v in hist.values())) + 1e-16
    nz = []
    for b, v in sorted(hist.items()):
        nz.append([b, (v / nrm).real])
        nz.append([D + b, (v / nrm).imag])
    out = dict(name=name, k=k, s=s, m=m, c=c, nonzero=nz)
    if P == 1:
        out["fill"] = pattern[0]
    else:
        out["codes"] = codes
    return out


cases = [
    case("all-A s8m4c4", [0], 2000, 8, 4, 4),
    case("all-C s5m1c5", [1], 2001, 5, 1, 5),
    case("all-T s16m2c4", [3], 17, 16, 2, 4),
    case("all-G s8m4c4 (key 0x20 survives)", [2], 2000, 8, 4, 4),
    case("AC repeat s8m4c4", [0, 1], 2001, 8, 4, 4),
    case("GT repeat s5m1c5", [2, 3], 600, 5, 1, 5),
]
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "encoder_closed_forms.json"), "w") as f:
    json.dump(cases, f)
print(len(cases), "cases")
