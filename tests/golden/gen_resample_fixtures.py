"""Writes resample_fixtures.json: resampler known answers in exact rational arithmetic.

Each case holds dyadic weights and uniforms, chosen so that every floating-point operation
the reference performs on them (the running cumsum, n * W, division by W_n when W_n is a
power of two, u0 + (k - 1), r + u) is exact in FP64.  The expected ancestors are then the
exact-arithmetic values of each definition, computed here with fractions.Fraction, so they
are also what the reference's floating-point code returns -- including exact ties, which
pin the strict / non-strict comparisons.

Definitions (0-based outputs; paths relative to the reference checkout):
  systematic         GF/src/resamplers.jl:216-238  a_k = min{j : n W_j > u0 + k}
                     (no j satisfies it only when sum(w) < 1: clamped to n-1)
  systematic_murray  GF/ext/CUDAExt.jl:49-58       O_j = min(n, floor(n W_j / W_n + u0))
  stratified         GF/ext/CUDAExt.jl:68-78       O_j = min(n, floor(r_j + u[min(n, floor(r_j + 1))]))
  multinomial        GF/ext/CUDAExt.jl:31-39       a_k = min{j : W_j >= u_k W_n}
  offspring->ancestors CUDAExt.jl:90-117 (slots past O_n are clamped to n-1)

Run:  python tests/golden/gen_resample_fixtures.py
"""
from __future__ import annotations

import json
import math
import os
from fractions import Fraction as F

import numpy as np

KINDS = {"systematic": 0, "systematic_murray": 1, "stratified": 2, "multinomial": 3}


def cumsum(w):
    out, s = [], F(0)
    for v in w:
        s += v
        out.append(s)
    return out


def o2a(O, n):
    anc, start = [], 0
    for i, fin in enumerate(O):
        fin = min(fin, n)
        anc.extend([i] * max(0, fin - start))
        start = max(start, fin)
    anc.extend([n - 1] * (n - len(anc)))
    return anc


def systematic(w, u):
    n = len(w)
    vs = [n * W for W in cumsum(w)]
    out = []
    for k in range(n):
        cands = [j for j in range(n) if vs[j] > u[0] + k]
        out.append(cands[0] if cands else n - 1)
    return out


def systematic_murray(w, u):
    n = len(w)
    W = cumsum(w)
    Wn = W[-1]
    O = [min(n, math.floor(n * Wj / Wn + u[0])) for Wj in W]
    return o2a(O, n)


def stratified(w, u):
    n = len(w)
    W = cumsum(w)
    Wn = W[-1]
    O = []
    for Wj in W:
        r = n * Wj / Wn
        k = min(n, math.floor(r + 1))
        O.append(min(n, math.floor(r + u[k - 1])))
    return o2a(O, n)


def multinomial(w, u):
    n = len(w)
    W = cumsum(w)
    Wn = W[-1]
    out = []
    for uk in u:
        q = uk * Wn
        cands = [j for j in range(n) if W[j] >= q]
        out.append(cands[0] if cands else n - 1)
    return out


FUN = {"systematic": systematic, "systematic_murray": systematic_murray,
       "stratified": stratified, "multinomial": multinomial}


def dyadic_weights(rng, n, bits, total_num=1, total_den=1):
    """n nonnegative dyadic weights k_i / 2^bits summing to total_num/total_den (a dyadic)."""
    scale = (2 ** bits) * total_num // total_den
    cuts = np.sort(rng.integers(0, scale + 1, size=n - 1))
    ks = np.diff(np.concatenate([[0], cuts, [scale]]))
    return [F(int(k), 2 ** bits) for k in ks]


def dyadic_uniforms(rng, m, bits):
    return [F(int(v), 2 ** bits) for v in rng.integers(0, 2 ** bits, size=m)]


def case(kind, w, u, tag):
    anc = FUN[kind](w, u)
    return {"kind": kind, "kind_id": KINDS[kind], "n": len(w), "tag": tag,
            "w": [float(v) for v in w], "u": [float(v) for v in u], "ancestors": anc}


def main():
    rng = np.random.default_rng(1234)
    cases = []
    # Hand-made ties and edges.
    quarter = [F(1, 4)] * 4
    for kind in KINDS:
        m = 1 if kind.startswith("systematic") else 4
        cases.append(case(kind, quarter, [F(0)] * m, "uniform-u0=0 (exact ties)"))
        cases.append(case(kind, quarter, [F(1, 2)] * m, "uniform-u=1/2"))
        cases.append(case(kind, [F(0), F(1), F(0), F(0)], [F(3, 8)] * m, "one-hot"))
        cases.append(case(kind, [F(1, 2), F(0), F(1, 2)], [F(1, 2)] * (1 if m == 1 else 3),
                          "tie-at-half"))
        cases.append(case(kind, [F(1)], [F(1, 4)], "n=1"))
        cases.append(case(kind, [F(3, 4), F(1, 4)], [F(3, 4)] * (1 if m == 1 else 2),
                          "two-particle"))
    # Systematic CPU form with sum(w) < 1: the reference's out-of-bounds edge, clamped.
    cases.append(case("systematic", [F(1, 8), F(1, 8), F(1, 4)], [F(7, 8)], "sum<1 clamp"))
    # Scan totals that are powers of two other than 1 (the GPU forms divide by W_n).
    for kind in ("systematic_murray", "stratified", "multinomial"):
        for n in (3, 5, 8):
            w = dyadic_weights(rng, n, 12, 2, 1)
            m = 1 if kind == "systematic_murray" else n
            cases.append(case(kind, w, dyadic_uniforms(rng, m, 10), "total=2"))
    # Random dyadic cases.
    for n in (2, 3, 5, 8, 16, 33, 64):
        for rep in range(3):
            w = dyadic_weights(rng, n, 16)
            for kind in KINDS:
                m = 1 if kind.startswith("systematic") else n
                u = dyadic_uniforms(rng, m, 12)
                cases.append(case(kind, w, u, f"random-{rep}"))
    # Known answers from the reference's own tests (test/components/resamplers.jl:76-92),
    # converted to 0-based indices.
    known = {
        "offspring_to_ancestors": {"offspring": [0, 2, 2, 3, 5], "ancestors": [1, 1, 3, 4, 4]},
        "ancestors_to_offspring": {"ancestors": [3, 1, 1, 2, 0], "offspring": [1, 2, 1, 1, 0]},
    }
    out = {"generator": os.path.basename(__file__), "cases": cases, "known": known}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "resample_fixtures.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0)
    print(f"wrote {len(cases)} cases to {path}")


if __name__ == "__main__":
    main()
