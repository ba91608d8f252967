"""Writes tests/golden/golden.npz: small input/expected-output vectors computed
independently of the oracle and of the HIP library.

The reference repository ships no fixtures or golden vectors (its tests draw
unseeded random inputs, test/runtests.jl), so the expected values here come
from textbook evaluations of the reference's formulas:
  * imfilter (imfilter.jl:5-26)  -> scipy.signal.correlate(mode="valid")
  * fastdistance (utils.jl:5-13) -> sum_j w_j (TI - K)^2 in exact rational arithmetic
  * band pick (iqsim.jl:237)     -> numpy closed form
  * taumodel (taumodel.jl:5-45)  -> scipy dense ranks + exact rationals, rounded once
  * graphcut (graphcut.jl:5-84)  -> brute-force enumeration of every s-t cut
                                    (exact rationals), mask = union of the
                                    source sides of all minimum cuts
  * relaxation (relaxation.jl)   -> set formulation with numpy lexsort
Run: python tests/golden/make_golden.py
"""
import itertools
import os
from fractions import Fraction

import numpy as np
from scipy.signal import correlate
from scipy.stats import rankdata

HERE = os.path.dirname(os.path.abspath(__file__))
EPS = Fraction(2) ** -52


def cap_exact(A, B, u, v, w):
    """graphcut.jl:37-52 for Float64 slabs, evaluated as exact rationals of the
    Float64 inputs (the reference rounds each operation; these small integer
    cases round exactly except for the final division)."""
    Du = abs(Fraction(A[u]) - Fraction(B[u]))
    Dv = abs(Fraction(A[v]) - Fraction(B[v]))
    gAu = abs(Fraction(A[v]) - Fraction(A[u]))
    gBu = abs(Fraction(B[v]) - Fraction(B[u]))
    gAv = abs(Fraction(A[w]) - Fraction(A[v])) if w is not None else gAu
    gBv = abs(Fraction(B[w]) - Fraction(B[v])) if w is not None else gBu
    return (Du + Dv) / (gAu + gAv + gBu + gBv + EPS)


def brute_cut(A, B, dim):
    shape = A.shape
    idx = list(np.ndindex(*shape))
    edges = []
    for d in range(A.ndim):
        for c in idx:
            if c[d] + 1 >= shape[d]:
                continue
            v = tuple(c[k] + (1 if k == d else 0) for k in range(A.ndim))
            w = tuple(c[k] + (2 if k == d else 0) for k in range(A.ndim))
            w = w if w[d] < shape[d] else None
            edges.append((c, v, cap_exact(A, B, c, v, w)))
    L = shape[dim]
    free = [c for c in idx if 0 < c[dim] < L - 1]
    best, union = None, set()
    for bits in itertools.product([0, 1], repeat=len(free)):
        S = {c for c in idx if c[dim] == 0} | {c for c, b in zip(free, bits) if b}
        val = sum(cp for u, v, cp in edges if (u in S) != (v in S))
        if best is None or val < best:
            best, union = val, set(S)
        elif val == best:
            union |= S
    mask = np.zeros(shape, dtype=bool)
    for c in union:
        mask[c] = True
    return mask


def main():
    rng = np.random.default_rng(20261202)
    out = {}
    # imfilter
    img = rng.random((13, 9, 5))
    krn = rng.random((4, 3, 2))
    out["imf_img"], out["imf_krn"], out["imf_out"] = img, krn, correlate(img, krn, mode="valid")
    # fastdistance (exact rationals of the Float64 inputs, rounded once)
    ti = rng.integers(-3, 4, (10, 8)).astype(np.float64) / 4
    kern = rng.integers(-3, 4, (4, 3)).astype(np.float64) / 4
    w = rng.random((4, 3)) < 0.6
    ds = (7, 6)
    D = np.zeros(ds)
    for c in np.ndindex(*ds):
        s = Fraction(0)
        for j in np.ndindex(4, 3):
            if w[j]:
                t = Fraction(ti[c[0] + j[0], c[1] + j[1]]) - Fraction(kern[j])
                s += t * t
        D[c] = float(s)
    out["fd_ti"], out["fd_kern"], out["fd_w"], out["fd_out"] = ti, kern, w, D
    # band pick
    Db = np.round(rng.random(400) * 40) / 8
    Db[rng.random(400) < 0.1] = np.inf
    out["band_D"] = Db
    out["band_out"] = np.flatnonzero(Db <= (1 + 0.1) * Db.min())
    # taumodel
    n = 37
    D1 = np.round(rng.random(60) * 9)
    D2 = np.round(rng.random(60) * 5)
    ev = np.sort(rng.choice(60, n, replace=False))
    R = np.stack([rankdata(D1[ev], method="dense"), rankdata(D2[ev], method="dense")], axis=1)
    Pm = [[Fraction(n - int(r) + 1) for r in R[:, j]] for j in range(2)]
    Pm = [[p / sum(col) for p in col] for col in Pm]
    x0 = Fraction(n - 1)
    p = []
    for i in range(n):
        prod = Fraction(1)
        for j in range(2):
            X = (1 - Pm[j][i]) / Pm[j][i]
            prod *= X / x0
        p.append(float(1 / (1 + x0 * prod)))
    out["tau_D1"], out["tau_D2"], out["tau_ev"], out["tau_p"] = D1, D2, ev, np.asarray(p)
    # relaxation (set formulation)
    nrel = 300
    Dr = np.round(rng.random(nrel) * 20)
    Ar = np.round(rng.random(nrel) * 20)
    dis = rng.random(nrel) < 0.1
    Dr[dis] = np.inf
    Ar[dis] = np.inf
    tol = 0.1
    en = int((~dis).sum())
    dbsize = int(np.ceil(tol * en))
    top = lambda v, k: set(np.lexsort((np.arange(v.size), v))[:k].tolist())
    over = top(Dr, dbsize)
    frac = 0.1 * (dbsize / en)
    while True:
        res = over & top(Ar, int(np.ceil(frac * en)))
        if res:
            break
        frac = min(frac + 0.1, 1.0)
    out["rel_D"], out["rel_A"], out["rel_out"] = Dr, Ar, np.asarray(sorted(res))
    # graphcut brute force (Float64 integer-valued slabs: exact up to the final division)
    for k, (shape, dim) in enumerate([((4, 3), 0), ((3, 5), 1), ((4, 2, 2), 0), ((2, 3, 4), 2)]):
        A = rng.integers(0, 4, shape).astype(np.float64)
        B = rng.integers(0, 4, shape).astype(np.float64)
        out[f"gc{k}_A"], out[f"gc{k}_B"], out[f"gc{k}_dim"] = A, B, np.int64(dim)
        out[f"gc{k}_mask"] = brute_cut(A, B, dim)
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    print("wrote", os.path.join(HERE, "golden.npz"), len(out), "arrays")


if __name__ == "__main__":
    main()
