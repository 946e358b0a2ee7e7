"""Helpers shared by the test modules (conversion between product and oracle structs)."""
import os

import numpy as np


def pm_to_oracle(orc, pm):
    return orc.PmParams(*[getattr(pm, f[0]) for f in orc.PmParams._fields_])


def fp_to_oracle(orc, fp):
    return orc.FilterParams(*[getattr(fp, f[0]) for f in orc.FilterParams._fields_])


def cone_hypotheses(lfpm, wl, n, seed, vmin=2.0, vmax=8.0, cone_deg=60.0):
    return lfpm.cone_hypotheses(wl, n, seed, vmin, vmax, cone_deg)


def rel_err(c, ref, floor=1e-3):
    return np.abs(c - ref) / np.maximum(np.abs(ref), floor)


# Exact fixed-plane cost bar.  An evaluation (hypothesis, view) may differ from the FP64 oracle
# by more than the tolerance only where FP32 arithmetic cannot decide it:
#  - a straddle: the anchor or a window sample's source point lies within GEO_STRADDLE px of a
#    crop edge (FP32 rounding of the source point, a few 1e-6 px, can move it across), or a
#    window variance lies within its FP32 error estimate of the variance gate;
#  - ill-conditioned: the oracle's first-order worst-case FP32 error estimate of the cost
#    exceeds ILL_FACTOR times the tolerance (the estimate is a bound: realised FP32 errors are
#    below 0.15 of it on every evaluation measured, and about 0.003 of it typically).
GEO_STRADDLE = 1e-4
ILL_FACTOR = 10.0


def cost_bar(out, ref, geo, gate, ferr, tol, floor=1e-3):
    """Classify GPU costs `out` against the FP64 oracle `ref` (both (n, K), 2 = not scored) at
    relative tolerance `tol`.  Returns a dict of statistics; 'unexplained' counts the evaluations
    outside the tolerance that are neither straddles nor ill-conditioned (the bar: 0)."""
    out = np.asarray(out, np.float64)
    ref = np.asarray(ref, np.float64)
    tol_abs = tol * np.maximum(np.abs(ref), floor)
    err = np.abs(out - ref)
    straddle = (geo < GEO_STRADDLE) | (gate < 1.0)
    ill = ferr >= ILL_FACTOR * tol_abs
    viol = err > tol_abs
    scored = (ref < 2) | (out < 2)
    both = (ref < 2) & (out < 2)
    e = err[both] / np.maximum(np.abs(ref[both]), floor)
    unexpl = viol & ~straddle & ~ill
    return dict(n_scored=int(scored.sum()), n_viol=int(viol.sum()),
                n_viol_straddle=int((viol & straddle).sum()),
                n_viol_ill=int((viol & ~straddle & ill).sum()),
                unexplained=int(unexpl.sum()),
                excused_frac=float((straddle | ill)[scored].mean()),
                flips=int(((ref < 2) != (out < 2)).sum()),
                frac_le_tol=float((e <= tol).mean()), max_rel=float(e.max()),
                max_rel_unexcused=float((err / np.maximum(np.abs(ref), floor))[
                    scored & ~straddle & ~ill].max()))


ORACLE_SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                          "oracle", "lfpm_oracle.cpp")


def build_oracle_mutant(i, edits, tmp):
    """Compile the oracle source with textual edits [(old, new), ...] (each old text must occur
    exactly once) into tmp/mut<i>.so and return its path."""
    import subprocess
    s = open(ORACLE_SRC).read()
    for old, new in edits:
        assert s.count(old) == 1, (i, old, s.count(old))
        s = s.replace(old, new)
    src = os.path.join(tmp, "mut%d.cpp" % i)
    so = os.path.join(tmp, "mut%d.so" % i)
    with open(src, "w") as f:
        f.write(s)
    subprocess.check_call(["g++", "-O0", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off",
                           "-shared", "-o", so, src])
    return so
