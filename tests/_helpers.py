"""Helpers shared by the test modules (conversion between product and oracle structs)."""
import numpy as np


def pm_to_oracle(orc, pm):
    return orc.PmParams(*[getattr(pm, f[0]) for f in orc.PmParams._fields_])


def fp_to_oracle(orc, fp):
    return orc.FilterParams(*[getattr(fp, f[0]) for f in orc.FilterParams._fields_])


def cone_hypotheses(lfpm, wl, n, seed, vmin=2.0, vmax=8.0, cone_deg=60.0):
    return lfpm.cone_hypotheses(wl, n, seed, vmin, vmax, cone_deg)


def rel_err(c, ref, floor=1e-3):
    return np.abs(c - ref) / np.maximum(np.abs(ref), floor)
