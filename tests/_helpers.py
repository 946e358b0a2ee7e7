"""Helpers shared by the test modules (conversion between product and oracle structs)."""
import numpy as np


def pm_to_oracle(orc, pm):
    return orc.PmParams(*[getattr(pm, f[0]) for f in orc.PmParams._fields_])


def fp_to_oracle(orc, fp):
    return orc.FilterParams(*[getattr(fp, f[0]) for f in orc.FilterParams._fields_])


def cone_hypotheses(lfpm, wl, n, seed, vmin=2.0, vmax=8.0, cone_deg=60.0):
    """n seeded (MI, pixel, plane) records sorted by MI: normals uniform in the cone of
    `cone_deg` about +z, virtual depth uniform in [vmin, vmax] -> disparity pitch/v."""
    rng = np.random.default_rng(seed)
    V, W = wl.V, wl.crop
    mi = np.sort(rng.integers(0, V, n)).astype(np.int32)
    pix = rng.integers(0, W * W, n).astype(np.int32)
    cz = 1.0 - rng.random(n) * (1.0 - np.cos(np.radians(cone_deg)))
    phi = 2 * np.pi * rng.random(n)
    sz = np.sqrt(np.maximum(0.0, 1.0 - cz * cz))
    v = vmin + (vmax - vmin) * rng.random(n)
    planes = np.stack([sz * np.cos(phi), sz * np.sin(phi), cz, wl.L["pitch"] / v], 1)
    return lfpm.hypotheses_from_arrays(mi, pix, planes.astype(np.float32))


def rel_err(c, ref, floor=1e-3):
    return np.abs(c - ref) / np.maximum(np.abs(ref), floor)
