"""Pins of the oracle's ACMM multi-scale variant (AdaptMIPMFrameACMM.{cpp,cu}).

Level schedule (AdaptMIPMFrameACMM.cpp:155-200, 300-345): two levels only for crops of at least
24 px.  INTER_AREA level images: constant and linear tiles (the area mean of a linear function
is its value at the cell centre), exact 2x2 block means for even crops.  Solve: every output
pixel's cost and selection are its plane's own top-k support (the metadata re-alignment that
ends every ACMM update and the post-refine pass), checked against the oracle's independent
cost-vector and top-k functions, and the solve converges to the analytic ground truth on one- and
two-level frames.
"""
import numpy as np
import pytest

from _helpers import pm_to_oracle


def test_level_schedule(orc):
    one = orc.acmm_levels(16)
    assert len(one) == 1 and one[0] == dict(width=16, scale=1.0, radius=2, iterations=5,
                                             top_k=6)
    two = orc.acmm_levels(24)
    assert len(two) == 2
    assert two[0] == dict(width=12, scale=0.5, radius=3, iterations=3, top_k=6)
    assert two[1] == dict(width=24, scale=1.0, radius=2, iterations=6, top_k=5)
    assert orc.acmm_levels(25)[0]["width"] == 13  # round(12.5) away from zero
    assert len(orc.acmm_levels(23)) == 1


def test_area_resize_closed_forms(orc):
    rng = np.random.default_rng(2)
    t = rng.random((16, 16))
    b = orc.area_resize(t, 8)
    want = t.reshape(8, 2, 8, 2).mean(axis=(1, 3))
    assert np.allclose(b, want, rtol=0, atol=1e-15)
    for W, W0 in ((25, 13), (27, 14), (24, 12)):
        assert np.allclose(orc.area_resize(np.full((W, W), 0.37), W0), 0.37, atol=1e-15)
        # separable tile g(x) h(y): the area mean of a pixel-wise constant image over a cell
        # is (G(b) - G(a)) / (b - a) per axis, with G the running integral of the steps
        g, h = rng.random(W), rng.random(W)
        S = W / W0

        def cell_means(f):
            F = np.concatenate([[0.0], np.cumsum(f)])

            def integ(x):
                i = min(int(np.floor(x)), W - 1)
                return F[i] + (x - i) * f[i] if x < W else F[W]
            return np.array([(integ(min((d + 1) * S, W)) - integ(d * S)) /
                             (min((d + 1) * S, W) - d * S) for d in range(W0)])
        want = np.outer(cell_means(h), cell_means(g))
        assert np.allclose(orc.area_resize(np.outer(h, g), W0), want, atol=1e-12)


def _support_check(orc, wl, tiles, out, pm):
    """cost / sel of every pixel == top-k support of its own plane (FP64 identity)."""
    V, W = wl.V, pm.crop
    rng = np.random.default_rng(4)
    idx = rng.choice(V * W * W, 400, replace=False)
    hyp = np.zeros((len(idx), 6), np.int32)
    hyp[:, 0] = idx // (W * W)
    hyp[:, 1] = idx % (W * W)
    hyp[:, 2:] = out["plane"][idx].astype(np.float32).view(np.int32)
    order = np.argsort(hyp[:, 0], kind="stable")
    hyp, idx = hyp[order], idx[order]
    cv = orc.eval_costs(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"], pm, hyp)
    mism = 0
    for r, i in enumerate(idx):
        c = int(wl.L["cnt"][hyp[r, 0]])
        mask, mean, _ = orc.topk(cv[r, :c], pm.top_k)
        mism += mask != out["sel"][i] or abs(mean - out["cost"][i]) > 1e-6
    return mism


def _gt_err(wl, out):
    gt_v = wl.gt_virtual_depth()
    o = wl.L["origins"]
    W = wl.crop
    errs = []
    for v in range(wl.V):
        g = gt_v[o[v, 1]:o[v, 1] + W, o[v, 0]:o[v, 0] + W].reshape(-1)
        d = out["plane"][v * W * W:(v + 1) * W * W, 3]
        ok = g > 0
        errs.append(np.abs(d[ok] - wl.L["pitch"] / g[ok]))
    return np.concatenate(errs)


@pytest.mark.parametrize("pitch", [23.0, 35.0], ids=["one_level", "two_levels"])
def test_acmm_solve(orc, lfpm, pitch):
    wl = lfpm.workload_c1(raw_w=400 if pitch < 30 else 560, raw_h=320 if pitch < 30 else 448,
                          pitch=pitch)
    tiles = orc.slice_u8(wl.raw(), wl.L["origins"], wl.crop)
    pm = pm_to_oracle(orc, wl.pm)
    out = orc.solve_acmm(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"], pm)
    levels = orc.acmm_levels(wl.crop)
    assert len(levels) == (2 if pitch > 30 else 1)
    # metadata of the final level: the reference's top-k of the final level
    pm.top_k = levels[-1]["top_k"]
    pm.radius = levels[-1]["radius"]
    assert _support_check(orc, wl, tiles, out, pm) == 0
    err = _gt_err(wl, out)
    assert np.median(err) < 0.25, np.median(err)


def test_acmm_blur_term(orc, lfpm):
    """ACMM with the blur term (lambda 0.8 as the ACMM twin, AdaptMIPMFrameACMM.cu:165) on the
    two-level schedule: equal blur maps everywhere leave every blur sum difference at zero, so
    the solve is the one without the term; the frame's defocus maps change it."""
    import copy
    wl = lfpm.workload_c1(raw_w=560, raw_h=448, pitch=35.0)
    tiles = orc.slice_u8(wl.raw(), wl.L["origins"], wl.crop)
    pm = pm_to_oracle(orc, wl.pm)
    base = orc.solve_acmm(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"], pm)
    pmb = copy.copy(pm)
    pmb.blur_lambda = 0.8
    flat = np.full(tiles.shape, 0.3, np.float64)
    same = orc.solve_acmm(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"], pmb, blur=flat)
    for k in ("plane", "cost", "sel"):
        assert np.allclose(same[k], base[k], rtol=0, atol=1e-12), k
    blur = orc.slice_u8(wl.blur_map(), wl.L["origins"], wl.crop).astype(np.float64) / 255.0
    out = orc.solve_acmm(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"], pmb, blur=blur)
    assert not np.array_equal(out["plane"], base["plane"])
