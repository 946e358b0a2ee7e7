"""Synthetic workload generator: deterministic, seeded, and geometrically consistent with the
solver's correspondence model (MI local pixel p in MI r maps to p + (c_r - c_n) * pitch / v in
neighbour n for a scene point at virtual depth v)."""
import numpy as np


def test_deterministic_and_seeded(lfpm):
    sp = lfpm.SynthParams(256, 200, 23.0, 4, 2.5, 6.0, 1234, 1235, 2.0)
    a, ga = lfpm.synth(sp)
    b, gb = lfpm.synth(sp)
    assert np.array_equal(a, b) and np.array_equal(ga, gb)
    sp2 = lfpm.SynthParams(256, 200, 23.0, 4, 2.5, 6.0, 1234, 999, 2.0)
    c, gc = lfpm.synth(sp2)
    assert not np.array_equal(a, c) and np.array_equal(ga, gc)  # noise seed: texture only


def test_depth_range_and_discs(lfpm):
    wl = lfpm.workload_c1(raw_w=512, raw_h=400)
    raw, gt = wl.raw(), wl.gt_virtual_depth()
    disc = gt > 0
    assert disc.mean() > 0.85  # hex discs of diameter = pitch cover ~90.7 % of the plane
    assert gt[disc].min() >= 2.5 - 1e-3 and gt[disc].max() <= 6.0 + 1e-3
    assert np.all(raw[~disc] == 0)


def test_ground_truth_consistent_with_correspondence_model(lfpm):
    """Raw intensities at GT-corresponding points of neighbouring MIs agree up to sensor noise."""
    wl = lfpm.workload_c1(raw_w=512, raw_h=400, noise_sigma=0.0)
    raw, gt = wl.raw().astype(np.float64), wl.gt_virtual_depth()
    L = wl.L
    W = wl.crop
    diffs = []
    for v in range(0, L["V"], 7):
        ox, oy = L["origins"][v]
        for k in range(L["cnt"][v]):
            n = L["nbr"][v, k]
            step = (L["centers"][v] - L["centers"][n]) / wl.L["pitch"]
            for (py, px) in [(W // 2, W // 2), (W // 2 - 2, W // 2 + 1)]:
                vv = gt[oy + py, ox + px]
                d = wl.L["pitch"] / vv
                qx, qy = px + step[0] * d, py + step[1] * d
                if not (0 <= qx < W - 1 and 0 <= qy < W - 1):
                    continue
                nx, ny = L["origins"][n]
                x0, y0 = int(qx), int(qy)
                fx, fy = qx - x0, qy - y0
                g = raw[ny + y0:ny + y0 + 2, nx + x0:nx + x0 + 2]
                s = (g[0, 0] * (1 - fx) + g[0, 1] * fx) * (1 - fy) + (g[1, 0] * (1 - fx) + g[1, 1] * fx) * fy
                diffs.append(abs(s - raw[oy + py, ox + px]))
    diffs = np.array(diffs)
    assert len(diffs) > 100
    # bilinear resampling of a smooth texture: small residual, far below texture contrast
    assert np.median(diffs) < 6.0
