"""Pins of the oracle's per-frame products (CPU only).

Disparity mosaic (DepthSolver::WriteDisMap_TileKey_new_AccuSequence, DepthSolver.cpp:6716-6897):
the crop corner is c - (round(crop/2) - 1) truncated, which for lattice-snapped centres equals the
slicing origin c - (crop-1)/2 for even and odd crops; the cost gate is inclusive; d <= 0 is
dropped; the range covers every d > 0 regardless of cost; overlapping crops keep the later MI.

Virtual-depth map (DepthSolver::Virtual_depth_map_TileKeySequence, DepthSolver.cpp:7832-8000):
offsets trunc(x - ((crop-1)/2 + 1)) make local columns crop/2 and crop/2 + 1 coincide for even
crops (one MI lands 15 x 15 distinct positions for crop 16, four samples at its centre); the
median follows the reference's index rule (upper pair for even n, clamped for n <= 2), positions
on row or column 0 are dropped, and the colour fill probes p + s before p - s (a straight edge
survives the ten 5x5 median passes, so the filled gap column shows which side was probed first).
"""
import numpy as np


def _round(v):
    """std::round (half away from zero) for the positive positions used here."""
    return int(np.floor(v + 0.5))


def _plane(n, d):
    p = np.zeros((n, 4), np.float32)
    p[:, 2] = 1.0
    p[:, 3] = d
    return p


def test_mosaic_corner_matches_slice_origin(orc):
    for crop, c in ((16, 30.5), (21, 30.0)):
        n = crop * crop
        plane = _plane(n, np.arange(1, n + 1, dtype=np.float32))
        m, rng = orc.disparity_mosaic(plane, np.zeros(n, np.float32),
                                      np.array([[c, c]], np.float32), crop, 0.2, 64, 64)
        o = int(c - (crop - 1) / 2)
        assert np.array_equal(m[o:o + crop, o:o + crop].ravel(), plane[:, 3])
        assert m.sum() == plane[:, 3].sum()
        assert rng.tolist() == [1.0, float(n)]


def test_mosaic_cost_gate_and_invalid(orc):
    crop, n = 16, 256
    plane = _plane(n, 3.0)
    plane[7, 3] = 0.0
    plane[8, 3] = -2.0
    plane[9, 3] = 11.0
    cost = np.zeros(n, np.float32)
    cost[1] = np.float32(0.2)
    cost[2] = np.nextafter(np.float32(0.2), np.float32(1))
    cost[9] = 1.5  # gated out of the map but still in the range
    m, rng = orc.disparity_mosaic(plane, cost, np.array([[20.5, 20.5]], np.float32), crop, 0.2,
                                  48, 48)
    flat = m[13:29, 13:29].ravel()
    assert flat[1] == 3.0 and flat[2] == 0.0
    assert flat[7] == 0.0 and flat[8] == 0.0 and flat[9] == 0.0
    assert rng.tolist() == [3.0, 11.0]


def test_mosaic_overlap_later_mi_wins(orc):
    crop, n = 16, 256
    plane = np.concatenate([_plane(n, 2.0), _plane(n, 5.0)])
    centers = np.array([[20.5, 20.5], [24.5, 20.5]], np.float32)
    m, _ = orc.disparity_mosaic(plane, np.zeros(2 * n, np.float32), centers, crop, 0.2, 64, 48)
    assert m[20, 15] == 2.0 and m[20, 17] == 5.0 and m[20, 32] == 5.0


def test_vd_constant_disparity_positions(orc):
    crop, n = 16, 256
    plane = _plane(n, 4.0)  # f_v = 16 / 4 = 4
    vd, _, rng = orc.vd_map(plane, np.array([[100.5, 100.5]], np.float32), None, crop, 16.0,
                            200, 200, color=False)
    offs = [int(x - 8.5) for x in range(16)]  # trunc toward zero: -8..-1, 0, 0, 1..6
    assert offs[8] == 0 and offs[9] == 0 and offs[0] == -8 and offs[15] == 6
    want = np.zeros_like(vd)
    for oy in set(offs):
        for ox in set(offs):
            want[_round(100.5 + 4 * oy), _round(100.5 + 4 * ox)] = 4.0
    assert np.count_nonzero(vd) == 15 * 15
    assert np.array_equal(vd, want)
    assert rng.tolist() == [4.0, 4.0]


def test_vd_median_index_rule(orc):
    crop, n = 16, 256
    plane = _plane(n, 16.0)
    # the four pixels with zero offset land on the centre: f_v = 16/d = {1, 2, 3, 5}
    for (x, y), fv in zip(((8, 8), (9, 8), (8, 9), (9, 9)), (3.0, 1.0, 5.0, 2.0)):
        plane[y * crop + x, 3] = 16.0 / fv
    vd, _, _ = orc.vd_map(plane, np.array([[50.0, 50.0]], np.float32), None, crop, 16.0, 100, 100,
                          color=False)
    assert vd[50, 50] == (5.0 + 3.0) / 2  # s[n/2+1], s[(n+1)/2] of the sorted samples
    # two identical MIs: every position holds n = 2 samples -> clamped rule takes s[1] (max)
    p2 = np.concatenate([_plane(n, 8.0), _plane(n, 4.0)])
    vd2, _, _ = orc.vd_map(p2, np.array([[50.0, 50.0], [50.0, 50.0]], np.float32), None, crop,
                           16.0, 100, 100, color=False)
    assert vd2[50, 50] == 4.0  # centre: 8 samples {2 x4, 4 x4} -> (s[5] + s[4]) / 2 = 4
    # (42, 42): offset -4 at f_v 2 and offset -2 at f_v 4 -> n = 2 {2, 4} -> clamped s[1] = 4
    assert vd2[42, 42] == 4.0
    # d <= 0 counts as d = 1
    p3 = _plane(n, 0.0)
    vd3, _, _ = orc.vd_map(p3, np.array([[50.0, 50.0]], np.float32), None, crop, 2.0, 100, 100,
                           color=False)
    assert vd3[50, 50] == 2.0


def test_vd_border_positions_dropped(orc):
    crop, n = 17, 289  # odd crop: offsets x - 9 exactly, f_v = 1
    plane = _plane(n, 4.0)
    vd, _, _ = orc.vd_map(plane, np.array([[9.0, 30.0]], np.float32), None, crop, 4.0, 64, 64,
                          color=False)
    assert vd[30, 0] == 0.0 and vd[30, 1] == 1.0  # X = 0 is outside (X > 0 required)


def test_vd_fill_probes_right_before_left(orc):
    crop, n = 16, 256
    # two MIs at f_v = 1, their 15-wide blocks one empty column apart
    plane = np.concatenate([_plane(n, 4.0), _plane(n, 4.0)])
    centers = np.array([[40.5, 40.5], [56.5, 40.5]], np.float32)
    gray = np.concatenate([np.full(n, 50, np.float32), np.full(n, 200, np.float32)])
    vd, col, _ = orc.vd_map(plane, centers, gray, crop, 4.0, 96, 80)
    blk0 = sorted({_round(40.5 + int(x - 8.5)) for x in range(16)})
    blk1 = sorted({_round(56.5 + int(x - 8.5)) for x in range(16)})
    gap = blk0[-1] + 1
    assert blk1[0] == gap + 1 and np.count_nonzero(vd[40, :]) == 30
    assert col[40, gap] == 200 and col[40, gap - 1] == 50 and col[40, gap + 1] == 200


def _solved_like_plane(V, crop, seed):
    rng = np.random.default_rng(seed)
    n = V * crop * crop
    p = np.zeros((n, 4), np.float32)
    p[:, 2] = 1.0
    p[:, 3] = rng.uniform(2.5, 9.0, n).astype(np.float32)
    return p


def test_vd_bgr_channels_are_independent(orc, lfpm):
    """The colour VD map treats B, G and R as three independent gray images (the reference sorts
    color0/1/2 separately, DepthSolver.cpp:7919-7927, and medianBlur filters each channel): its
    channel c equals the gray map run on channel c alone, and the VD values are the gray run's.
    The channels are chosen so that sorting them together (one sample's colour for all three)
    would differ."""
    wl = lfpm.workload_c1(raw_w=256, raw_h=200)
    crop, V = wl.crop, wl.V
    plane = _solved_like_plane(V, crop, 5)
    rng = np.random.default_rng(9)
    bgr = rng.integers(0, 256, (V, crop, crop, 3)).astype(np.uint8)
    vd, col, r = orc.vd_map_bgr(plane, wl.L["centers"], bgr, crop, wl.L["pitch"], 256, 200)
    assert col.shape == (200, 256, 3)
    for c in range(3):
        vdc, colc, rc = orc.vd_map(plane, wl.L["centers"], bgr[..., c].astype(np.float32).ravel(),
                                   crop, wl.L["pitch"], 256, 200)
        assert np.array_equal(vdc.view(np.uint32), vd.view(np.uint32))
        assert np.array_equal(rc, r)
        assert np.array_equal(colc, col[..., c]), c
    # co-sorted colours would make the three channels one sample's colour; here they differ
    assert (col[..., 0] != col[..., 1]).mean() > 0.5


def test_vd_bgr_point_colours(orc):
    """Point records take red, green, blue from channels 2, 1, 0 of the B, G, R image (the
    writer's r/g/b = color.z/.y/.x, MVStereo/ACMP.cpp:479-481)."""
    vd = np.zeros((4, 5), np.float32)
    vd[1, 2] = 3.0
    vd[3, 4] = 5.0
    col = np.zeros((4, 5, 3), np.uint8)
    col[1, 2] = (10, 20, 30)
    col[3, 4] = (200, 150, 100)
    rec = orc.vd_point_cloud(vd, col, (2.0, 2.0), 1.0, 50.0, 1.0, 0.0)
    assert rec.shape == (2, 27)
    assert rec[0, 24:].tolist() == [30, 20, 10]
    assert rec[1, 24:].tolist() == [100, 150, 200]
