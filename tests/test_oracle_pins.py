"""Pins of the CPU oracle (and of the product's host-side tables) against known answers,
closed forms and independent implementations — not against a retyped copy of its formulas.

* Philox4x32-10: Random123 known-answer vectors (Salmon et al., SC'11 distribution).
* Lens layout: product and oracle builders agree; ring distances and counts from hex geometry;
  neighbour sets agree with a KD-tree radius query.
* Bilateral NCC: closed-form cases (identity, affine, inverted, constant, anchor outside, slant
  gate) and an independent implementation (scipy map_coordinates bilinear + numpy weighted
  covariance) on random tiles and planes.
* Top-k / weighted cost / cross-view re-encoding: hand-computed values.
* Marsaglia normals: unit norm and uniform-sphere moments.
"""
import math

import numpy as np
import pytest

from _helpers import pm_to_oracle

KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_kat_oracle(orc, ctr, key, want):
    assert orc.philox(ctr, key) == want


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_kat_product(lfpm, ctr, key, want):
    assert lfpm.philox(ctr, key) == want


LAYOUTS = [(23.0, 1024, 800, 6, 1, (6, 6, 6)), (23.0, 2048, 1536, 18, 1, (18, 18, 18)),
           (15.0, 1024, 800, 18, 3, (6, 12, 18)), (31.0, 1024, 800, 18, 3, (6, 12, 18)),
           (23.0, 1024, 800, 12, 1, (12, 12, 12))]


@pytest.mark.parametrize("cfg", LAYOUTS)
def test_layout_product_matches_oracle(lfpm, orc, cfg):
    pitch, w, h, K, lt, per = cfg
    L = lfpm.layout(lfpm.layout_params(pitch, w, h, K, lens_types=lt, per_type=per))
    O = orc.layout(orc.LayoutParams(pitch, 0, w, h, K, lt, per))
    for k in ("centers", "origins", "tilekeys", "lens_type", "nbr", "cnt"):
        assert np.array_equal(L[k], O[k]), k


def test_layout_hex_geometry(lfpm):
    from scipy.spatial import cKDTree
    pitch = 23.0
    L = lfpm.layout(lfpm.layout_params(pitch, 2048, 1536, 18))
    c = L["centers"].astype(np.float64)
    crop = L["crop"]
    assert crop == int(pitch * 0.707106781186)
    # crop origins are exact integers: c - (crop-1)/2
    assert np.array_equal(c - (crop - 1) / 2.0, L["origins"].astype(np.float64))
    # ring distances (each centre snapped to the crop grid moves by <= 0.5 px per axis, so a
    # pair distance moves by <= sqrt(2))
    for v in range(0, L["V"], 97):
        n = L["cnt"][v]
        d = np.hypot(*(c[L["nbr"][v, :n]] - c[v]).T)
        ring1, ring2a, ring2b = d[:min(n, 6)], d[6:12], d[12:18]
        if n == 18:
            assert np.all(np.abs(ring1 - pitch) <= 1.42)
            assert np.all(np.abs(ring2a - math.sqrt(3) * pitch) <= 1.42)
            assert np.all(np.abs(ring2b - 2 * pitch) <= 1.42)
    # interior MIs have exactly 18 neighbours; neighbour sets equal a KD-tree radius query
    tree = cKDTree(c)
    for v in range(0, L["V"], 53):
        got = set(L["nbr"][v, :L["cnt"][v]].tolist())
        want = set(tree.query_ball_point(c[v], 2.0 * pitch + 1.5)) - {v}
        assert got == want
    interior = (c[:, 0] > 3 * pitch) & (c[:, 0] < 2048 - 3 * pitch) & \
               (c[:, 1] > 3 * pitch) & (c[:, 1] < 1536 - 3 * pitch)
    assert np.all(L["cnt"][interior] == 18)


def test_lens_types_three_colouring(lfpm):
    L = lfpm.layout(lfpm.layout_params(15.0, 1024, 800, 18, lens_types=3, per_type=(6, 12, 18)))
    t = L["lens_type"]
    c = L["centers"].astype(np.float64)
    # ring-1 neighbours (distance ~ pitch) of every MI are of the other two types
    for v in range(L["V"]):
        nn = L["nbr"][v, :L["cnt"][v]]
        ring1 = nn[np.hypot(*(c[nn] - c[v]).T) < 1.5 * 15.0]
        assert np.all(t[ring1] != t[v])
    assert set(np.unique(L["cnt"][t == 0])) <= set(range(0, 7))


# ----------------------------------------------------------------------------------------
# Bilateral NCC
# ----------------------------------------------------------------------------------------
def _two_mi_frame(orc, ref, nei, shift=(23.0, 0.0), K=1):
    """Two MIs: 0 = reference, 1 = its single neighbour at centre offset `shift` (c0 - c1)."""
    W = ref.shape[0]
    tiles = np.stack([ref, nei]).astype(np.uint8)
    centers = np.array([[100.0, 100.0], [100.0 - shift[0], 100.0 - shift[1]]], np.float32)
    nbr = np.array([[1], [0]], np.int32)
    cnt = np.array([1, 1], np.int32)
    return tiles, centers, nbr, cnt


def _pm(orc, W, r=3, sc=0.1, ss=3.0, cos=0.5):
    return orc.PmParams(W, 23.0, r, 1, ss, sc, 1e-5 / 255.0 ** 2, 5, 0.0, 11.5, 1, 1, 1234, 0,
                        0, cos)


def _hyp(lfpm, pix, plane):
    return lfpm.hypotheses_from_arrays(np.array([0]), np.array([pix]),
                                       np.array([plane], np.float32))


def test_ncc_closed_forms(orc, lfpm):
    rng = np.random.default_rng(0)
    W = 16
    ref = rng.integers(20, 100, (W, W))
    pm = _pm(orc, W)
    h0 = _hyp(lfpm, 8 * W + 8, [0, 0, 1, 0.0])  # zero disparity: source = same pixel
    # identical tiles -> 0
    c = orc.eval_costs(*_two_mi_frame(orc, ref, ref), pm, h0)
    assert c[0, 0] == 0.0
    # positive affine map of intensities -> 0
    c = orc.eval_costs(*_two_mi_frame(orc, ref, 2 * ref + 10), pm, h0)
    assert abs(c[0, 0]) < 1e-12
    # negative affine map -> 2 (correlation -1)
    c = orc.eval_costs(*_two_mi_frame(orc, ref, 250 - 2 * ref), pm, h0)
    assert abs(c[0, 0] - 2.0) < 1e-12
    # constant reference window -> variance gate -> 2
    c = orc.eval_costs(*_two_mi_frame(orc, np.full((W, W), 77), ref), pm, h0)
    assert c[0, 0] == 2.0
    # anchor outside the neighbour crop -> 2 (step (1,0) * d=9 from x=8)
    c = orc.eval_costs(*_two_mi_frame(orc, ref, ref), pm, _hyp(lfpm, 8 * W + 8, [0, 0, 1, 9.0]))
    assert c[0, 0] == 2.0
    # slant gate: nz = 0.4 < 0.5 -> 2; without the gate the plane is scored
    steep = [math.sqrt(1 - 0.16), 0.0, 0.4, 0.0]
    c = orc.eval_costs(*_two_mi_frame(orc, ref, ref), pm, _hyp(lfpm, 8 * W + 8, steep))
    assert c[0, 0] == 2.0
    c = orc.eval_costs(*_two_mi_frame(orc, ref, ref), _pm(orc, W, cos=0.0),
                       _hyp(lfpm, 8 * W + 8, steep))
    assert c[0, 0] < 2.0


def _independent_ncc(ref, nei, px, py, plane, step, r, ss, sc, min_var):
    """Bilateral NCC built from scipy's bilinear sampler and numpy's weighted covariance."""
    from scipy.ndimage import map_coordinates
    W = ref.shape[0]
    nx, ny, nz, w = plane
    A, B = -nx / nz, -ny / nz
    refn, nein = ref / 255.0, nei / 255.0
    ax, ay = px + step[0] * w, py + step[1] * w
    if not (0 <= ax < W and 0 <= ay < W):
        return 2.0
    ii, jj = np.meshgrid(np.arange(-r, r + 1), np.arange(-r, r + 1), indexing="ij")
    rx, ry = px + ii, py + jj
    inside = (rx >= 0) & (rx < W) & (ry >= 0) & (ry < W)
    d = w + A * ii + B * jj
    qx, qy = rx + step[0] * d, ry + step[1] * d
    ok = inside & (qx >= 0) & (qx < W) & (qy >= 0) & (qy < W)
    rv = refn[ry[ok], rx[ok]]
    sv = map_coordinates(nein, [qy[ok], qx[ok]], order=1, mode="nearest")
    wt = np.exp(-np.hypot(ii[ok], jj[ok]) / (2 * ss * ss) - np.abs(rv - refn[py, px]) / (2 * sc * sc))
    if wt.sum() <= 1e-6:
        return 2.0
    cov = np.cov(np.stack([rv, sv]), aweights=wt, bias=True)
    if cov[0, 0] < min_var or cov[1, 1] < min_var:
        return 2.0
    return float(np.clip(1 - cov[0, 1] / math.sqrt(cov[0, 0] * cov[1, 1]), 0, 2))


def test_ncc_matches_independent_implementation(orc, lfpm):
    rng = np.random.default_rng(1)
    W, r = 16, 4
    ref = rng.integers(0, 256, (W, W))
    nei = np.clip(ref + rng.integers(-30, 30, (W, W)), 0, 255)
    pm = _pm(orc, W, r=r, sc=0.15, ss=2.5)
    step = (23.0 / 23.0, 0.0)
    n = 300
    planes, pix = [], []
    for _ in range(n):
        th = math.radians(50) * rng.random()
        ph = 2 * math.pi * rng.random()
        planes.append([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th),
                       6 * rng.random()])
        pix.append(int(rng.integers(0, W * W)))
    hyp = lfpm.hypotheses_from_arrays(np.zeros(n, np.int32), np.array(pix),
                                      np.array(planes, np.float32))
    got = orc.eval_costs(*_two_mi_frame(orc, ref, nei), pm, hyp)[:, 0]
    pl32 = np.array(planes, np.float32).astype(np.float64)
    # parameters as the oracle receives them (float32 fields)
    sc, mv = float(np.float32(0.15)), float(np.float32(1e-5 / 255.0 ** 2))
    want = [_independent_ncc(ref, nei, p % W, p // W, pl32[i], step, r, 2.5, sc, mv)
            for i, p in enumerate(pix)]
    assert np.max(np.abs(got - np.array(want))) < 1e-9
    assert (got < 2).sum() > n // 3


def test_topk_hand_values(orc):
    m, mean, best = orc.topk([0.5, 0.2, 2.0, 0.2, 0.9, 0.1], 3)
    assert best == 5 and abs(mean - (0.1 + 0.2 + 0.2) / 3) < 1e-15
    assert m == 0b101010  # costs <= 3rd smallest (0.2): indices 1, 3, 5
    m, mean, best = orc.topk([2.0, 2.0], 5)
    assert (m, mean, best) == (0, 2.0, -1)
    m, mean, best = orc.topk([0.3, 2.0, 0.4], 5)  # fewer valid views than top_k
    assert m == 0b101 and abs(mean - 0.35) < 1e-15 and best == 0


def test_weighted_cost_hand_values(orc):
    c, am = orc.weighted([0.5, 0.2, 1.0], [2, 0, 1])
    assert abs(c - (2 * 0.5 + 1.0) / 3) < 1e-15 and am == 1
    c, am = orc.weighted([2.0, 2.0], [0, 0])
    assert c == 2.0 and am == -1


def test_reencode_fronto_parallel_and_slope(orc):
    # a fronto-parallel neighbour plane keeps its disparity at every reference pixel
    o = orc.reencode_plane([0, 0, 1, 5.25], (3, 4), 23.0, 0.0, (9, 2))
    assert np.allclose(o, [0, 0, 1, 5.25], atol=1e-15)
    # slanted: alpha = 0.01 -> reference slope alpha / (1 + alpha*du) (reference formula)
    n = np.array([-0.01, 0.0, 1.0])
    n /= np.linalg.norm(n)
    o = orc.reencode_plane([*n, 5.0], (3, 4), 23.0, 0.0, (3, 4))
    den = 1 + 0.01 * 23.0
    assert abs(-o[0] / o[2] - 0.01 / den) < 1e-12
    assert abs(o[3] - (0.01 * 3 + (5.0 - 0.01 * 3)) / den) < 1e-12


def test_marsaglia_normals_uniform_sphere(orc):
    n = orc.random_normals(1234, 0, 0, 200000)
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-12)
    assert np.all(np.abs(n.mean(0)) < 0.01)
    assert np.all(np.abs((n ** 2).mean(0) - 1.0 / 3.0) < 0.01)


def test_oracle_solve_thread_independent(lfpm, orc):
    wl = lfpm.workload_c1(raw_w=200, raw_h=160, iterations=2)
    tiles = orc.slice_u8(wl.raw(), wl.L["origins"], wl.crop)
    pm = pm_to_oracle(orc, wl.pm)
    a = orc.solve(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"], pm, threads=1)
    b = orc.solve(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"], pm, threads=4)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_oracle_solve_recovers_ground_truth(lfpm, orc):
    """On a textured synthetic frame the restated solver converges to the analytic disparity
    (pitch / virtual depth) — it is a working PatchMatch, not only a self-consistent one."""
    wl = lfpm.workload_c1(raw_w=400, raw_h=320)
    tiles = orc.slice_u8(wl.raw(), wl.L["origins"], wl.crop)
    o = orc.solve(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"], pm_to_oracle(orc, wl.pm))
    W = wl.crop
    gt = wl.gt_virtual_depth()
    gtv = np.stack([gt[y:y + W, x:x + W] for x, y in wl.L["origins"]]).reshape(-1)
    err = np.abs(o["plane"][:, 3] - wl.L["pitch"] / gtv)
    assert np.median(err) < 0.2
    assert (err < 0.5).mean() > 0.9


@pytest.mark.parametrize("shift,pix,d,want", [
    ((23.0, 0.0), (8, 8), 0.25, 4.75),     # samples x 5..11 -> 5.25..11.25: nearest edge 16
    ((23.0, 0.0), (2, 8), 0.25, 0.25),     # window clipped at x = 0: sample 0.25 from the edge
    ((23.0, 0.0), (2, 8), -0.5, 0.5),      # sample at -0.5 (outside by 0.5), next at 0.5
    ((0.0, 23.0), (8, 8), 0.25, 4.75),     # vertical step: x is exact, y decides
    ((0.0, 23.0), (2, 13), 0.25, 0.75),    # y samples 10..15 -> 10.25..15.25; x exact
])
def test_eval_margin_geo_closed_form(orc, lfpm, shift, pix, d, want):
    """The straddle margin of a fronto-parallel hypothesis: the smallest distance of the anchor
    or a window sample's source point (x + step * d) from a change of its inside-the-crop test,
    axes with a zero step component not counted."""
    rng = np.random.default_rng(3)
    W = 16
    ref = rng.integers(20, 200, (W, W))
    pm = _pm(orc, W)
    h = _hyp(lfpm, pix[1] * W + pix[0], [0, 0, 1, d])
    _, geo, _, _ = orc.eval_margins(*_two_mi_frame(orc, ref, ref, shift=shift), pm, h)
    assert abs(geo[0, 0] - want) < 1e-12


def test_eval_margin_ferr_bounds_fp32(orc, lfpm):
    """The FP32 error estimate bounds the FP32 twin's deviation from the FP64 cost on every
    evaluation that is not a straddle, and is not vacuous (the median ratio is above 1e-4)."""
    wl = lfpm.workload_c2(raw_w=320, raw_h=256)
    from _helpers import cone_hypotheses
    hyp = cone_hypotheses(lfpm, wl, 4000, 5)
    pm = pm_to_oracle(orc, wl.pm)
    tiles = orc.slice_u8(wl.raw(), wl.L["origins"], wl.crop)
    L = wl.L
    c64, geo, gate, ferr = orc.eval_margins(tiles, L["centers"], L["nbr"], L["cnt"], pm, hyp)
    c32 = orc.eval_costs(tiles, L["centers"], L["nbr"], L["cnt"], pm, hyp, dtype=np.float32)
    e = np.abs(c32.astype(np.float64) - c64)
    ok = (geo >= 1e-4) & (gate >= 1) & (c64 < 2)
    assert ok.sum() > 10000
    assert np.all(e[ok] <= ferr[ok]), float(np.max(e[ok] / np.maximum(ferr[ok], 1e-300)))
    ratio = e[ok & (e > 0)] / ferr[ok & (e > 0)]
    assert np.median(ratio) > 1e-4
