"""Pins of the oracle's CheckModule filter, refocus and V2R fit.

Filter (MIDisparityFilterCPU::FilterTile, CheckModule/MIDisparityFilter.cpp:423-693): exactly
consistent fronto-parallel scenes give zero cycle error and are kept; injected spikes are removed
with reason PR_Spike; the cost gate is inclusive at 1.8; invalid input pixels are untouched.
Refocus: linear in the input tiles, constant tiles refocus to the constant, a single MI reduces
to one bilinear sample.  Fit: exact model data recovers the model; noisy data agrees with numpy's
least squares.
"""
import numpy as np
import pytest

from _helpers import fp_to_oracle

PR_KEEP, PR_COST, PR_NOSUPPORT, PR_CYCLE, PR_SPIKE = 0, 80, 140, 200, 255


@pytest.fixture(scope="module")
def scene(lfpm):
    wl = lfpm.workload_c1(raw_w=320, raw_h=256)
    V, W = wl.V, wl.crop
    rng = np.random.default_rng(5)
    gray = rng.integers(0, 256, (V, W, W)).astype(np.float32)
    return wl, gray


def _inputs(wl, d=5.0, cost=0.5):
    V, W, K = wl.V, wl.crop, wl.pm.max_neighbors
    n = V * W * W
    plane = np.zeros((n, 4), np.float32)
    plane[:, 2] = 1.0
    plane[:, 3] = d
    costs = np.full(n, cost, np.float32)
    disp = np.zeros((n, 4), np.float32)
    sel = np.zeros(n, np.uint32)
    for v in range(V):
        sel[v * W * W:(v + 1) * W * W] = (1 << int(wl.L["cnt"][v])) - 1
    return plane, costs, disp, sel


def _run(orc, wl, gray, fp, plane, costs, disp, sel):
    return orc.filter_f32(gray, wl.L["centers"], wl.L["nbr"], wl.L["cnt"], wl.crop,
                          wl.pm.max_neighbors, wl.pm.base, fp, plane, costs, disp, sel)


def _fp(orc, wl, **kw):
    fp = fp_to_oracle(orc, wl.fp)
    for k, v in kw.items():
        setattr(fp, k, v)
    return fp


def test_consistent_scene_kept(orc, scene):
    wl, gray = scene
    fp = _fp(orc, wl, enable_cycle_photo_filter=0)
    out = _run(orc, wl, gray, fp, *_inputs(wl))
    r = out["reason"]
    # constant disparity: geometric cycle error is exactly zero wherever the neighbour sample
    # lands inside its crop, so every rejection is for missing support, never for the cycle
    assert not np.any(r == PR_CYCLE) and not np.any(r == PR_SPIKE) and not np.any(r == PR_COST)
    assert (r == PR_KEEP).mean() > 0.5
    kept = r == PR_KEEP
    assert np.all(out["plane"][kept, 3] == 5.0)
    dropped = r == PR_NOSUPPORT
    assert np.all(out["plane"][dropped, 3] == -1.0) and np.all(out["cost"][dropped] == 2.0)
    assert np.all(out["sel"][dropped] == 0)


def test_spike_removed(orc, scene):
    wl, gray = scene
    fp = _fp(orc, wl, enable_cycle_photo_filter=0, enable_cycle_geo_filter=0)
    plane, costs, disp, sel = _inputs(wl)
    W = wl.crop
    # the centre pixel of an interior MI kept in the clean run: raise its disparity by 1.5 px
    # (> spike_abs_diff = 1), its neighbour samples stay inside their crops
    clean = _run(orc, wl, gray, fp, plane, costs, disp, sel)["reason"]
    v = int(np.flatnonzero(wl.L["cnt"] == wl.L["cnt"].max())[5])
    idx = v * W * W + (W // 2) * W + W // 2
    assert clean[idx] == PR_KEEP
    plane[idx, 3] = 6.5
    out = _run(orc, wl, gray, fp, plane, costs, disp, sel)
    assert out["reason"][idx] == PR_SPIKE
    assert out["plane"][idx, 3] == -1.0


def test_cost_gate_inclusive(orc, scene):
    wl, gray = scene
    fp = _fp(orc, wl, enable_cycle_photo_filter=0)
    p, c, d, s = _inputs(wl)
    base = _run(orc, wl, gray, fp, p, c, d, s)["reason"]
    i = int(np.flatnonzero(base == PR_KEEP)[10])
    c2 = c.copy()
    c2[i] = np.float32(1.8)
    assert _run(orc, wl, gray, fp, p, c2, d, s)["reason"][i] != PR_COST
    c2[i] = np.nextafter(np.float32(1.8), np.float32(2))
    assert _run(orc, wl, gray, fp, p, c2, d, s)["reason"][i] == PR_COST


def test_invalid_pixels_untouched(orc, scene):
    wl, gray = scene
    fp = _fp(orc, wl)
    p, c, d, s = _inputs(wl)
    p[::7, 3] = -1.0
    p[::11, 3] = np.nan
    out = _run(orc, wl, gray, fp, p, c, d, s)
    bad = (p[:, 3] <= 0) | np.isnan(p[:, 3])
    assert np.all(out["reason"][bad] == PR_KEEP)
    assert np.array_equal(out["plane"][bad].view(np.uint32), p[bad].view(np.uint32))


def test_photo_residual_zero_on_identical_neighbours(orc, lfpm):
    """Identical gray content in every MI and zero disparity: photometric cycle residual 0."""
    wl = lfpm.workload_c1(raw_w=320, raw_h=256)
    V, W = wl.V, wl.crop
    rng = np.random.default_rng(9)
    one = rng.integers(0, 256, (W, W)).astype(np.float32)
    gray = np.broadcast_to(one, (V, W, W)).copy()
    p, c, d, s = _inputs(wl, d=1e-3)
    out = _run(orc, wl, gray, _fp(orc, wl), p, c, d, s)
    assert not np.any(out["reason"] == PR_CYCLE)


# ----------------------------------------------------------------------------------------
# Refocus
# ----------------------------------------------------------------------------------------
def test_refocus_linear_and_constant(orc, lfpm):
    wl = lfpm.workload_c1(raw_w=256, raw_h=200)
    rng = np.random.default_rng(2)
    V, W = wl.V, wl.crop
    t1 = rng.integers(0, 128, (V, W, W)).astype(np.uint8)
    t2 = rng.integers(0, 128, (V, W, W)).astype(np.uint8)
    args = (wl.L["centers"], wl.L["pitch"], W, 256, 200, 3, 2.0, 8.0)
    r1, r2 = orc.refocus(t1, *args), orc.refocus(t2, *args)
    r12 = orc.refocus((t1 + t2).astype(np.uint8), *args)
    assert np.max(np.abs(r12 - (r1 + r2))) < 1e-12
    rc = orc.refocus(np.full((V, W, W), 100, np.uint8), *args)
    covered = rc > 0
    assert covered.mean() > 0.8
    assert np.max(np.abs(rc[covered] - 100 / 255.0)) < 1e-12


def test_refocus_single_mi_is_bilinear(orc):
    from scipy.ndimage import map_coordinates
    W = 16
    rng = np.random.default_rng(4)
    tile = rng.integers(0, 256, (1, W, W)).astype(np.uint8)
    c = np.array([[40.0, 30.0]], np.float32)
    v = 3.0
    out = orc.refocus(tile, c, 23.0, W, 90, 70, 1, v, v)[0]
    ys, xs = np.mgrid[0:70, 0:90]
    ux = (xs - 40.0) / v + 7.5
    uy = (ys - 30.0) / v + 7.5
    inside = (ux >= 0) & (ux <= W - 1) & (uy >= 0) & (uy <= W - 1)
    want = map_coordinates(tile[0] / 255.0, [uy[inside], ux[inside]], order=1, mode="nearest")
    assert np.max(np.abs(out[inside] - want)) < 1e-6
    assert np.all(out[~inside] == 0.0)


# ----------------------------------------------------------------------------------------
# V2R fit
# ----------------------------------------------------------------------------------------
def test_fit_recovers_exact_model(orc):
    c_true = np.array([0.04, 1.3, 0.7])
    v = np.linspace(2.0, 8.0, 5000)
    aL = (c_true[1] * v + c_true[2]) / (1 - c_true[0] * v)
    # float32 inputs (as the reference's vectors): compare against the model on those inputs
    c = orc.fit_v2r(aL.astype(np.float32), v.astype(np.float32))
    assert np.max(np.abs(c - c_true) / np.abs(c_true)) < 2e-5
    v32 = v.astype(np.float32).astype(np.float64)
    a32 = aL.astype(np.float32).astype(np.float64)
    X = np.stack([a32 * v32, v32, np.ones_like(v32)], 1)
    want = np.linalg.lstsq(X, a32, rcond=None)[0]
    assert np.max(np.abs(c - want) / np.abs(want)) < 1e-9


def test_fit_matches_lstsq_on_noisy_data(orc):
    rng = np.random.default_rng(8)
    v = (2 + 6 * rng.random(20000)).astype(np.float32)
    aL = ((1.5 * v - 0.2) / (1 - 0.03 * v) * (1 + 0.01 * rng.standard_normal(v.size)))
    aL = aL.astype(np.float32)
    c = orc.fit_v2r(aL, v)
    a, vv = aL.astype(np.float64), v.astype(np.float64)
    want = np.linalg.lstsq(np.stack([a * vv, vv, np.ones_like(vv)], 1), a, rcond=None)[0]
    assert np.max(np.abs(c - want) / np.abs(want)) < 1e-9


def test_v2r_convert_clamps(orc):
    v = np.array([2.0, 10.0, 25.0, np.nan], np.float32)
    z = orc.v2r_convert(v, np.array([0.04, 1.0, 0.5]))
    assert abs(z[0] - (2.0 + 0.5) / (1 - 0.08)) < 1e-12
    assert z[2] == 0.0          # negative -> 0
    assert z[3] == 0.0          # nan -> 0
    assert z[1] > 0


def _run_inplace(orc, wl, gray, fp, plane, costs, disp, sel):
    return orc.filter_f32_inplace(gray, wl.L["centers"], wl.L["nbr"], wl.L["cnt"], wl.crop,
                                  wl.pm.max_neighbors, wl.pm.base, fp, plane, costs, disp, sel)


def test_inplace_order_equals_snapshot_when_nothing_is_dropped(orc, scene):
    """With every gate open (no pixel invalidated) the reference's in-place order
    (FilterGlobal, CheckModule/MIDisparityFilter.cpp:973-985) and the snapshot agree exactly."""
    wl, gray = scene
    fp = _fp(orc, wl, enable_cycle_photo_filter=0, enable_cycle_geo_filter=0,
             enable_spike_filter=0, min_good_neighbors=0)
    args = _inputs(wl)
    a = _run(orc, wl, gray, fp, *args)
    b = _run_inplace(orc, wl, gray, fp, *args)
    assert np.all(a["reason"] == PR_KEEP)
    for k in a:
        assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), k


def test_inplace_order_only_later_mis_change(orc, scene):
    """The in-place order differs from the snapshot only in MIs that read a pixel an earlier MI
    invalidated: MI 0 is identical, and every MI whose reasons differ has a neighbour with a
    smaller index that dropped a pixel."""
    wl, gray = scene
    fp = _fp(orc, wl)
    plane, costs, disp, sel = _inputs(wl)
    rng = np.random.default_rng(9)
    W, V, K = wl.crop, wl.V, wl.pm.max_neighbors
    # spikes and high costs in a tenth of the MIs invalidate pixels their neighbours read
    for v in rng.choice(V, V // 10, replace=False):
        idx = v * W * W + rng.integers(0, W * W, 12)
        plane[idx[:6], 3] += 1.5
        costs[idx[6:]] = 1.9
    a = _run(orc, wl, gray, fp, plane, costs, disp, sel)["reason"].reshape(V, W * W)
    b = _run_inplace(orc, wl, gray, fp, plane, costs, disp, sel)["reason"].reshape(V, W * W)
    assert np.array_equal(a[0], b[0])
    dropped = (b != PR_KEEP).any(axis=1)
    diff = np.flatnonzero((a != b).any(axis=1))
    assert diff.size > 0, "the scene must exercise the order dependence"
    for u in diff:
        nb = wl.L["nbr"][u * K:(u + 1) * K] if wl.L["nbr"].ndim == 1 else wl.L["nbr"][u]
        earlier = [n for n in nb[:int(wl.L["cnt"][u])] if 0 <= n < u]
        assert any(dropped[n] for n in earlier), u


def _levels_brute(nbr, cnt, K):
    """Longest chain of smaller-index neighbours ending at each MI, by plain recursion."""
    V = len(cnt)
    nbr = np.asarray(nbr).reshape(V, K)
    memo = {}

    def lev(v):
        if v not in memo:
            ds = [int(n) for n in nbr[v, :min(int(cnt[v]), K)] if 0 <= n < v]
            memo[v] = 1 + max(lev(n) for n in ds) if ds else 0
        return memo[v]

    return np.array([lev(v) for v in range(V)])


@pytest.mark.parametrize("which", ["c1", "c2"])
def test_filter_levels_schedule(lfpm, which):
    """lfpm_filter_order_levels: a permutation of the MIs, level by level in index order within
    a level; each MI's level is one more than the largest level of its smaller-index neighbours
    (brute force), so no MI reads the filtered state of an MI of its own or a later level."""
    wl = (lfpm.workload_c1(raw_w=400, raw_h=320) if which == "c1"
          else lfpm.workload_c2(raw_w=1024, raw_h=800))
    K = wl.pm.max_neighbors
    order, start = lfpm.filter_levels(wl.L["nbr"], wl.L["cnt"], K)
    want = _levels_brute(wl.L["nbr"], wl.L["cnt"], K)
    assert np.array_equal(np.sort(order), np.arange(wl.V))
    assert start[0] == 0 and start[-1] == wl.V and np.all(np.diff(start) > 0)
    got = np.empty(wl.V, np.int64)
    for l in range(len(start) - 1):
        members = order[start[l]:start[l + 1]]
        assert np.all(np.diff(members) > 0)
        got[members] = l
    assert np.array_equal(got, want)
    assert len(start) - 1 == want.max() + 1


def test_filter_levels_small_cases(lfpm):
    """A chain 0 <- 1 <- 2 gives three levels; MIs without smaller-index neighbours share
    level 0; -1 padding and counts below K are respected."""
    K = 2
    nbr = np.array([[1, -1], [0, 2], [1, -1]], np.int32)
    order, start = lfpm.filter_levels(nbr, np.array([1, 2, 1], np.int32), K)
    assert list(order) == [0, 1, 2] and list(start) == [0, 1, 2, 3]
    # MI 2's only listed neighbour is beyond its count: it joins level 0
    order, start = lfpm.filter_levels(nbr, np.array([1, 1, 0], np.int32), K)
    assert list(order) == [0, 2, 1] and list(start) == [0, 2, 3]


def test_filter_ordered_rejects_bad_schedules(lfpm):
    """Host-side validation of lfpm_filter_frame_ordered, before any device work."""
    import ctypes as C
    fp = lfpm.filter_default()
    fake = C.c_void_p(256)

    def call(start, nl, order=fake, out_plane=C.c_void_p(512)):
        s = np.asarray(start, np.int32)
        return lfpm.lib().lfpm_filter_frame_ordered(
            fake, fake, fake, fake, 3, 16, 2, C.c_float(23.0), C.byref(fp), fake, fake, fake,
            fake, out_plane, C.c_void_p(1024), fake, fake, fake, order,
            s.ctypes.data_as(C.POINTER(C.c_int)), nl, None)

    assert call([0, 2, 3], 2, order=None) == -1
    assert call([0, 2, 4], 2) == -1          # does not end at num_mis
    assert b"cover" in lfpm.lib().lfpm_last_error()
    assert call([0, 2, 2, 3], 3) == -1       # empty level
    assert call([0, 3], 0) == -1
    assert call([0, 2, 3], 2, out_plane=fake) == -1   # output aliases the input
    assert b"alias" in lfpm.lib().lfpm_last_error()


def test_inplace_order_equals_iterated_snapshot(orc, lfpm):
    """The reference's order restated from the snapshot form: MI v's result is the snapshot
    filter of the state that MIs 0..v-1 left, restricted to MI v (one full snapshot call per MI
    on a small frame).  The in-place twin must equal it bit for bit."""
    wl = lfpm.workload_c1(raw_w=256, raw_h=200)
    V, W = wl.V, wl.crop
    WW = W * W
    rng = np.random.default_rng(17)
    gray = rng.integers(0, 256, (V, W, W)).astype(np.float32)
    plane, costs, disp, sel = _inputs(wl)
    # a varied state: disparities around 5 px with spikes, some high costs
    plane[:, 3] = (5.0 + rng.normal(0.0, 0.3, plane.shape[0])).astype(np.float32)
    plane[rng.random(plane.shape[0]) < 0.05, 3] += 2.0
    costs[rng.random(costs.shape[0]) < 0.1] = 1.95
    fp = _fp(orc, wl, enable_cycle_photo_filter=0)
    want = {k: np.empty_like(a) for k, a in
            dict(plane=plane, cost=costs, disp=disp, sel=sel).items()}
    want["reason"] = np.empty(V * WW, np.uint8)
    cur_plane, cur_cost = plane.copy(), costs.copy()
    for v in range(V):
        o = _run(orc, wl, gray, fp, cur_plane, cur_cost, disp, sel)
        sl = slice(v * WW, (v + 1) * WW)
        for k in ("plane", "cost", "disp", "sel", "reason"):
            want[k][sl] = o[k][sl]
        cur_plane[sl] = o["plane"][sl]
        cur_cost[sl] = o["cost"][sl]
    got = _run_inplace(orc, wl, gray, fp, plane, costs, disp, sel)
    assert (want["reason"] != PR_KEEP).any() and (want["reason"] == PR_KEEP).any()
    for k in want:
        assert np.array_equal(got[k].view(np.uint8), want[k].view(np.uint8)), k
