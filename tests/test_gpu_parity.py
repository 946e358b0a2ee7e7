"""GPU parity tests: the HIP path (through the C ABI) against the CPU oracle on the same seeded
inputs.  Bars: bit-exact for gathers, masks and determinism; fixed-plane costs within the
stated relative tolerance of the FP64 oracle; full-solve disparities within 0.05 px of the FP64
oracle on >= 99.5 % of valid pixels (valid = oracle final cost <= 1.8, the CheckModule cost gate).
"""
import ctypes as C
import time

import numpy as np
import pytest

from _helpers import cone_hypotheses, cost_bar, fp_to_oracle, pm_to_oracle
from conftest import stats_dump

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


@pytest.fixture(scope="module")
def c1(lfpm, torch_cuda):
    wl = lfpm.workload_c1()
    df = lfpm.DeviceFrame(wl)
    return wl, df


def test_slice_exact(lfpm, orc, c1, torch_cuda):
    wl, df = c1
    tiles_u8 = orc.slice_u8(wl.raw(), wl.L["origins"], wl.crop)
    ref = tiles_u8.astype(np.float32) / np.float32(255.0)
    torch_cuda.cuda.synchronize()
    assert np.array_equal(df.tiles.cpu().numpy(), ref)
    assert np.array_equal(df.gray.cpu().numpy(), tiles_u8.astype(np.float32))


@pytest.mark.parametrize("cfg", ["c1", "c2params", "c2params_fp16"])
def test_eval_costs_vs_fp64(lfpm, orc, torch_cuda, cfg):
    """Fixed-plane bilateral NCC (ComputeBilateralNCC_Frame) vs the FP64 oracle on 40,000 cone
    hypotheses x all views.  Relative error |c - c*| / max(|c*|, 1e-3) at most 1e-5 (FP32 tiles)
    or 2e-3 (FP16 tiles in LDS, FP32 arithmetic) on every evaluation, except where FP32 cannot
    decide it (cost_bar: straddles of the crop edge or the variance gate, ill-conditioned
    windows); those excused evaluations are a small stated fraction."""
    _say("case eval_costs", cfg)
    if cfg == "c1":
        wl = lfpm.workload_c1()
    else:  # C2 window/neighbour parameters on a 1024x800 frame
        wl = lfpm.workload_c2(raw_w=1024, raw_h=800)
    fp16 = cfg.endswith("fp16")
    if fp16:
        wl.pm.storage = 1
    df = lfpm.DeviceFrame(wl)
    hyp = cone_hypotheses(lfpm, wl, 40000, seed=7)
    out = lfpm.eval_costs(df, torch_cuda.from_numpy(hyp).cuda()).cpu().numpy()
    tiles = orc.slice_u8(wl.raw(), wl.L["origins"], wl.crop)
    ref, geo, gate, ferr = orc.eval_margins(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"],
                                            pm_to_oracle(orc, wl.pm), hyp)
    tol = 2e-3 if fp16 else 1e-5
    st = dict(cfg=cfg, tol=tol, **cost_bar(out, ref, geo, gate, ferr, tol))
    _say("eval_costs", st)
    stats_dump("eval_costs_" + cfg, st)
    assert st["unexplained"] == 0, st
    assert st["excused_frac"] <= 0.15, st
    assert st["frac_le_tol"] >= 0.995, st


def _solve_gpu(lfpm, df, torch_cuda, pm=None):
    bufs = lfpm.patchmatch_frame(df, pm=pm)
    torch_cuda.cuda.synchronize()
    return dict(plane=bufs.plane.cpu().numpy(), cost=bufs.cost.cpu().numpy(),
                disp=bufs.disp.cpu().numpy(), sel=bufs.sel.cpu().numpy().view(np.uint32)), bufs


def test_solve_c1_vs_fp64(lfpm, orc, c1, torch_cuda):
    """Full solve (init + 3 red-black iterations + refinement) on the 1024x800 frame."""
    wl, df = c1
    _say("case solve c1 V=%d" % wl.V)
    g, _ = _solve_gpu(lfpm, df, torch_cuda)
    tiles = orc.slice_u8(wl.raw(), wl.L["origins"], wl.crop)
    t0 = time.time()
    o = orc.solve(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"], pm_to_oracle(orc, wl.pm))
    t_orc = time.time() - t0
    dg, do = g["plane"][:, 3].astype(np.float64), o["plane"][:, 3]
    valid = o["cost"] <= 1.8
    err = np.abs(dg - do)
    frac = float((err[valid] <= 0.05).mean())
    # ground truth (synthetic scene): GT disparity = pitch / virtual depth
    W = wl.crop
    gt = wl.gt_virtual_depth()
    gtv = np.stack([gt[y:y + W, x:x + W] for x, y in wl.L["origins"]]).reshape(-1)
    gtd = wl.L["pitch"] / np.maximum(gtv, 1e-6)
    st = dict(V=wl.V, frac_le_005=frac, valid=float(valid.mean()),
              frac_all_le_005=float((err <= 0.05).mean()),
              sel_equal=float((g["sel"] == o["sel"]).mean()),
              cost_med_abs=float(np.median(np.abs(g["cost"] - o["cost"]))),
              gt_med_err_gpu=float(np.median(np.abs(dg - gtd))),
              gt_med_err_orc=float(np.median(np.abs(do - gtd))),
              gt_frac05_gpu=float((np.abs(dg - gtd) < 0.5).mean()),
              oracle_s=t_orc)
    _say("solve_c1", st)
    stats_dump("solve_c1", st)
    assert frac >= 0.995


def test_solve_c1_fp16_storage_vs_fp64(lfpm, orc, torch_cuda):
    """FP16 tiles in LDS: full-solve disparities within 0.05 px of the FP64 oracle on >= 99.5 %
    of valid pixels (the FP32 bar)."""
    import copy
    wl = lfpm.workload_c1()
    pm = copy.copy(wl.pm)
    pm.storage = 1
    df = lfpm.DeviceFrame(wl)
    g, _ = _solve_gpu(lfpm, df, torch_cuda, pm=pm)
    tiles = orc.slice_u8(wl.raw(), wl.L["origins"], wl.crop)
    o = orc.solve(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"], pm_to_oracle(orc, wl.pm))
    valid = o["cost"] <= 1.8
    err = np.abs(g["plane"][:, 3].astype(np.float64) - o["plane"][:, 3])
    st = dict(frac_le_005=float((err[valid] <= 0.05).mean()),
              frac_le_01=float((err[valid] <= 0.1).mean()), median=float(np.median(err)))
    _say("solve_c1_fp16", st)
    stats_dump("solve_c1_fp16", st)
    assert st["frac_le_005"] >= 0.995


def test_solve_deterministic(lfpm, c1, torch_cuda):
    wl, df = c1
    a, _ = _solve_gpu(lfpm, df, torch_cuda)
    b, _ = _solve_gpu(lfpm, df, torch_cuda)
    for k in a:
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k


def test_solve_seed_changes_result(lfpm, c1, torch_cuda):
    wl, df = c1
    import copy
    pm = copy.copy(wl.pm)
    pm.frame_uid = 99
    a, _ = _solve_gpu(lfpm, df, torch_cuda)
    b, _ = _solve_gpu(lfpm, df, torch_cuda, pm=pm)
    assert not np.array_equal(a["plane"], b["plane"])


def _solve_steps_without_records(lfpm, df, torch, pm):
    """The same solve through the step calls with no workspace, so that no per-pixel cost
    records are kept and every update scores the pixel's current plane afresh."""
    V, W = df.wl.V, df.wl.crop
    n = V * W * W
    L, st, dp = lfpm.lib(), lfpm._stream(), lfpm._dp
    a = (torch.empty((n, 4), device="cuda"), torch.empty(n, device="cuda"),
         torch.empty(n, dtype=torch.int32, device="cuda"))
    b = (torch.empty((n, 4), device="cuda"), torch.empty(n, device="cuda"),
         torch.empty(n, dtype=torch.int32, device="cuda"))
    disp = torch.empty((n, 4), device="cuda")
    args = (dp(df.tiles), dp(df.centers), dp(df.nbr), dp(df.cnt), V, C.byref(pm))
    lfpm._check(L.lfpm_patchmatch_init(*args, dp(a[0]), dp(a[1]), dp(a[2]), None,
                                       C.c_size_t(0), st))
    for it in range(pm.iterations):
        for ph, (src, dst) in enumerate(((a, b), (b, a))):
            lfpm._check(L.lfpm_patchmatch_phase(*args, it, ph, dp(src[0]), dp(src[1]),
                                                dp(src[2]), dp(dst[0]), dp(dst[1]), dp(dst[2]),
                                                None, C.c_size_t(0), st))
    lfpm._check(L.lfpm_patchmatch_finalize(*args, dp(a[0]), dp(disp), st))
    torch.cuda.synchronize()
    return dict(plane=a[0].cpu().numpy(), cost=a[1].cpu().numpy(), disp=disp.cpu().numpy(),
                sel=a[2].cpu().numpy().view(np.uint32))


@pytest.mark.parametrize("blur", [False, True])
def test_cost_records_keep_results(lfpm, torch_cuda, blur):
    """The frame solve reuses each pixel's recorded all-view costs while its plane is unchanged;
    its outputs equal, bit for bit, those of the step calls without a workspace (no records)."""
    import copy
    wl = lfpm.workload_c2(raw_w=800, raw_h=640)
    pm = copy.copy(wl.pm)
    if blur:
        pm.blur_lambda = 0.25
    df = lfpm.DeviceFrame(wl, blur_raw="laplacian" if blur else None)
    _say("case cost records blur=%s V=%d" % (blur, wl.V))
    g, bufs = _solve_gpu(lfpm, df, torch_cuda, pm=pm)
    s = _solve_steps_without_records(lfpm, df, torch_cuda, pm)
    for k in g:
        assert np.array_equal(g[k].view(np.uint32), s[k].view(np.uint32)), k
    # the step-API final record on the solve's workspace (records of the final planes)
    V, W = df.wl.V, df.wl.crop
    n = V * W * W
    off = (n * 24 + 255) // 256 * 256
    disp = torch_cuda.empty((n, 4), device="cuda")
    L, dp = lfpm.lib(), lfpm._dp
    lfpm._check(L.lfpm_patchmatch_finalize_ws(
        dp(df.tiles), dp(df.centers), dp(df.nbr), dp(df.cnt), V, C.byref(pm), dp(bufs.plane),
        dp(disp), dp(bufs.ws[off:]), C.c_size_t(bufs.ws_bytes - off), lfpm._stream()))
    torch_cuda.cuda.synchronize()
    assert np.array_equal(disp.cpu().numpy().view(np.uint32), g["disp"].view(np.uint32))


def test_filter_bitexact(lfpm, orc, c1, torch_cuda):
    """Fused CheckModule pass vs the FP32 CPU twin on the GPU solve's outputs: identical reason
    masks and outputs, bit for bit."""
    wl, df = c1
    g, bufs = _solve_gpu(lfpm, df, torch_cuda)
    fo = lfpm.filter_frame(df, bufs)
    torch_cuda.cuda.synchronize()
    gf = {k: v.cpu().numpy() for k, v in fo.items()}
    gray = df.gray.cpu().numpy()
    of = orc.filter_f32(gray, wl.L["centers"], wl.L["nbr"], wl.L["cnt"], wl.crop, wl.pm.max_neighbors,
                        wl.pm.base, fp_to_oracle(orc, wl.fp), g["plane"], g["cost"], g["disp"],
                        g["sel"])
    reasons = {int(r): int((gf["reason"] == r).sum()) for r in (0, 80, 140, 200, 255)}
    _say("filter reasons", reasons)
    stats_dump("filter_c1", dict(reasons=reasons))
    assert np.array_equal(gf["reason"], of["reason"])
    for k in ("plane", "cost", "disp"):
        assert np.array_equal(gf[k].view(np.uint32), of[k].view(np.uint32)), k
    assert np.array_equal(gf["sel"].view(np.uint32), of["sel"])


def test_refocus_vs_fp64(lfpm, orc, torch_cuda):
    wl = lfpm.workload_c1(raw_w=320, raw_h=256)
    df = lfpm.DeviceFrame(wl)
    out = lfpm.refocus(df, slices=5, vmin=2.0, vmax=8.0)
    torch_cuda.cuda.synchronize()
    g = out.cpu().numpy().astype(np.float64)
    tiles = orc.slice_u8(wl.raw(), wl.L["origins"], wl.crop)
    o = orc.refocus(tiles, wl.L["centers"], wl.L["pitch"], wl.crop, 320, 256, 5, 2.0, 8.0)
    err = np.abs(g - o)
    _say("refocus max abs err", float(err.max()))
    stats_dump("refocus", dict(max_abs=float(err.max())))
    assert err.max() <= 1e-5


@pytest.mark.parametrize("w,h,pitch,slices,vmin,vmax", [
    (330, 290, 23.0, 6, 1.5, 9.0),    # partial 16x16 output tiles on both axes
    (400, 300, 31.0, 4, 2.0, 8.0),    # crop 21
    (300, 250, 15.0, 4, 2.0, 8.0),    # crop 10
    (256, 208, 23.0, 3, 2.0, 30.0),   # a reach beyond the tile kernel's candidate bound:
                                      # the per-pixel kernel
])
def test_refocus_shapes_vs_fp64(lfpm, orc, torch_cuda, w, h, pitch, slices, vmin, vmax):
    """Refocus stacks on frames whose size is not a multiple of the 16x16 output tile, at three
    lens pitches, and with a slice range whose reach sends the call to the per-pixel kernel:
    within 1e-5 of the FP64 oracle everywhere."""
    radius = {15.0: 5, 23.0: 7, 31.0: 9}[pitch]
    wl = lfpm.workload_c2(raw_w=w, raw_h=h, pitch=pitch, radius=radius)
    _say("case refocus %dx%d pitch %g slices %d v %g..%g" % (w, h, pitch, slices, vmin, vmax))
    df = lfpm.DeviceFrame(wl)
    out = lfpm.refocus(df, slices=slices, vmin=vmin, vmax=vmax)
    torch_cuda.cuda.synchronize()
    g = out.cpu().numpy().astype(np.float64)
    tiles = orc.slice_u8(wl.raw(), wl.L["origins"], wl.crop)
    o = orc.refocus(tiles, wl.L["centers"], wl.L["pitch"], wl.crop, w, h, slices, vmin, vmax)
    err = np.abs(g - o)
    _say("refocus shapes max abs err", float(err.max()), "covered", float((o > 0).mean()))
    assert (o > 0).mean() > 0.3
    assert err.max() <= 1e-5


def test_fit_v2r_fp64(lfpm, orc, torch_cuda):
    rng = np.random.default_rng(3)
    c_true = np.array([0.05, 1.7, -0.4])
    v = (2 + 6 * rng.random(200000)).astype(np.float32)
    aL = ((c_true[1] * v + c_true[2]) / (1 - c_true[0] * v)).astype(np.float32)
    aL_noisy = (aL * (1 + 1e-3 * rng.standard_normal(aL.size))).astype(np.float32)
    cg = lfpm.fit_v2r(torch_cuda.from_numpy(aL_noisy).cuda(), torch_cuda.from_numpy(v).cuda())
    co = orc.fit_v2r(aL_noisy, v)
    st = dict(gpu=cg.tolist(), oracle=co.tolist(), rel=float(np.max(np.abs(cg - co) / np.abs(co))))
    _say("fit", st)
    stats_dump("fit_v2r", st)
    assert np.max(np.abs(cg - co) / np.abs(co)) <= 1e-8
    z = lfpm.v2r_convert(torch_cuda.from_numpy(v).cuda(), cg).cpu().numpy()
    zo = orc.v2r_convert(v, cg)
    assert np.array_equal(z, zo)


def test_ctx_api_matches_device_path(lfpm, c1, torch_cuda):
    """lfpm_ctx_solve_frame (host in/out) gives the device path's results bit for bit."""
    wl, df = c1
    g, bufs = _solve_gpu(lfpm, df, torch_cuda)
    fo = lfpm.filter_frame(df, bufs)
    torch_cuda.cuda.synchronize()
    L = lfpm.lib()
    ctx = C.c_void_p()
    assert L.lfpm_create(0, C.byref(ctx)) == 0
    try:
        n = wl.V * wl.crop ** 2
        plane = np.zeros((n, 4), np.float32)
        cost = np.zeros(n, np.float32)
        disp = np.zeros((n, 4), np.float32)
        sel = np.zeros(n, np.uint32)
        reason = np.zeros(n, np.uint8)
        raw = np.ascontiguousarray(wl.raw())
        rc = L.lfpm_ctx_solve_frame(ctx, C.byref(wl.lp), raw.ctypes.data_as(C.c_void_p),
                                    C.byref(wl.pm), C.byref(wl.fp),
                                    plane.ctypes.data_as(C.c_void_p),
                                    cost.ctypes.data_as(C.c_void_p),
                                    disp.ctypes.data_as(C.c_void_p),
                                    sel.ctypes.data_as(C.c_void_p),
                                    reason.ctypes.data_as(C.c_void_p))
        assert rc == 0, L.lfpm_last_error()
    finally:
        L.lfpm_destroy(ctx)
    assert np.array_equal(reason, fo["reason"].cpu().numpy())
    assert np.array_equal(plane.view(np.uint32), fo["plane"].cpu().numpy().view(np.uint32))


@pytest.mark.parametrize("pitch", [23.0, 35.0], ids=["one_level", "two_levels"])
def test_solve_acmm_vs_fp64(lfpm, orc, torch_cuda, pitch):
    """ACMM variant (AdaptMIPMFrameACMM) vs the FP64 oracle: final-level disparities within
    0.05 px on >= 99.5 % of pixels; cost/sel are the top-k support of the final planes."""
    w, h = (400, 320) if pitch < 30 else (560, 448)
    wl = lfpm.workload_c1(raw_w=w, raw_h=h, pitch=pitch)
    _say("case acmm pitch %g V=%d crop=%d" % (pitch, wl.V, wl.crop))
    df = lfpm.DeviceFrame(wl)
    bufs = lfpm.patchmatch_frame_acmm(df)
    torch_cuda.cuda.synchronize()
    d = bufs.plane[:, 3].cpu().numpy().astype(np.float64)
    tiles = orc.slice_u8(wl.raw(), wl.L["origins"], wl.crop)
    o = orc.solve_acmm(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"],
                       pm_to_oracle(orc, wl.pm))
    err = np.abs(d - o["plane"][:, 3])
    frac = float((err <= 0.05).mean())
    sel_eq = float((bufs.sel.cpu().numpy().view(np.uint32) == o["sel"]).mean())
    st = dict(pitch=pitch, V=wl.V, frac_le_005=frac, sel_equal=sel_eq,
              median=float(np.median(err)))
    _say("acmm", st)
    stats_dump("acmm_%g" % pitch, st)
    assert frac >= 0.995


@pytest.mark.parametrize("pitch", [23.0, 35.0], ids=["one_level", "two_levels"])
def test_solve_acmm_blur_vs_fp64(lfpm, orc, torch_cuda, pitch):
    """ACMM with the blur term (lambda 0.8, the frame's defocus maps; a coarse level reads them
    area-resized) vs the FP64 oracle: final-level disparities within 0.05 px on >= 99.5 % of
    pixels."""
    import copy
    w, h = (400, 320) if pitch < 30 else (560, 448)
    wl = lfpm.workload_c1(raw_w=w, raw_h=h, pitch=pitch)
    pm = copy.copy(wl.pm)
    pm.blur_lambda = 0.8
    _say("case acmm blur pitch %g V=%d crop=%d" % (pitch, wl.V, wl.crop))
    blur = wl.blur_map()
    df = lfpm.DeviceFrame(wl, blur_raw=blur)
    bufs = lfpm.patchmatch_frame_acmm(df, pm=pm)
    torch_cuda.cuda.synchronize()
    d = bufs.plane[:, 3].cpu().numpy().astype(np.float64)
    tiles = orc.slice_u8(wl.raw(), wl.L["origins"], wl.crop)
    bv = orc.slice_u8(blur, wl.L["origins"], wl.crop).astype(np.float64) / 255.0
    o = orc.solve_acmm(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"], pm_to_oracle(orc, pm),
                       blur=bv)
    err = np.abs(d - o["plane"][:, 3])
    frac = float((err <= 0.05).mean())
    st = dict(pitch=pitch, V=wl.V, frac_le_005=frac, median=float(np.median(err)))
    _say("acmm blur", st)
    stats_dump("acmm_blur_%g" % pitch, st)
    assert frac >= 0.995


def _blur_inputs(orc, wl, src):
    """(DeviceFrame blur argument, oracle blur values) for a blur source: the synthetic defocus
    map, or the reference's Laplacian blur score of the frame."""
    if src == "laplacian":
        return "laplacian", orc.laplacian_blur_tiles(wl.raw(), wl.L["origins"],
                                                     wl.crop).astype(np.float64)
    blur = wl.blur_map()
    return blur, orc.slice_u8(blur, wl.L["origins"], wl.crop).astype(np.float64) / 255.0


def test_slice_blur_tiles_exact(lfpm, orc, torch_cuda):
    """lfpm_slice_blur_tiles == the oracle's Laplacian blur tiles, bit for bit."""
    wl = lfpm.workload_c1()
    df = lfpm.DeviceFrame(wl, blur_raw="laplacian")
    got = df.tiles[wl.V:].cpu().numpy()
    want = orc.laplacian_blur_tiles(wl.raw(), wl.L["origins"], wl.crop)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.abs(want).max() > 0.5


@pytest.mark.parametrize("cfg,src", [("c1", "defocus"), ("c2params", "defocus"),
                                     ("c2params", "laplacian")])
def test_eval_costs_blur_vs_fp64(lfpm, orc, torch_cuda, cfg, src):
    """Blur term of the cost vector (lambda 2, AdaptMIPMFrame.cu:211-218) with per-MI blur
    values from the synthetic defocus map or the Laplacian blur score: fixed-plane costs vs the
    FP64 oracle at the FP32 bar of test_eval_costs_vs_fp64 (1e-5 on every evaluation that FP32
    can decide; the error estimate includes the blur sums)."""
    _say("case eval_costs_blur", cfg, src)
    wl = lfpm.workload_c1() if cfg == "c1" else lfpm.workload_c2(raw_w=1024, raw_h=800)
    wl.pm.blur_lambda = 2.0
    barg, bval = _blur_inputs(orc, wl, src)
    df = lfpm.DeviceFrame(wl, blur_raw=barg)
    hyp = cone_hypotheses(lfpm, wl, 40000, seed=11)
    out = lfpm.eval_costs(df, torch_cuda.from_numpy(hyp).cuda()).cpu().numpy()
    tiles = orc.slice_u8(wl.raw(), wl.L["origins"], wl.crop)
    ref, geo, gate, ferr = orc.eval_margins(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"],
                                            pm_to_oracle(orc, wl.pm), hyp, blur=bval)
    pm0 = pm_to_oracle(orc, wl.pm)
    pm0.blur_lambda = 0.0
    noblur = orc.eval_costs(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"], pm0, hyp)
    both = (out < 2) & (ref < 2)
    st = dict(cfg=cfg, src=src, **cost_bar(out, ref, geo, gate, ferr, 1e-5),
              term_active=float((ref[both] > noblur[both] + 1e-6).mean()))
    _say("eval_costs_blur", st)
    stats_dump("eval_costs_blur_%s_%s" % (cfg, src), st)
    assert st["term_active"] > 0.5
    assert st["unexplained"] == 0, st
    assert st["excused_frac"] <= 0.15, st
    assert st["frac_le_tol"] >= 0.995, st


@pytest.mark.parametrize("src", ["defocus", "laplacian"])
def test_solve_blur_vs_fp64(lfpm, orc, torch_cuda, src):
    """Full solve with the blur term (lambda 2) on a 400x320 C1-parameter frame: disparities
    within 0.05 px of the FP64 oracle on >= 99.5 % of valid pixels."""
    wl = lfpm.workload_c1(raw_w=400, raw_h=320)
    wl.pm.blur_lambda = 2.0
    barg, bval = _blur_inputs(orc, wl, src)
    _say("case solve_blur %s V=%d" % (src, wl.V))
    df = lfpm.DeviceFrame(wl, blur_raw=barg)
    g, _ = _solve_gpu(lfpm, df, torch_cuda)
    tiles = orc.slice_u8(wl.raw(), wl.L["origins"], wl.crop)
    o = orc.solve(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"], pm_to_oracle(orc, wl.pm),
                  blur=bval)
    valid = o["cost"] <= 1.8
    err = np.abs(g["plane"][:, 3].astype(np.float64) - o["plane"][:, 3])
    st = dict(V=wl.V, src=src, frac_le_005=float((err[valid] <= 0.05).mean()),
              valid=float(valid.mean()), sel_equal=float((g["sel"] == o["sel"]).mean()))
    _say("solve_blur", st)
    stats_dump("solve_blur_" + src, st)
    assert st["frac_le_005"] >= 0.995


def test_blur_rejected_where_not_provided(lfpm, torch_cuda):
    """The blur term needs FP32 blur tiles: FP16 storage and caller-supplied tiles in the host
    context (no raw frame to take the Laplacian of) refuse it with LFPM_E_UNSUPPORTED instead
    of solving without it."""
    import copy
    wl = lfpm.workload_c1(raw_w=320, raw_h=256)
    pm = copy.copy(wl.pm)
    pm.blur_lambda = 2.0
    df = lfpm.DeviceFrame(wl, blur_raw=wl.blur_map())
    pm16 = copy.copy(pm)
    pm16.storage = 1
    with pytest.raises(lfpm.LfpmError):
        lfpm.patchmatch_frame(df, pm=pm16)
    ctx = lfpm.Context()
    try:
        tiles = np.ascontiguousarray(
            np.stack([wl.raw()[y:y + wl.crop, x:x + wl.crop] for x, y in wl.L["origins"]]))
        with pytest.raises(lfpm.LfpmError):
            ctx.solve_tiles(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"], pm)
    finally:
        ctx.close()


def test_ctx_blur_term_matches_device_path(lfpm, torch_cuda):
    """The host context takes the raw frame's Laplacian blur tiles for the blur term: its
    single-frame and batched solves equal the device path with those tiles, bit for bit."""
    import copy
    wl = lfpm.workload_c1(raw_w=320, raw_h=256)
    pm = copy.copy(wl.pm)
    pm.blur_lambda = 1.0
    df = lfpm.DeviceFrame(wl, blur_raw="laplacian")
    g, _ = _solve_gpu(lfpm, df, torch_cuda, pm=pm)
    ctx = lfpm.Context()
    try:
        o = ctx.solve_frame(wl, filt=False, pm=pm)
        outs = ctx.solve_frames(wl, [wl.raw()], filt=False, pm=pm)
    finally:
        ctx.close()
    for out in (o, outs[0]):
        for k in ("plane", "cost", "sel", "disp"):
            assert np.array_equal(np.asarray(out[k]).view(np.uint32), g[k].view(np.uint32)), k


def test_counting_kernels_identical(lfpm, torch_cuda):
    """The counting kernel variants (work counters attached) solve a C2-parameter frame to the
    same planes, costs and selections as the plain kernels, and count the views the two-round
    candidate scoring leaves out."""
    torch = torch_cuda
    wl = lfpm.workload_c2(raw_w=1024, raw_h=800, iterations=3)
    df = lfpm.DeviceFrame(wl)
    _say("case counting kernels V=%d" % wl.V)
    plain = lfpm.patchmatch_frame(df)
    torch.cuda.synchronize()
    cnt = torch.zeros(4, dtype=torch.int64, device="cuda")
    L = lfpm.lib()
    L.lfpm_set_work_counters(C.c_void_p(cnt.data_ptr()))
    try:
        counted = lfpm.patchmatch_frame(df)
        torch.cuda.synchronize()
    finally:
        L.lfpm_set_work_counters(None)
    for k in ("plane", "cost", "disp", "sel"):
        a, b = getattr(plain, k).cpu().numpy(), getattr(counted, k).cpu().numpy()
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), k
    c = cnt.cpu().numpy()
    _say("work", c.tolist(), "left out %.1f %% of samples" % (100.0 * c[2] / (c[0] + c[2])))
    assert c[0] > 0 and c[1] > 0 and c[2] > 0 and c[3] > 0
    # a left-out triple has at most the full window of samples
    assert c[2] <= c[3] * (2 * wl.pm.radius + 1) ** 2
