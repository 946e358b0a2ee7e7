"""C++ facades (namespace LFMVS, include/lfmvs.hpp) on the GPU, driven like the reference's
main() drives DepthSolverManager: DepthSolverPipelineBySequence and DepthSolverPipelineByBatch
over two synthetic frames give, per frame, the C-ABI path's filtered planes, costs, selections
and CheckModule reasons (the reference's MI order, the facades' default) bit for bit, and the oracle's disparity mosaic, virtual-depth map and
colour image; AdaptMIPM on one micro-image equals lfpm_ctx_solve_tiles on the same work list and
agrees with the FP64 oracle."""
import copy
import os
import struct
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "_build", "facade_driver")


def _say(*a):
    print(*a, flush=True)


def _run(lfpm, wl, nframes, mode, tmp_path):
    cfg = tmp_path / "cfg.bin"
    cfg.write_bytes(bytes(wl.lp) + bytes(wl.sp) + bytes(wl.pm) + bytes(wl.fp) +
                    struct.pack("ii", nframes, mode))
    out = tmp_path / ("out%d" % mode)
    out.mkdir()
    _say("facade_driver mode %d frames %d" % (mode, nframes))
    r = subprocess.run([DRIVER, str(cfg), str(out)], capture_output=True, text=True, timeout=300)
    _say(r.stdout.strip(), r.stderr.strip())
    assert r.returncode == 0, r.stderr
    return out


def _frame_workload(lfpm, i):
    wl = lfpm.workload_c1(raw_w=320, raw_h=256, scene_seed=1234 + 17 * i,
                          noise_seed=1235 + 17 * i)
    wl.pm.frame_uid = 5 + i
    return wl


@pytest.fixture(scope="module")
def expected(lfpm, orc):
    import torch
    assert torch.cuda.is_available()
    if not os.path.exists(DRIVER):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")])
    ctx = lfpm.Context()
    ctx.set_filter_order("reference")  # the facades' default (LightFieldParams)
    exp = []
    try:
        for i in range(2):
            wl = _frame_workload(lfpm, i)
            o = ctx.solve_frame(wl)
            tiles = orc.slice_u8(wl.raw(), wl.L["origins"], wl.crop).astype(np.float32)
            o["mosaic"], _ = orc.disparity_mosaic(o["plane"], o["cost"], wl.L["centers"],
                                                  wl.crop, 0.2, 320, 256)
            o["vd"], o["color"], _ = orc.vd_map(o["plane"], wl.L["centers"], tiles.reshape(-1),
                                                wl.crop, wl.L["pitch"], 320, 256)
            exp.append(o)
    finally:
        ctx.close()
    return exp


def _compare(out, exp, i):
    b = str(out / ("frame%03d" % i))
    got = dict(plane=np.fromfile(b + "_plane.bin", np.float32).reshape(-1, 4),
               cost=np.fromfile(b + "_cost.bin", np.float32),
               sel=np.fromfile(b + "_sel.bin", np.uint32),
               reason=np.fromfile(b + "_reason.bin", np.uint8),
               mosaic=np.fromfile(b + "_mosaic.bin", np.float32).reshape(256, 320),
               vd=np.fromfile(b + "_vd.bin", np.float32).reshape(256, 320),
               color=np.fromfile(b + "_color.bin", np.uint8).reshape(256, 320))
    for k, v in got.items():
        assert np.array_equal(v.view(np.uint8), exp[k].view(np.uint8)), (i, k)


@pytest.mark.parametrize("mode", [0, 1], ids=["sequence", "batch"])
def test_facade_pipelines(lfpm, expected, mode, tmp_path):
    wl = _frame_workload(lfpm, 0)
    out = _run(lfpm, wl, 2, mode, tmp_path)
    for i in range(2):
        _compare(out, expected[i], i)


def test_facade_adaptmipm_single_mi(lfpm, orc, tmp_path):
    wl = _frame_workload(lfpm, 0)
    pm = copy.copy(wl.pm)
    out = _run(lfpm, wl, 1, 2, tmp_path)
    V, K, px = np.fromfile(str(out / "mi_meta.bin"), np.int32)
    tiles = np.fromfile(str(out / "mi_tiles.bin"), np.uint8)
    centers = np.fromfile(str(out / "mi_centers.bin"), np.float32).reshape(-1, 2)
    nbr = np.fromfile(str(out / "mi_nbr.bin"), np.int32)
    cnt = np.fromfile(str(out / "mi_cnt.bin"), np.int32)
    plane = np.fromfile(str(out / "mi_plane.bin"), np.float32).reshape(-1, 4)
    assert V == centers.shape[0] and cnt[0] == V - 1 and K == pm.max_neighbors
    ctx = lfpm.Context()
    try:
        ref = ctx.solve_tiles(tiles, centers, nbr, cnt, pm)
    finally:
        ctx.close()
    assert np.array_equal(plane.view(np.uint32), ref["plane"][:px].view(np.uint32))
    # AdaptMIPMPFPGR: both stages give the per-MI solve; no proxy is collected
    for f in ("pf_collect_plane.bin", "pf_repair_plane.bin"):
        pf = np.fromfile(str(out / f), np.float32).reshape(-1, 4)
        assert np.array_equal(pf.view(np.uint32), plane.view(np.uint32)), f
    assert np.all(np.fromfile(str(out / "pf_pgr.bin"), np.int32) == -1)
    o = orc.solve(tiles.reshape(V, wl.crop, wl.crop), centers, nbr.reshape(V, K), cnt,
                  orc.PmParams(*[getattr(pm, f[0]) for f in orc.PmParams._fields_]))
    err = np.abs(plane[:, 3].astype(np.float64) - o["plane"][:px, 3])
    _say("AdaptMIPM views", V, "frac<=0.05px", float((err <= 0.05).mean()))
    assert (err <= 0.05).mean() >= 0.99


def test_facade_acmm_matcher(lfpm, tmp_path):
    """DepthSolverPipelineBySequence with LightFieldParams::frame_matcher_acmm (the
    AdaptMIPMFrameACMM matcher) gives lfpm_ctx_solve_frame's ACMM results, filter included,
    bit for bit, on a two-level layout."""
    wl = lfpm.workload_c1(raw_w=560, raw_h=448, pitch=35.0, scene_seed=1234, noise_seed=1235)
    wl.pm.frame_uid = 5
    out = _run(lfpm, wl, 1, 3, tmp_path)
    ctx = lfpm.Context()
    try:
        ctx.set_variant(lfpm.Context.ACMM)
        ctx.set_filter_order("reference")
        exp = ctx.solve_frame(wl)
    finally:
        ctx.close()
    b = str(out / "frame000")
    got = dict(plane=np.fromfile(b + "_plane.bin", np.float32).reshape(-1, 4),
               cost=np.fromfile(b + "_cost.bin", np.float32),
               sel=np.fromfile(b + "_sel.bin", np.uint32),
               reason=np.fromfile(b + "_reason.bin", np.uint8))
    for k, v in got.items():
        assert np.array_equal(v.view(np.uint8), exp[k].view(np.uint8)), k
