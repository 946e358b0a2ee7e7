"""Parity at the full-frame configuration's parameters (18 ring-1 + ring-2 views, 15 x 15
bilateral window with sigma_s 5 / sigma_c 0.12, 6 red-black iterations, cycle check at 0.5 px,
32 refocus slices), GPU (through the C ABI) against the CPU oracle:

- full solve on a 1024 x 800 frame vs the FP64 oracle, FP32 tiles and FP16 tiles, each at the
  bar of >= 99.5 % of valid pixels within 0.05 px (RunPatchMatchCUDAForFrame_Impl,
  MIStereo/AdaptMIPMFrame.cu:949-995);
- the fused CheckModule pass on that solve, bit-exact vs the FP32 twin;
- the 32-slice refocus of that frame vs the FP64 oracle;
- the 4080 x 3068 frame: two solves bit-identical, and the frame batched with a second frame
  equal to its plain solve bit for bit;
- the precision set: 10^6 seeded cone hypotheses x 18 views on the 4080 x 3068 frame, FP32 and
  FP16 tiles vs the FP64 oracle at the exact cost bar of _helpers.cost_bar.
"""
import copy
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


def _solve(lfpm, torch, df, pm):
    bufs = lfpm.patchmatch_frame(df, pm=pm)
    torch.cuda.synchronize()
    return bufs, dict(plane=bufs.plane.cpu().numpy(), cost=bufs.cost.cpu().numpy(),
                      disp=bufs.disp.cpu().numpy(), sel=bufs.sel.cpu().numpy().view(np.uint32))


@pytest.fixture(scope="module")
def c2sub(lfpm, orc, torch_cuda):
    """The 1024 x 800 frame at full-frame parameters: its device frame and the FP64 oracle
    solve (shared by the FP32 and FP16 comparisons)."""
    wl = lfpm.workload_c2(raw_w=1024, raw_h=800)
    _say("case c2 parameters 1024x800: V=%d K=%d window=%d iterations=%d"
         % (wl.V, wl.pm.max_neighbors, 2 * wl.pm.radius + 1, wl.pm.iterations))
    df = lfpm.DeviceFrame(wl)
    tiles = orc.slice_u8(wl.raw(), wl.L["origins"], wl.crop)
    t0 = time.time()
    o = orc.solve(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"], pm_to_oracle(orc, wl.pm))
    _say("oracle solve %.1f s" % (time.time() - t0))
    return wl, df, tiles, o


def _disp_parity(wl, g, o):
    valid = o["cost"] <= 1.8
    d_g = g["plane"][:, 3].astype(np.float64)
    d_o = o["plane"][:, 3]
    err = np.abs(d_g - d_o)
    W = wl.crop
    gt = wl.gt_virtual_depth()
    gtv = np.stack([gt[y:y + W, x:x + W] for x, y in wl.L["origins"]]).reshape(-1)
    gtd = wl.L["pitch"] / np.maximum(gtv, 1e-6)
    return dict(V=wl.V, valid=float(valid.mean()),
                frac_le_005=float((err[valid] <= 0.05).mean()),
                frac_all_le_005=float((err <= 0.05).mean()),
                sel_equal=float((g["sel"] == o["sel"]).mean()),
                gt_med_err_gpu=float(np.median(np.abs(d_g - gtd))),
                gt_med_err_orc=float(np.median(np.abs(d_o - gtd))))


@pytest.mark.parametrize("storage", [0, 1], ids=["fp32", "fp16"])
def test_solve_c2params_vs_fp64(lfpm, torch_cuda, c2sub, storage):
    wl, df, _, o = c2sub
    pm = copy.copy(wl.pm)
    pm.storage = storage
    _say("case solve c2 parameters storage=%d" % storage)
    _, g = _solve(lfpm, torch_cuda, df, pm)
    st = _disp_parity(wl, g, o)
    st["storage"] = storage
    _say("solve_c2params", st)
    stats_dump("solve_c2params_%s" % ("fp16" if storage else "fp32"), st)
    assert st["frac_le_005"] >= 0.995, st


def test_filter_c2params_bitexact(lfpm, orc, torch_cuda, c2sub):
    """CheckModule (cycle check at 0.5 px, photo check, spike filter) on the GPU solve: masks
    and outputs equal the FP32 twin's bit for bit."""
    wl, df, _, _ = c2sub
    _say("case filter c2 parameters")
    bufs, g = _solve(lfpm, torch_cuda, df, wl.pm)
    fo = lfpm.filter_frame(df, bufs)
    torch_cuda.cuda.synchronize()
    gf = {k: v.cpu().numpy() for k, v in fo.items()}
    of = orc.filter_f32(df.gray.cpu().numpy(), wl.L["centers"], wl.L["nbr"], wl.L["cnt"],
                        wl.crop, wl.pm.max_neighbors, wl.pm.base, fp_to_oracle(orc, wl.fp),
                        g["plane"], g["cost"], g["disp"], g["sel"])
    reasons = {int(r): int((gf["reason"] == r).sum()) for r in (0, 80, 140, 200, 255)}
    _say("filter c2 reasons", reasons)
    stats_dump("filter_c2params", dict(reasons=reasons))
    assert sum(v for r, v in reasons.items() if r) > 0 and reasons[0] > 0
    assert np.array_equal(gf["reason"], of["reason"])
    for k in ("plane", "cost", "disp"):
        assert np.array_equal(gf[k].view(np.uint32), of[k].view(np.uint32)), k
    assert np.array_equal(gf["sel"].view(np.uint32), of["sel"])


def test_refocus_32_slices_vs_fp64(lfpm, orc, torch_cuda, c2sub):
    wl, df, tiles, _ = c2sub
    _say("case refocus 32 slices 1024x800")
    out = lfpm.refocus(df, slices=32, vmin=2.0, vmax=8.0)
    torch_cuda.cuda.synchronize()
    g = out.cpu().numpy().astype(np.float64)
    o = orc.refocus(tiles, wl.L["centers"], wl.L["pitch"], wl.crop, 1024, 800, 32, 2.0, 8.0)
    err = np.abs(g - o)
    st = dict(max_abs=float(err.max()), covered=float((o > 0).mean()))
    _say("refocus32", st)
    stats_dump("refocus_32", st)
    assert st["covered"] > 0.5
    assert err.max() <= 1e-5


@pytest.fixture(scope="module")
def full_frame(lfpm, torch_cuda):
    wl = lfpm.workload_c2()
    _say("case 4080x3068: V=%d" % wl.V)
    return wl, lfpm.DeviceFrame(wl)


def test_full_frame_deterministic_and_batched(lfpm, torch_cuda, full_frame):
    """4080 x 3068: two solves are bit-identical, and the frame solved in a batch with a second
    frame (its own seed and uid) equals its plain solve bit for bit."""
    wl, df = full_frame
    _, a = _solve(lfpm, torch_cuda, df, wl.pm)
    _, b = _solve(lfpm, torch_cuda, df, wl.pm)
    for k in a:
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k
    other = lfpm.workload_c2(seed=99, scene_seed=98, noise_seed=97, frame_uid=1)
    slots = [lfpm.FrameSlot(wl, filt=False, refocus=False),
             lfpm.FrameSlot(other, filt=False, refocus=False)]
    batch = lfpm.Batch(slots)
    batch.run()
    torch_cuda.cuda.synchronize()
    s = slots[0]
    got = dict(plane=s.plane.cpu().numpy(), cost=s.cost.cpu().numpy(),
               disp=s.disp.cpu().numpy(), sel=s.sel.cpu().numpy().view(np.uint32))
    for k in a:
        assert np.array_equal(got[k].view(np.uint32), a[k].view(np.uint32)), k
    _, c = _solve(lfpm, torch_cuda, slots[1].df, other.pm)
    assert np.array_equal(slots[1].plane.cpu().numpy().view(np.uint32),
                          c["plane"].view(np.uint32))
    del batch, slots


@pytest.mark.parametrize("storage", [0, 1], ids=["fp32", "fp16"])
def test_precision_set_1e6(lfpm, orc, torch_cuda, full_frame, storage):
    """10^6 cone hypotheses (normals within 60 degrees of +z, virtual depth in [2, 8]) x 18
    views on the 4080 x 3068 frame vs the FP64 oracle: every evaluation within 1e-5 (FP32 tiles)
    or 2e-3 (FP16 tiles) relative, except the straddles and ill-conditioned windows that
    cost_bar excuses."""
    wl, df = full_frame
    pm = copy.copy(wl.pm)
    pm.storage = storage
    _say("case precision 1e6 storage=%d" % storage)
    hyp = cone_hypotheses(lfpm, wl, 1000000, seed=2024)
    out = lfpm.eval_costs(df, torch_cuda.from_numpy(hyp).cuda(), pm=pm).cpu().numpy()
    tiles = orc.slice_u8(wl.raw(), wl.L["origins"], wl.crop)
    ref, geo, gate, ferr = orc.eval_margins(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"],
                                            pm_to_oracle(orc, wl.pm), hyp)
    tol = 2e-3 if storage else 1e-5
    st = dict(storage=storage, tol=tol, **cost_bar(out, ref, geo, gate, ferr, tol))
    _say("precision", st)
    stats_dump("precision_1e6_%s" % ("fp16" if storage else "fp32"), st)
    assert st["unexplained"] == 0, st
    assert st["excused_frac"] <= 0.15, st
