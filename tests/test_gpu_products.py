"""Per-frame products on the GPU vs the oracle, bit for bit: the raw-size disparity mosaic with
its range (DepthSolver::WriteDisMap_TileKey_new_AccuSequence, DepthSolver.cpp:6716-6897) and the
virtual-depth median splat with the filled, median-smoothed gray image
(DepthSolver::Virtual_depth_map_TileKeySequence, DepthSolver.cpp:7832-8000).  Inputs: a solved
frame (the GPU solve's own planes and costs) and hand-built edge cases (every pixel gated out,
non-positive disparities, colliding splats)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available()
    return torch


def _check(lfpm, orc, torch, df, plane, cost, tag, max_cost=0.2):
    wl = df.wl
    h, w = df.raw.shape
    m, mr = lfpm.disparity_mosaic(df, plane, cost, max_cost=max_cost)
    vd, col, vr = lfpm.virtual_depth_map(df, plane)
    torch.cuda.synchronize()
    p = plane.cpu().numpy()
    c = cost.cpu().numpy()
    g = df.gray.cpu().numpy().reshape(-1)
    om, omr = orc.disparity_mosaic(p, c, wl.L["centers"], wl.crop, max_cost, w, h)
    ovd, ocol, ovr = orc.vd_map(p, wl.L["centers"], g, wl.crop, wl.L["pitch"], w, h)
    m, vd, col = m.cpu().numpy(), vd.cpu().numpy(), col.cpu().numpy()
    _say(tag, "mosaic px", int(np.count_nonzero(m)), "vd px", int(np.count_nonzero(vd)),
         "ranges", mr.cpu().numpy().tolist(), vr.cpu().numpy().tolist())
    assert np.array_equal(m.view(np.uint32), om.view(np.uint32)), tag + " mosaic"
    assert np.array_equal(mr.cpu().numpy().view(np.uint32), omr.view(np.uint32)), tag
    assert np.array_equal(vd.view(np.uint32), ovd.view(np.uint32)), tag + " vd"
    assert np.array_equal(vr.cpu().numpy().view(np.uint32), ovr.view(np.uint32)), tag
    assert np.array_equal(col, ocol), tag + " colour"


def test_products_solved_c1(lfpm, orc, torch_cuda):
    wl = lfpm.workload_c1()
    _say("case products c1 V=%d" % wl.V)
    df = lfpm.DeviceFrame(wl)
    bufs = lfpm.patchmatch_frame(df)
    _check(lfpm, orc, torch_cuda, df, bufs.plane, bufs.cost, "c1")


def test_products_solved_mixed_pitch(lfpm, orc, torch_cuda):
    wl = lfpm.workload_c2(raw_w=720, raw_h=560, pitch=31.0, iterations=2)
    _say("case products pitch31 V=%d crop=%d" % (wl.V, wl.crop))
    df = lfpm.DeviceFrame(wl)
    bufs = lfpm.patchmatch_frame(df)
    _check(lfpm, orc, torch_cuda, df, bufs.plane, bufs.cost, "pitch31", max_cost=0.5)


def test_products_edge_cases(lfpm, orc, torch_cuda):
    torch = torch_cuda
    wl = lfpm.workload_c1(raw_w=320, raw_h=256)
    _say("case products edges V=%d" % wl.V)
    df = lfpm.DeviceFrame(wl)
    n = wl.V * wl.crop ** 2
    rng = np.random.default_rng(3)
    plane = np.zeros((n, 4), np.float32)
    plane[:, 2] = 1.0
    plane[:, 3] = rng.uniform(-2.0, 12.0, n).astype(np.float32)  # some d <= 0
    cost = np.full(n, 1.5, np.float32)  # all gated out of the mosaic
    _check(lfpm, orc, torch, df, torch.from_numpy(plane).cuda(), torch.from_numpy(cost).cuda(),
           "gated")
    plane[:, 3] = 23.0  # f_v = 1: dense splats, many collisions at MI centres
    cost[:] = 0.0
    _check(lfpm, orc, torch, df, torch.from_numpy(plane).cuda(), torch.from_numpy(cost).cuda(),
           "dense")


def test_products_bgr_colour(lfpm, orc, torch_cuda):
    """Colour VD map (lfpm_slice_bgr_tiles, lfpm_virtual_depth_map_bgr) and its coloured point
    cloud (lfpm_vd_point_cloud_bgr) on a solved frame with a colour raw image, bit for bit vs the
    oracle: VD values, the B, G, R image (per-channel medians, fill, ten 5x5 passes per channel)
    and the 27-byte records with red, green, blue from channels 2, 1, 0."""
    torch = torch_cuda
    wl = lfpm.workload_c1(raw_w=640, raw_h=512)
    _say("case products bgr V=%d" % wl.V)
    df = lfpm.DeviceFrame(wl)
    bufs = lfpm.patchmatch_frame(df)
    raw = wl.raw().astype(np.int64)
    rng = np.random.default_rng(21)
    bgr = np.stack([raw, (raw * 7 + 11) % 256, rng.integers(0, 256, raw.shape)], -1).astype(np.uint8)
    bgr_dev = torch.from_numpy(bgr).cuda()
    tiles = lfpm.slice_bgr_tiles(df, bgr_dev)
    vd, col, vr = lfpm.virtual_depth_map(df, bufs.plane, bgr_tiles=tiles)
    recs, n = lfpm.vd_point_cloud(vd, col, origin=(8.0, 8.0), pixel_size=0.0055,
                                  main_lens_f=50.0, B=1.3, bl0=70.0)
    torch.cuda.synchronize()
    p = bufs.plane.cpu().numpy()
    otiles = orc.slice_bgr(bgr, wl.L["origins"], wl.crop)
    assert np.array_equal(tiles.cpu().numpy(), otiles)
    ovd, ocol, ovr = orc.vd_map_bgr(p, wl.L["centers"], otiles, wl.crop, wl.L["pitch"], 640, 512)
    g_vd, g_col = vd.cpu().numpy(), col.cpu().numpy()
    assert np.array_equal(g_vd.view(np.uint32), ovd.view(np.uint32))
    assert np.array_equal(vr.cpu().numpy().view(np.uint32), ovr.view(np.uint32))
    assert np.array_equal(g_col, ocol)
    orec = orc.vd_point_cloud(ovd, ocol, (8.0, 8.0), 0.0055, 50.0, 1.3, 70.0)
    assert n == orec.shape[0] > 1000
    assert np.array_equal(recs.cpu().numpy(), orec)
    assert (ocol[..., 0] != ocol[..., 2]).mean() > 0.3
