"""Host-level context API on the GPU: the batched lfpm_ctx_solve_frames equals one
lfpm_ctx_solve_frame per frame (frame_uid + i) bit for bit, including the CheckModule reasons,
and lfpm_ctx_frame_products of any frame of the batch equals the oracle's mosaic and
virtual-depth map of that frame's filtered results."""
import copy

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def test_ctx_batched_frames_match_single(lfpm, orc):
    import torch
    assert torch.cuda.is_available()
    wls = [lfpm.workload_c1(raw_w=320, raw_h=256, scene_seed=900 + i, noise_seed=950 + i)
           for i in range(3)]
    wl = wls[0]
    _say("case ctx batched 3 frames V=%d" % wl.V)
    ctx = lfpm.Context()
    try:
        raws = [w.raw() for w in wls]
        outs = ctx.solve_frames(wl, raws)
        prods = [ctx.frame_products(320, 256, wl.L["pitch"], frame=i) for i in range(3)]
        for i in range(3):
            pm = copy.copy(wl.pm)
            pm.frame_uid = wl.pm.frame_uid + i
            one = ctx.solve_frame(wl, raws[i], pm=pm)
            for k in one:
                assert np.array_equal(one[k].view(np.uint8), outs[i][k].view(np.uint8)), (i, k)
            # products of the batched frame i vs the oracle on its filtered result
            tiles = orc.slice_u8(raws[i], wl.L["origins"], wl.crop).astype(np.float32)
            om, omr = orc.disparity_mosaic(outs[i]["plane"], outs[i]["cost"], wl.L["centers"],
                                           wl.crop, 0.2, 320, 256)
            ovd, ocol, ovr = orc.vd_map(outs[i]["plane"], wl.L["centers"], tiles.reshape(-1),
                                        wl.crop, wl.L["pitch"], 320, 256)
            p = prods[i]
            assert np.array_equal(p["mosaic"].view(np.uint32), om.view(np.uint32)), i
            assert np.array_equal(p["vd"].view(np.uint32), ovd.view(np.uint32)), i
            assert np.array_equal(p["color"], ocol), i
            assert np.array_equal(p["mosaic_range"], omr) and np.array_equal(p["vd_range"], ovr)
            _say("frame", i, "mosaic px", int(np.count_nonzero(p["mosaic"])), "ok")
    finally:
        ctx.close()


def test_ctx_solve_tiles_matches_frame(lfpm):
    """Caller-supplied crops + neighbour table (the MLA_Problem data model) give the frame
    solve's results when they are the frame's own crops and table."""
    import torch
    assert torch.cuda.is_available()
    import oracle as orc
    wl = lfpm.workload_c1(raw_w=320, raw_h=256)
    _say("case ctx solve_tiles V=%d" % wl.V)
    ctx = lfpm.Context()
    try:
        whole = ctx.solve_frame(wl)
        tiles = orc.slice_u8(wl.raw(), wl.L["origins"], wl.crop)
        part = ctx.solve_tiles(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"], wl.pm, wl.fp)
    finally:
        ctx.close()
    for k in whole:
        assert np.array_equal(whole[k].view(np.uint8), part[k].view(np.uint8)), k


def test_ctx_acmm_variant_matches_device_entry(lfpm):
    """lfpm_ctx_set_variant(ACMM) + lfpm_ctx_solve_tiles (the AdaptMIPMFrameACMM seam) equals
    lfpm_patchmatch_frame_acmm on the same crops bit for bit; switching back to PLAIN gives the
    plain solve again."""
    import torch
    assert torch.cuda.is_available()
    import oracle as orc
    wl = lfpm.workload_c1(raw_w=560, raw_h=448, pitch=35.0)
    _say("case ctx acmm variant V=%d crop=%d" % (wl.V, wl.crop))
    df = lfpm.DeviceFrame(wl)
    bufs = lfpm.patchmatch_frame_acmm(df)
    torch.cuda.synchronize()
    tiles = orc.slice_u8(wl.raw(), wl.L["origins"], wl.crop)
    ctx = lfpm.Context()
    try:
        plain0 = ctx.solve_tiles(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"], wl.pm)
        ctx.set_variant(lfpm.Context.ACMM)
        ac = ctx.solve_tiles(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"], wl.pm)
        ctx.set_variant(lfpm.Context.PLAIN)
        plain1 = ctx.solve_tiles(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"], wl.pm)
        with pytest.raises(lfpm.LfpmError):
            ctx.set_variant(7)
    finally:
        ctx.close()
    for k, dev in (("plane", bufs.plane), ("cost", bufs.cost), ("disp", bufs.disp),
                   ("sel", bufs.sel)):
        assert np.array_equal(ac[k].view(np.uint8), dev.cpu().numpy().view(np.uint8)), k
    for k in plain0:
        assert np.array_equal(plain0[k].view(np.uint8), plain1[k].view(np.uint8)), k
    assert not np.array_equal(plain0["plane"], ac["plane"])
