"""The CheckModule pass in the reference's own MI order (lfpm_filter_frame_ordered) on the GPU,
against the FP32 twin of that order (oracle filter_f32_inplace): MIs in index order, each
reading the filtered states its smaller-index neighbours left (MIDisparityFilterCPU::
FilterGlobal, CheckModule/MIDisparityFilter.cpp:973-985, invalidation :671-686).

- the 320 x 256 frame at the small-frame parameters and the 4080 x 3068 frame at the full-frame
  parameters, one launch per level of the schedule: reason masks and outputs bit-exact vs the
  twin, and the result differs from the snapshot pass (rejections cascade);
- a three-lens-type layout (asymmetric neighbour table), a crop-21 frame and eight checked views
  per pixel likewise, the snapshot pass against its own twin each time;
- two calls on the full frame are identical;
- the ordered pass captured in a batch graph replays bit-exactly.
"""
import time

import numpy as np
import pytest

from _helpers import fp_to_oracle

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _host(d):
    return {k: v.cpu().numpy() for k, v in d.items()}


def _check_ordered(lfpm, orc, torch, wl, tag):
    df = lfpm.DeviceFrame(wl)
    bufs = lfpm.patchmatch_frame(df)
    snap = lfpm.filter_frame(df, bufs)
    lfpm.filter_levels(wl.L["nbr"], wl.L["cnt"], wl.pm.max_neighbors)
    torch.cuda.synchronize()
    t0 = time.time()
    ordr = lfpm.filter_frame(df, bufs, order="reference")
    torch.cuda.synchronize()
    _say("%s: V=%d, %d levels, ordered pass (first call, schedule built) + sync %.1f ms"
         % (tag, wl.V, len(df._filter_levels[1]) - 1, 1e3 * (time.time() - t0)))
    g = _host(ordr)
    s = _host(snap)
    plane = bufs.plane.cpu().numpy()
    cost = bufs.cost.cpu().numpy()
    disp = bufs.disp.cpu().numpy()
    sel = bufs.sel.cpu().numpy().view(np.uint32)
    o = orc.filter_f32_inplace(df.gray.cpu().numpy(), wl.L["centers"], wl.L["nbr"], wl.L["cnt"],
                               wl.crop, wl.pm.max_neighbors, wl.pm.base,
                               fp_to_oracle(orc, wl.fp), plane, cost, disp, sel)
    os_ = orc.filter_f32(df.gray.cpu().numpy(), wl.L["centers"], wl.L["nbr"], wl.L["cnt"],
                         wl.crop, wl.pm.max_neighbors, wl.pm.base, fp_to_oracle(orc, wl.fp),
                         plane, cost, disp, sel)
    for k in ("reason", "plane", "cost", "disp", "sel"):
        assert np.array_equal(s[k].view(np.uint8), os_[k].view(np.uint8)), ("snapshot", k)
    reasons = {int(r): int((g["reason"] == r).sum()) for r in (0, 80, 140, 200, 255)}
    differ = float((g["reason"] != s["reason"]).mean())
    _say("%s reasons" % tag, reasons, "differ from snapshot %.4f" % differ)
    assert np.array_equal(g["reason"], o["reason"])
    for k in ("plane", "cost", "disp"):
        assert np.array_equal(g[k].view(np.uint32), o[k].view(np.uint32)), k
    assert np.array_equal(g["sel"].view(np.uint32), o["sel"])
    # the orders differ only by cascades: MI 0 reads no filtered state
    n0 = wl.crop * wl.crop
    assert np.array_equal(g["reason"][:n0], s["reason"][:n0])
    assert differ > 0.0
    return df, bufs, g


def test_filter_ordered_small_frame(lfpm, orc, torch_cuda):
    wl = lfpm.workload_c1(raw_w=320, raw_h=256)
    _say("case filter ordered 320x256")
    _check_ordered(lfpm, orc, torch_cuda, wl, "ordered c1")


def test_filter_ordered_three_lens_types(lfpm, orc, torch_cuda):
    """Per-type neighbour caps (6, 12, 18) make the neighbour table asymmetric: an MI can read a
    neighbour that does not read it back."""
    wl = lfpm.workload_c2(raw_w=300, raw_h=260, pitch=15.0, radius=5, lens_types=3,
                          per_type=(6, 12, 18), iterations=2)
    _say("case filter ordered, 3 lens types")
    _check_ordered(lfpm, orc, torch_cuda, wl, "ordered 3 types")


def test_filter_ordered_full_frame(lfpm, orc, torch_cuda):
    wl = lfpm.workload_c2()
    _say("case filter ordered 4080x3068")
    df, bufs, g = _check_ordered(lfpm, orc, torch_cuda, wl, "ordered c2")
    again = _host(lfpm.filter_frame(df, bufs, order="reference"))
    torch_cuda.cuda.synchronize()
    for k in g:
        assert np.array_equal(g[k].view(np.uint8), again[k].view(np.uint8)), k


def test_ctx_and_batch_reference_order(lfpm, torch_cuda):
    """The context (lfpm_ctx_set_filter_order) and the batched form (lfpm_batch_frame's schedule
    fields) run the same ordered pass: ctx.solve_frame, ctx.solve_frames and a Batch of
    FrameSlots with filter_order="reference" equal the plain solve + ordered filter bit for
    bit; switching the context back to snapshot gives the snapshot pass."""
    wl = lfpm.workload_c1(raw_w=320, raw_h=256, frame_uid=11)
    _say("case ctx / batch reference order: V=%d" % wl.V)
    df = lfpm.DeviceFrame(wl)
    bufs = lfpm.patchmatch_frame(df)
    ref = _host(lfpm.filter_frame(df, bufs, order="reference"))
    snap = _host(lfpm.filter_frame(df, bufs))
    torch_cuda.cuda.synchronize()
    ctx = lfpm.Context()
    try:
        ctx.set_filter_order("reference")
        one = ctx.solve_frame(wl)
        many = ctx.solve_frames(wl, [wl.raw(), wl.raw()])
        ctx.set_filter_order("snapshot")
        back = ctx.solve_frame(wl)
    finally:
        ctx.close()
    for k in ("plane", "cost", "sel", "reason"):
        assert np.array_equal(one[k].view(np.uint8), ref[k].view(np.uint8)), k
        assert np.array_equal(back[k].view(np.uint8), snap[k].view(np.uint8)), k
        # frame 0 of the batch has frame_uid = wl.pm.frame_uid, the single-frame solve's
        assert np.array_equal(many[0][k].view(np.uint8), ref[k].view(np.uint8)), k
    slot = lfpm.FrameSlot(wl, refocus=False, filter_order="reference")
    b = lfpm.Batch([slot])
    b.run()
    torch_cuda.cuda.synchronize()
    got = dict(plane=slot.f_plane, cost=slot.f_cost, sel=slot.f_sel, reason=slot.reason)
    for k, v in got.items():
        assert np.array_equal(v.cpu().numpy().view(np.uint8), ref[k].view(np.uint8)), k


def test_filter_ordered_higher_index_neighbour_earlier(lfpm, orc, torch_cuda):
    """An edited neighbour table in which odd MIs keep only their higher-index neighbours: those
    MIs sit in level 0, so an even MI reads higher-index neighbours that the schedule filters
    before it (or beside it).  The pass must still give them their pre-filter values, as the
    reference's map order does: bit-exact vs the in-place twin on the same table."""
    wl = lfpm.workload_c1(raw_w=320, raw_h=256, frame_uid=21)
    V, K = wl.V, wl.pm.max_neighbors
    nbr = np.asarray(wl.L["nbr"]).reshape(V, K).copy()
    cnt = np.asarray(wl.L["cnt"]).copy()
    for v in range(1, V, 2):
        keep = [int(n) for n in nbr[v, :cnt[v]] if n > v]
        if keep:
            nbr[v] = -1
            nbr[v, :len(keep)] = keep
            cnt[v] = len(keep)
    wl.L["nbr"] = nbr.reshape(np.asarray(wl.L["nbr"]).shape)
    wl.L["cnt"] = cnt.astype(np.asarray(wl.L["cnt"]).dtype)
    order, start = lfpm.filter_levels(wl.L["nbr"], wl.L["cnt"], K)
    lev = np.empty(V, np.int64)
    for l in range(len(start) - 1):
        lev[order[start[l]:start[l + 1]]] = l
    early = sum(1 for v in range(V) for n in nbr[v, :cnt[v]] if n > v and lev[n] <= lev[v])
    _say("case filter ordered, edited table: %d reads of a higher-index MI filtered no later"
         % early)
    assert early > 0
    _check_ordered(lfpm, orc, torch_cuda, wl, "ordered edited table")


@pytest.mark.parametrize("case", ["crop21", "eight_views"])
def test_filter_orders_other_shapes(lfpm, orc, torch_cuda, case):
    """Both orders vs their twins on a crop-21 frame (448-thread workgroups, 19 x 19 window) and
    with every one of the first 8 neighbours checked (max_triplet 8, selection ignored: four
    rounds of paired views per pixel)."""
    if case == "crop21":
        wl = lfpm.workload_c2(raw_w=330, raw_h=280, pitch=31.0, radius=9, iterations=2)
    else:
        wl = lfpm.workload_c2(raw_w=400, raw_h=320, iterations=2)
        wl.fp.max_triplet = 8
        wl.fp.use_selected_views_only = 0
    _say("case filter orders, %s: crop %d, V=%d" % (case, wl.crop, wl.V))
    _check_ordered(lfpm, orc, torch_cuda, wl, "ordered " + case)


def test_batch_graph_reference_order_replays(lfpm, torch_cuda):
    """The ordered pass inside a captured batch graph (one kernel node per level): three
    replays, with the outputs scribbled over between them, each equal the plain solve + ordered
    pass bit for bit."""
    wl = lfpm.workload_c1(raw_w=320, raw_h=256, frame_uid=31)
    _say("case batch graph, reference order: V=%d" % wl.V)
    df = lfpm.DeviceFrame(wl)
    bufs = lfpm.patchmatch_frame(df)
    ref = _host(lfpm.filter_frame(df, bufs, order="reference"))
    torch_cuda.cuda.synchronize()
    slot = lfpm.FrameSlot(wl, refocus=False, filter_order="reference")
    b = lfpm.Batch([slot])
    b.capture()
    try:
        for rep in range(3):
            for t in (slot.f_plane, slot.f_cost, slot.plane, slot.cost):
                t.fill_(-7.0)
            slot.reason.fill_(9)
            torch_cuda.cuda.synchronize()
            b.replay()
            torch_cuda.cuda.synchronize()
            got = dict(plane=slot.f_plane, cost=slot.f_cost, sel=slot.f_sel, reason=slot.reason)
            for k, v in got.items():
                assert np.array_equal(v.cpu().numpy().view(np.uint8),
                                      ref[k].view(np.uint8)), (rep, k)
    finally:
        b.close()
