"""Batched and served forms on the GPU: a frame's results are the same alone and in a mixed batch
(bit for bit — the random stream is keyed by (seed, frame_uid), never by batch position), through
the eager schedule, the captured hipGraph (replayed several times with eager work in between) and
the continuous-batching step API with staggered frames; autotuning changes no result."""
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


def _workloads(lfpm):
    return [
        lfpm.workload_c1(raw_w=320, raw_h=256, slices=4, frame_uid=1),
        lfpm.workload_c1(raw_w=256, raw_h=200, slices=4, frame_uid=2, seed=77),
        lfpm.workload_c2(raw_w=300, raw_h=260, pitch=15.0, lens_types=3, per_type=(6, 12, 18),
                         iterations=2, slices=3, frame_uid=3),
        lfpm.workload_c2(raw_w=330, raw_h=280, pitch=31.0, iterations=2, slices=3, frame_uid=4),
    ]


def _plain(lfpm, torch, slot):
    bufs = lfpm.patchmatch_frame(slot.df)
    fo = lfpm.filter_frame(slot.df, bufs)
    rf = lfpm.refocus(slot.df, slices=slot.slices)
    torch.cuda.synchronize()
    out = dict(plane=bufs.plane, cost=bufs.cost, sel=bufs.sel, disp=bufs.disp,
               reason=fo["reason"], fplane=fo["plane"], refocus=rf)
    return {k: v.cpu().numpy().copy() for k, v in out.items()}


def _batched(slot):
    out = dict(plane=slot.plane, cost=slot.cost, sel=slot.sel, disp=slot.disp,
               reason=slot.reason, fplane=slot.f_plane, refocus=slot.refocus)
    return {k: v.cpu().numpy().copy() for k, v in out.items()}


def _same(a, b, tag):
    for k in a:
        assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), (tag, k)


@pytest.fixture(scope="module")
def setup(lfpm, torch_cuda):
    slots = [lfpm.FrameSlot(wl) for wl in _workloads(lfpm)]
    ref = [_plain(lfpm, torch_cuda, s) for s in slots]
    return slots, ref


def test_batch_eager_matches_plain(lfpm, torch_cuda, setup):
    slots, ref = setup
    _say("case batch eager: %d frames, MIs %s" % (len(slots), [s.wl.V for s in slots]))
    b = lfpm.Batch(slots)
    b.run()
    torch_cuda.cuda.synchronize()
    for i, s in enumerate(slots):
        _same(ref[i], _batched(s), "frame %d" % i)


def test_batch_graph_replays_match(lfpm, torch_cuda, setup):
    slots, ref = setup
    _say("case batch graph")
    b = lfpm.Batch(slots)
    b.capture()
    try:
        for rep in range(3):
            for s in slots:  # eager work between replays: scribble over every output
                for t in (s.plane, s.cost, s.disp, s.f_plane, s.refocus):
                    t.fill_(-7.0)
                s.reason.fill_(9)
            torch_cuda.cuda.synchronize()
            b.replay()
            torch_cuda.cuda.synchronize()
            for i, s in enumerate(slots):
                _same(ref[i], _batched(s), "replay %d frame %d" % (rep, i))
    finally:
        b.close()


def test_serving_steps_staggered(lfpm, torch_cuda, setup):
    """Frames join at different steps (continuous batching); each still matches its solo run."""
    slots, ref = setup
    _say("case serving steps")
    b = lfpm.Batch(slots)
    n = len(slots)
    start = [0, 2, 5, 1]
    # per-frame schedule: init, iterations x (phase 0, phase 1), final
    sched = []
    for s in slots:
        q = [(lfpm.STAGE_INIT, 0, 0)]
        for it in range(s.pm.iterations):
            q += [(lfpm.STAGE_PHASE, it, 0), (lfpm.STAGE_PHASE, it, 1)]
        q.append((lfpm.STAGE_FINAL, 0, 0))
        sched.append(q)
    total = max(start[i] + len(sched[i]) for i in range(n))
    for t in range(total):
        st, it, ph = [], [], []
        for i in range(n):
            k = t - start[i]
            e = sched[i][k] if 0 <= k < len(sched[i]) else (lfpm.STAGE_IDLE, 0, 0)
            st.append(e[0]); it.append(e[1]); ph.append(e[2])
        b.step(st, it, ph)
    torch_cuda.cuda.synchronize()
    for i, s in enumerate(slots):
        _same(ref[i], _batched(s), "served frame %d" % i)


def test_autotune_keeps_results(lfpm, torch_cuda, setup):
    """The autotuner searches wavefronts per workgroup x workgroups per CU; the solve under the
    chosen launch shape equals the one under the default shape bit for bit."""
    slots, ref = setup
    b = lfpm.Batch(slots[:1])
    best, ms = b.autotune(0)
    shape = lfpm.autotune_get(slots[0].wl.pm)
    _say("autotune", best, ms, "shape", shape)
    assert best in (2, 4, 6, 8) and ms > 0
    # wavefronts per workgroup: the searched 1, 2, 4, 8, 12, or 6 where 8 do not fit the LDS
    assert shape[0] == best and shape[1] in (1, 2, 4, 6, 8, 12)
    again = _plain(lfpm, torch_cuda, slots[0])
    _same(ref[0], again, "after tune")


def test_default_launch_shape_keeps_wavefronts(lfpm, torch_cuda):
    """Until tuned, a class runs the workgroup size that keeps the most wavefronts per CU: three
    4-wavefront workgroups for crop 16 (48 KB of LDS each), one 12-wavefront workgroup for
    crop 21 with a 19 x 19 window (138 KB: a workgroup may use the device's whole LDS)."""
    c16 = lfpm.workload_c2(raw_w=320, raw_h=256)
    c21 = lfpm.workload_c2(raw_w=330, raw_h=280, pitch=31.0, radius=9)
    _say("case default launch shapes: crop %d, crop %d" % (c16.crop, c21.crop))
    assert lfpm.autotune_get(c16.pm)[1] == 4
    assert lfpm.autotune_get(c21.pm)[1] == 12


def test_batch_blur_term_matches_plain(lfpm, torch_cuda):
    """The blur term in a batch (Laplacian blur tiles after the intensity tiles): the batched
    solve equals the single-frame solve bit for bit, and both differ from the solve without it."""
    import copy
    wl = lfpm.workload_c2(raw_w=400, raw_h=320, iterations=2, slices=2, frame_uid=5)
    pm = copy.copy(wl.pm)
    pm.blur_lambda = 0.5
    _say("case batch blur: V=%d" % wl.V)
    slot = lfpm.FrameSlot(wl, filt=False, refocus=False, blur_raw="laplacian", pm=pm)
    b = lfpm.Batch([slot])
    b.run()
    torch_cuda.cuda.synchronize()
    plain = lfpm.patchmatch_frame(slot.df, pm=pm)
    off = lfpm.patchmatch_frame(slot.df)
    torch_cuda.cuda.synchronize()
    for k in ("plane", "cost", "sel", "disp"):
        a = getattr(slot, k).cpu().numpy()
        assert np.array_equal(a.view(np.uint8), getattr(plain, k).cpu().numpy().view(np.uint8)), k
    assert not np.array_equal(slot.cost.cpu().numpy(), off.cost.cpu().numpy())

