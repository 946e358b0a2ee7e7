"""The served form behind the C ABI (lfpm_server_create / lfpm_frame_submit / lfpm_frame_wait):
frames of several parameter classes arriving on a Poisson stream are admitted while others are on
the device, share launches with their class and overlap with the other classes, and each result
(filtered plane / cost / disp / sel, reject mask, refocus stack) equals the same frame solved
alone, bit for bit.  A frame larger than the HBM budget is refused with LFPM_E_NOMEM."""
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
        lfpm.workload_c1(raw_w=320, raw_h=256, slices=4, frame_uid=11),
        lfpm.workload_c2(raw_w=400, raw_h=320, iterations=3, slices=3, frame_uid=12),
        lfpm.workload_c2(raw_w=300, raw_h=260, pitch=15.0, radius=5, lens_types=3,
                         per_type=(6, 12, 18), iterations=2, slices=2, frame_uid=13),
        lfpm.workload_c1(raw_w=320, raw_h=256, slices=4, seed=77, scene_seed=77, noise_seed=78,
                         frame_uid=14),
        lfpm.workload_c2(raw_w=330, raw_h=280, pitch=31.0, radius=9, iterations=2, slices=2,
                         frame_uid=15),
        lfpm.workload_c2(raw_w=400, raw_h=320, iterations=3, slices=3, seed=5, scene_seed=6,
                         noise_seed=7, frame_uid=16),
    ]


def _solo(lfpm, torch, wl, order="snapshot"):
    df = lfpm.DeviceFrame(wl)
    bufs = lfpm.patchmatch_frame(df)
    fo = lfpm.filter_frame(df, bufs, order=order)
    rf = lfpm.refocus(df, slices=wl.slices)
    torch.cuda.synchronize()
    return dict(plane=fo["plane"].cpu().numpy(), cost=fo["cost"].cpu().numpy(),
                disp=fo["disp"].cpu().numpy(), sel=fo["sel"].cpu().numpy().view(np.uint32),
                reason=fo["reason"].cpu().numpy(), refocus=rf.cpu().numpy())


def _serve(lfpm, torch, arrivals, order="snapshot"):
    import serving
    wls = _workloads(lfpm)
    solo = [_solo(lfpm, torch, wl, order) for wl in wls]
    reqs = [serving.Request(i, float(arrivals[i]), wl) for i, wl in enumerate(wls)]
    srv = serving.Server(hbm_budget_bytes=8 << 30, max_slots=8, streams=4, filter_order=order)
    try:
        _say("case serve: %d requests, MIs %s" % (len(reqs), [wl.V for wl in wls]))
        done = srv.run(reqs)
        st = srv.stats()
    finally:
        srv.close()
    _say("serve stats", st, "latency ms", [round(r.latency_ms, 1) for r in done])
    for r in done:
        ref = solo[r.rid]
        for k in ("plane", "cost", "disp", "sel", "reason", "refocus"):
            a, b = r.result[k], ref[k]
            assert a.shape == b.shape, (r.rid, k)
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (r.rid, k)
        assert r.latency_ms > 0 and r.queue_ms >= 0
    assert st["completed"] == len(wls) and st["classes"] >= 4
    stages = sum(2 + 2 * wl.pm.iterations for wl in wls)
    return st, stages


def test_served_poisson_stream_equals_solo(lfpm, torch_cuda):
    """A Poisson stream (1000 req/s, seed 7): every result equals its solo solve."""
    rng = np.random.default_rng(7)
    _serve(lfpm, torch_cuda, np.cumsum(rng.exponential(1.0 / 1000.0, 6)))


def test_served_burst_shares_launches(lfpm, torch_cuda):
    """Six requests at once: frames of one class share their stage launches (fewer launches
    than frame-stages) and every result still equals its solo solve."""
    st, stages = _serve(lfpm, torch_cuda, np.zeros(6))
    assert st["max_active"] == 6
    assert st["launches"] < stages


def test_served_reference_order_equals_solo(lfpm, torch_cuda):
    """A server created with the reference filter order (lfpm_serve_opts.filter_order): every
    request's reject mask and filtered outputs equal its solo solve + ordered pass."""
    _serve(lfpm, torch_cuda, np.zeros(6), order="reference")


def test_served_frame_over_budget_is_refused(lfpm, torch_cuda):
    import serving
    wl = lfpm.workload_c1(raw_w=320, raw_h=256, slices=2)
    srv = serving.Server(hbm_budget_bytes=1 << 20, max_slots=4)
    try:
        r = serving.Request(0, 0.0, wl)
        srv.submit(r)
        with pytest.raises(lfpm.LfpmError) as ei:
            srv.wait(r)
        assert ei.value.code == -3
    finally:
        srv.close()
