"""lfpm_eval_costs_noalloc (the allocation-free evaluation, its segment table in a library-owned
device array) against lfpm_eval_costs through the C ABI: bit-identical costs on two calls in a
row, the second on a smaller frame with fewer hypotheses (the array keeps the first call's
longer table), and the size limit rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _eval(lfpm, torch, fn, df, hyp_dev, n):
    wl = df.wl
    K = wl.pm.max_neighbors
    out = torch.full((n, K), -1.0, dtype=torch.float32, device="cuda")
    L, dp = lfpm.lib(), lfpm._dp
    rc = fn(dp(df.tiles), dp(df.centers), dp(df.nbr), dp(df.cnt), wl.V, C.byref(wl.pm),
            dp(hyp_dev), n, dp(out), lfpm._stream())
    assert rc == 0, L.lfpm_last_error()
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_matches_eval_costs(lfpm, torch_cuda):
    L = lfpm.lib()
    cases = [(lfpm.workload_c2(raw_w=1024, raw_h=800), 30000, 5),
             (lfpm.workload_c1(raw_w=320, raw_h=256), 4000, 6)]
    for i, (wl, n, seed) in enumerate(cases):
        _say("case noalloc call %d: V=%d n=%d" % (i, wl.V, n))
        df = lfpm.DeviceFrame(wl)
        hyp = torch_cuda.from_numpy(lfpm.cone_hypotheses(wl, n, seed=seed)).cuda()
        ref = _eval(lfpm, torch_cuda, L.lfpm_eval_costs, df, hyp, n)
        got = _eval(lfpm, torch_cuda, L.lfpm_eval_costs_noalloc, df, hyp, n)
        assert np.array_equal(ref.view(np.uint32), got.view(np.uint32))
        assert (ref < 2).mean() > 0.2


def test_noalloc_rejects_beyond_limit(lfpm, torch_cuda):
    L = lfpm.lib()
    wl = lfpm.workload_c1(raw_w=256, raw_h=200)
    df = lfpm.DeviceFrame(wl)
    hyp = torch_cuda.from_numpy(lfpm.cone_hypotheses(wl, 10, seed=1)).cuda()
    out = torch_cuda.empty((10, wl.pm.max_neighbors), device="cuda")
    dp = lfpm._dp
    rc = L.lfpm_eval_costs_noalloc(dp(df.tiles), dp(df.centers), dp(df.nbr), dp(df.cnt),
                                   262144, C.byref(wl.pm), dp(hyp), 10, dp(out), lfpm._stream())
    assert rc == -1
    assert b"scratch" in L.lfpm_last_error()
