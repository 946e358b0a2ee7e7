"""Pins of the oracle's blur term (ComputeMultiViewCostVector_Frame, MIStereo/AdaptMIPMFrame.cu:
211-218, with the blur sums of ComputeBilateralNCC_Frame :140-144).

* Closed form: identical intensity tiles at zero disparity (NCC 0, all 49 window samples inside)
  with constant blur maps differing by D codes give lambda * (1 - exp(-0.007368 (49 D/255)^2)).
* lambda = 0 ignores the blur maps; the term never lifts a cost above 2.
* An independent implementation (scipy bilinear sampler for intensities and blur, numpy sums)
  on 300 random planes with smooth random blur maps, lambda 2.0 (plain) and 0.8 (ACMM value).
"""
import math

import numpy as np
import pytest

from test_oracle_pins import _hyp, _independent_ncc, _pm, _two_mi_frame

K_BLUR = float(np.float32(0.007368))


def _pm_blur(orc, W, lam, **kw):
    pm = _pm(orc, W, **kw)
    pm.blur_lambda = lam
    return pm


@pytest.mark.parametrize("lam", [2.0, 0.8])
@pytest.mark.parametrize("D", [1, 10, 40])
def test_blur_closed_form(orc, lfpm, lam, D):
    rng = np.random.default_rng(3)
    W = 16
    ref = rng.integers(20, 200, (W, W))
    tiles, centers, nbr, cnt = _two_mi_frame(orc, ref, ref)
    blur = np.stack([np.full((W, W), 100 + D), np.full((W, W), 100)]).astype(np.uint8)
    h0 = _hyp(lfpm, 8 * W + 8, [0, 0, 1, 0.0])
    c = orc.eval_costs(tiles, centers, nbr, cnt, _pm_blur(orc, W, lam), h0, blur_u8=blur)[0, 0]
    db = 49 * D / 255.0
    want = float(np.float32(lam)) * (1 - math.exp(-K_BLUR * db * db))
    assert abs(c - min(2.0, want)) < 1e-12


def test_blur_off_and_cap(orc, lfpm):
    rng = np.random.default_rng(4)
    W = 16
    ref = rng.integers(20, 200, (W, W))
    nei = np.clip(ref + rng.integers(-40, 40, (W, W)), 0, 255)
    tiles, centers, nbr, cnt = _two_mi_frame(orc, ref, nei)
    blur = rng.integers(0, 256, (2, W, W)).astype(np.uint8)
    h = _hyp(lfpm, 5 * W + 9, [0.1, -0.05, math.sqrt(1 - 0.0125), 1.5])
    base = orc.eval_costs(tiles, centers, nbr, cnt, _pm(orc, W), h)[0, 0]
    off = orc.eval_costs(tiles, centers, nbr, cnt, _pm_blur(orc, W, 0.0), h, blur_u8=blur)
    assert off[0, 0] == base
    big = orc.eval_costs(tiles, centers, nbr, cnt, _pm_blur(orc, W, 50.0), h, blur_u8=blur)
    assert base < 2.0 and big[0, 0] <= 2.0 and big[0, 0] >= base


def _independent_blur_cost(ref, nei, rb, nb, px, py, plane, step, r, ss, sc, mv, lam):
    from scipy.ndimage import map_coordinates
    c = _independent_ncc(ref, nei, px, py, plane, step, r, ss, sc, mv)
    W = ref.shape[0]
    nx, ny, nz, w = plane
    A, B = -nx / nz, -ny / nz
    ax, ay = px + step[0] * w, py + step[1] * w
    if not (0 <= ax < W and 0 <= ay < W):
        return 2.0
    ii, jj = np.meshgrid(np.arange(-r, r + 1), np.arange(-r, r + 1), indexing="ij")
    rx, ry = px + ii, py + jj
    inside = (rx >= 0) & (rx < W) & (ry >= 0) & (ry < W)
    d = w + A * ii + B * jj
    qx, qy = rx + step[0] * d, ry + step[1] * d
    ok = inside & (qx >= 0) & (qx < W) & (qy >= 0) & (qy < W)
    b_ref = (rb[ry[ok], rx[ok]] / 255.0).sum()
    b_src = map_coordinates(nb / 255.0, [qy[ok], qx[ok]], order=1, mode="nearest").sum()
    db = b_ref - b_src
    return float(min(2.0, max(0.0, c + lam * (1 - math.exp(-K_BLUR * db * db)))))


@pytest.mark.parametrize("lam", [2.0, 0.8])
def test_blur_matches_independent_implementation(orc, lfpm, lam):
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(5)
    W, r = 16, 4
    ref = rng.integers(0, 256, (W, W))
    nei = np.clip(ref + rng.integers(-30, 30, (W, W)), 0, 255)
    rb = np.clip(gaussian_filter(rng.random((W, W)) * 255, 2) * 2 - 120, 0, 255).round()
    nb = np.clip(gaussian_filter(rng.random((W, W)) * 255, 2) * 2 - 120, 0, 255).round()
    blur = np.stack([rb, nb]).astype(np.uint8)
    pm = _pm_blur(orc, W, lam, r=r, sc=0.15, ss=2.5)
    step = (1.0, 0.0)
    n = 300
    planes, pix = [], []
    for _ in range(n):
        th = math.radians(50) * rng.random()
        ph = 2 * math.pi * rng.random()
        planes.append([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th),
                       6 * rng.random()])
        pix.append(int(rng.integers(0, W * W)))
    hyp = lfpm.hypotheses_from_arrays(np.zeros(n, np.int32), np.array(pix),
                                      np.array(planes, np.float32))
    got = orc.eval_costs(*_two_mi_frame(orc, ref, nei), pm, hyp, blur_u8=blur)[:, 0]
    pl32 = np.array(planes, np.float32).astype(np.float64)
    sc, mv = float(np.float32(0.15)), float(np.float32(1e-5 / 255.0 ** 2))
    want = [_independent_blur_cost(ref, nei, rb, nb, p % W, p // W, pl32[i], step, r, 2.5, sc,
                                   mv, float(np.float32(lam))) for i, p in enumerate(pix)]
    assert np.max(np.abs(got - np.array(want))) < 1e-9
    no_blur = orc.eval_costs(*_two_mi_frame(orc, ref, nei), _pm(orc, W, r=r, sc=0.15, ss=2.5),
                             hyp)[:, 0]
    assert np.sum(got > no_blur + 1e-6) > n // 4  # the term is active on many planes


def test_laplacian_blur_tiles_restatement(orc):
    """The oracle's Laplacian blur tiles equal scipy's Laplacian with mirrored borders (OpenCV's
    BORDER_REFLECT_101) and hit the closed forms of a linear ramp: 0 inside, +-2 at the edges
    (the reflected neighbour repeats the inner one)."""
    from scipy.ndimage import laplace
    rng = np.random.default_rng(8)
    raw = rng.integers(0, 256, (40, 56)).astype(np.uint8)
    origins = np.array([[0, 0], [56 - 16, 40 - 16], [20, 11]], np.int32)
    got = orc.laplacian_blur_tiles(raw, origins, 16)
    full = laplace(raw.astype(np.float64), mode="mirror")
    want = np.stack([full[y:y + 16, x:x + 16] for x, y in origins]).astype(np.float32) / \
        np.float32(255.0)
    assert np.array_equal(got, want)
    ramp = np.tile(np.arange(56, dtype=np.uint8), (40, 1))
    t = orc.laplacian_blur_tiles(ramp, np.array([[0, 0], [40, 0]], np.int32), 16) * 255.0
    assert np.all(t[0][:, 1:] == 0) and np.all(t[0][:, 0] == 2)
    assert np.all(t[1][:, :15] == 0) and np.all(t[1][:, 15] == -2)
