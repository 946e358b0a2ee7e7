"""Colour raw-frame ingest: cv::cvtColor(..., COLOR_BGR2GRAY) of the MI crops cut from the colour
frame (DepthSolver::Slice_RawMLAImage, DepthSolver.cpp:1985-1990) as one device pass over the
frame (lfpm_bgr_to_gray).

The oracle's rule (OpenCV's 8-bit fixed-point form of 0.299 R + 0.587 G + 0.114 B) is pinned
by what the mathematics and OpenCV's documented behaviour fix: gray pixels keep their value,
the primaries give OpenCV's well-known 29 / 150 / 76 (blue / green / red at 255), and over all
2^24 colours the result stays within 0.52 of the real-valued formula (rounded, not truncated).
The GPU path is compared with the oracle bit for bit: the vector path (aligned rows of whole
16-pixel chunks), the byte path (odd pitches, row tails) and the chain into the MI slicer."""
import ctypes as C

import numpy as np
import pytest


def test_gray_pixels_keep_their_value(orc):
    v = np.arange(256, dtype=np.uint8)
    bgr = np.stack([v, v, v], axis=-1)[None]
    assert np.array_equal(orc.bgr_to_gray(bgr)[0], v)


def test_primaries_give_opencv_values(orc):
    px = np.array([[[255, 0, 0], [0, 255, 0], [0, 0, 255], [255, 255, 255], [0, 0, 0]]],
                  np.uint8)
    assert orc.bgr_to_gray(px)[0].tolist() == [29, 150, 76, 255, 0]


def test_within_half_of_real_formula_on_all_colours(orc):
    g, r = np.meshgrid(np.arange(256), np.arange(256), indexing="ij")
    worst = 0.0
    for b in range(256):
        bgr = np.stack([np.full_like(g, b), g, r], axis=-1).astype(np.uint8)
        y = orc.bgr_to_gray(bgr).astype(np.float64)
        exact = 0.114 * b + 0.587 * g + 0.299 * r
        worst = max(worst, float(np.abs(y - exact).max()))
    # coefficient rounding moves the value by at most 255 * 5e-5; truncation would reach 1
    assert worst <= 0.52, worst


def test_bad_arguments_are_rejected_without_a_device(lfpm):
    L = lfpm.lib()
    buf = C.c_void_p(1)
    assert L.lfpm_bgr_to_gray(None, 4, 4, C.c_size_t(12), buf, C.c_size_t(4), None) != 0
    assert L.lfpm_bgr_to_gray(buf, 4, 4, C.c_size_t(11), buf, C.c_size_t(4), None) != 0
    assert L.lfpm_bgr_to_gray(buf, 4, 4, C.c_size_t(12), buf, C.c_size_t(3), None) != 0
    assert L.lfpm_bgr_to_gray(buf, 0, 4, C.c_size_t(12), buf, C.c_size_t(4), None) != 0


def _say(*a):
    print(*a, flush=True)


# (width, height, extra bytes per colour row, extra bytes per gray row)
FRAMES = [(4080, 3068, 0, 0),     # whole 16-pixel chunks, 16-byte aligned rows: vector path
          (333, 77, 7, 3),        # odd pitches: byte path everywhere
          (17, 5, 13, 15),        # 48 + 13 = 61-byte rows: alignment varies by row
          (40, 9, 0, 0),          # 120-byte rows: aligned on every second row, 8-pixel tails
          (1, 1, 0, 0)]


@pytest.mark.gpu
@pytest.mark.parametrize("w,h,xb,xg", FRAMES)
def test_bgr_to_gray_matches_oracle(lfpm, orc, w, h, xb, xg):
    import torch
    _say("bgr_to_gray", w, h, xb, xg)
    rng = np.random.default_rng(w * 7 + h)
    bp, gp = 3 * w + xb, w + xg
    store = rng.integers(0, 256, size=h * bp, dtype=np.uint8)
    dev = torch.from_numpy(store).cuda()
    bgr = dev.as_strided((h, w, 3), (bp, 3, 1))
    gbuf = torch.full((h * gp,), 7, dtype=torch.uint8, device="cuda")
    gray = gbuf.as_strided((h, w), (gp, 1))
    lfpm.bgr_to_gray(bgr, out=gray)
    torch.cuda.synchronize()
    host = np.lib.stride_tricks.as_strided(store, (h, w, 3), (bp, 3, 1))
    want = orc.bgr_to_gray(host)
    got = gbuf.cpu().numpy()
    assert np.array_equal(np.lib.stride_tricks.as_strided(got, (h, w), (gp, 1)), want)
    # the pitch padding of the gray rows is left alone
    pad = np.lib.stride_tricks.as_strided(got, (h, gp), (gp, 1))[:, w:]
    assert (pad == 7).all()


@pytest.mark.gpu
def test_colour_frame_slices_like_converted_crops(lfpm, orc):
    """Convert-then-slice on the device equals converting each colour crop (the reference's
    order) in the oracle."""
    import torch
    wl = lfpm.workload_c1()
    h, w = wl.raw().shape
    _say("colour slice", w, h, wl.V)
    rng = np.random.default_rng(5)
    bgr = rng.integers(0, 256, size=(h, w, 3), dtype=np.uint8)
    gray = lfpm.bgr_to_gray(torch.from_numpy(bgr).cuda())
    org = torch.from_numpy(np.ascontiguousarray(wl.L["origins"], np.int32)).cuda()
    V, W = wl.V, wl.crop
    out = torch.empty((V, W, W), dtype=torch.float32, device="cuda")
    lfpm._check(lfpm.lib().lfpm_slice_gray(lfpm._dp(gray), w, h, C.c_size_t(w), lfpm._dp(org),
                                           V, W, lfpm._dp(out), lfpm._stream()))
    torch.cuda.synchronize()
    crops = np.stack([bgr[y:y + W, x:x + W] for x, y in wl.L["origins"]])
    want = orc.bgr_to_gray(crops).astype(np.float32)
    assert np.array_equal(out.cpu().numpy(), want)
