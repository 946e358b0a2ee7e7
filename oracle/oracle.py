"""ctypes binding of the CPU oracle (oracle/liblfpm_oracle.so) — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module; the
product path never does.  Structures mirror the product ABI field for field.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblfpm_oracle.so")


class PmParams(C.Structure):
    _fields_ = [("crop", C.c_int), ("base", C.c_float), ("radius", C.c_int),
                ("radius_increment", C.c_int), ("sigma_spatial", C.c_float),
                ("sigma_color", C.c_float), ("min_var", C.c_float), ("top_k", C.c_int),
                ("depth_min", C.c_float), ("depth_max", C.c_float), ("iterations", C.c_int),
                ("max_neighbors", C.c_int), ("seed", C.c_uint32), ("frame_uid", C.c_uint32),
                ("storage", C.c_int), ("cos_max_slant", C.c_float),
                ("blur_lambda", C.c_float)]


class FilterParams(C.Structure):
    _fields_ = [("max_triplet", C.c_int), ("min_valid_disp", C.c_float),
                ("use_selected_views_only", C.c_int),
                ("clear_selected_views_when_invalid", C.c_int),
                ("enable_cost_filter", C.c_int), ("max_cost", C.c_float),
                ("enable_cycle_geo_filter", C.c_int), ("max_geo_err_px", C.c_float),
                ("enable_cycle_photo_filter", C.c_int), ("max_photo_err_u", C.c_float),
                ("min_good_neighbors", C.c_int), ("enable_spike_filter", C.c_int),
                ("spike_abs_diff", C.c_float), ("spike_min_neighbors", C.c_int)]


class LayoutParams(C.Structure):
    _fields_ = [("pitch", C.c_float), ("crop", C.c_int), ("raw_w", C.c_int),
                ("raw_h", C.c_int), ("max_neighbors", C.c_int), ("lens_types", C.c_int),
                ("neighbors_per_type", C.c_int * 3)]


def build():
    """Compile the oracle with its Makefile (g++ only)."""
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def nthreads():
    return max(1, min(16, os.cpu_count() or 1))


def set_progress(on):
    """Stage lines on stderr from solve (init, then each iteration's two phases)."""
    lib().orc_set_progress(1 if on else 0)


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox(c, k, o)
    return list(o)


def layout(lp):
    L = lib()
    L.orc_layout.restype = C.c_int
    V = L.orc_layout(C.byref(lp), 0, None, None, None, None, None, None)
    K = lp.max_neighbors
    centers = np.zeros((V, 2), np.float32)
    origins = np.zeros((V, 2), np.int32)
    keys = np.zeros((V, 2), np.int32)
    ltype = np.zeros(V, np.int32)
    nbr = np.zeros((V, K), np.int32)
    cnt = np.zeros(V, np.int32)
    L.orc_layout(C.byref(lp), V, _p(centers), _p(origins), _p(keys), _p(ltype), _p(nbr),
                 _p(cnt))
    return dict(V=V, centers=centers, origins=origins, tilekeys=keys, lens_type=ltype,
                nbr=nbr, cnt=cnt)


def fixed_offsets(X, Y, u, v):
    """FixedPosition mode-1 key offsets of key (X, Y) in a u x v lattice, collection order."""
    out = np.zeros(24, np.int32)
    n = lib().orc_fixed_offsets(int(X), int(Y), int(u), int(v), _p(out))
    return [tuple(int(t) for t in out[2 * e:2 * e + 2]) for e in range(n)]


def fixed_neighbors(lp):
    """FixedPosition neighbour table of a layout: (nbr (V, K), cnt (V,), garbage (V,))."""
    L = lib()
    V = L.orc_layout(C.byref(lp), 0, None, None, None, None, None, None)
    K = lp.max_neighbors
    nbr = np.zeros((V, K), np.int32)
    cnt = np.zeros(V, np.int32)
    garbage = np.zeros(V, np.int32)
    L.orc_fixed_neighbors(C.byref(lp), _p(nbr), _p(cnt), _p(garbage))
    return nbr, cnt, garbage


def slice_u8(raw, origins, crop):
    raw = np.ascontiguousarray(raw, np.uint8)
    V = origins.shape[0]
    tiles = np.zeros((V, crop, crop), np.uint8)
    lib().orc_slice_u8(_p(raw), raw.shape[1], raw.shape[0], _p(np.ascontiguousarray(origins)),
                       V, crop, _p(tiles))
    return tiles


def _blur_values(tiles_u8, pm, blur_u8=None, blur=None):
    """Blur values as the NCC reads them (score / 255, FP64) when the blur term is on: from
    8-bit blur maps (blur_u8) or from blur scores already divided by 255 (blur)."""
    if pm.blur_lambda <= 0:
        return None
    if blur is None:
        assert blur_u8 is not None and blur_u8.shape == tiles_u8.shape
        blur = blur_u8.astype(np.float64) / 255.0
    assert blur.shape == tiles_u8.shape
    return np.ascontiguousarray(blur, np.float64)


def bgr_to_gray(bgr):
    """cv::cvtColor(..., cv::COLOR_BGR2GRAY) on an 8-bit (H, W, 3) B, G, R image, the conversion
    of the colour MI crops in DepthSolver::Slice_RawMLAImage (DepthSolver.cpp:1985-1990).
    OpenCV's published 8-bit form of Y = 0.299 R + 0.587 G + 0.114 B: coefficients rounded to
    14-bit fixed point (4899, 9617, 1868, summing to 2^14), result rounded half up."""
    c = np.asarray(bgr, np.int64)
    return ((c[..., 0] * 1868 + c[..., 1] * 9617 + c[..., 2] * 4899 + 8192) >> 14).astype(np.uint8)


def vd_point_cloud(vd, color, origin, pixel_size, f, B, bl0):
    """Records of the VD point cloud (DepthSolver::Virtual_depth_map_TileKeySequence,
    DepthSolver.cpp:8019-8053) in the binary layout of StoreColorPlyFileBinaryPointCloud
    (MVStereo/ACMP.cpp:451-500): bytes (n, 27) = x, y, z, nx, ny, nz float32 LE + red, green,
    blue, for the pixels with vd > 0 in row-major order.  The image-plane point is FP64, the
    virtual-depth line vdm * B + bl0 FP32 with two roundings, the thin-lens map FP64; non-finite
    points are written as 0 as the writer does; normals 0; colour = the gray value (color
    (h, w)) or red, green, blue of a B, G, R colour image (color (h, w, 3))."""
    vd = np.asarray(vd, np.float32)
    ys, xs = np.nonzero(vd > 0)
    vdm = vd[ys, xs]
    f32 = np.float32
    tz = ((vdm * f32(B)).astype(f32) + f32(bl0)).astype(f32).astype(np.float64)
    tx = (xs.astype(np.float64) - float(np.float32(origin[0]))) * float(np.float32(pixel_size))
    ty = (ys.astype(np.float64) - float(np.float32(origin[1]))) * float(np.float32(pixel_size))
    fd = float(np.float32(f))
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        z = (fd / (tz - fd) * tz).astype(f32)
        x = (-fd / (tz - fd) * tx).astype(f32)
        y = (-fd / (tz - fd) * ty).astype(f32)
    fmax = np.finfo(np.float32).max
    bad = ~((x < fmax) & (x > -fmax)) | ~((y < fmax) & (y > -fmax)) | ~((z < fmax) & (z >= -fmax))
    x[bad] = 0
    y[bad] = 0
    z[bad] = 0
    n = len(vdm)
    rec = np.zeros((n, 27), np.uint8)
    xyz = np.stack([x, y, z, np.zeros(n, f32), np.zeros(n, f32), np.zeros(n, f32)], axis=1)
    rec[:, :24] = np.ascontiguousarray(xyz.astype("<f4")).view(np.uint8).reshape(n, 24)
    if color is None:
        rec[:, 24:27] = 0
    elif np.asarray(color).ndim == 3:
        # colour image B, G, R: the writer emits red = channel 2, green = 1, blue = 0
        c = np.asarray(color, np.uint8)[ys, xs]
        rec[:, 24] = c[:, 2]
        rec[:, 25] = c[:, 1]
        rec[:, 26] = c[:, 0]
    else:
        g = np.asarray(color, np.uint8)[ys, xs]
        rec[:, 24] = g
        rec[:, 25] = g
        rec[:, 26] = g
    return rec


def laplacian_blur_tiles(raw, origins, crop):
    """Blur tiles of ImageQualityEstimate::QuantizeBlurByLaplacian (Util/ImageQualityEstimate.cpp:
    211-230): cv::Laplacian with ksize 1 (the 4-neighbour stencil) of the gray frame with
    BORDER_REFLECT_101, cropped per MI and divided by 255 in FP32 as the kernel reads it."""
    g = np.asarray(raw, np.int64)
    p = np.pad(g, 1, mode="reflect")  # numpy 'reflect' = edge pixel not repeated
    lap = p[1:-1, :-2] + p[1:-1, 2:] + p[:-2, 1:-1] + p[2:, 1:-1] - 4 * g
    t = np.stack([lap[y:y + crop, x:x + crop] for x, y in np.asarray(origins)])
    return t.astype(np.float32) / np.float32(255.0)


def solve(tiles_u8, centers, nbr, cnt, pm, dtype=np.float64, threads=None, blur_u8=None,
          blur=None):
    V, W, _ = tiles_u8.shape
    tiles_u8 = np.ascontiguousarray(tiles_u8, np.uint8)
    bv = _blur_values(tiles_u8, pm, blur_u8, blur)
    n = V * W * W
    plane = np.zeros((n, 4), dtype)
    cost = np.zeros(n, dtype)
    disp = np.zeros((n, 4), dtype)
    sel = np.zeros(n, np.uint32)
    f = lib().orc_solve_f64 if dtype == np.float64 else lib().orc_solve_f32
    f(_p(tiles_u8), _p(bv) if bv is not None else None,
      _p(np.ascontiguousarray(centers, np.float32)),
      _p(np.ascontiguousarray(nbr, np.int32)), _p(np.ascontiguousarray(cnt, np.int32)), V,
      C.byref(pm), _p(plane), _p(cost), _p(disp), _p(sel), threads or nthreads())
    return dict(plane=plane, cost=cost, disp=disp, sel=sel)


def eval_costs(tiles_u8, centers, nbr, cnt, pm, hyp, dtype=np.float64, threads=None,
               blur_u8=None, blur=None):
    """hyp: structured (n, 6) int32 view: mi, pixel, then 4 float32 bit patterns."""
    V = tiles_u8.shape[0]
    tiles_u8 = np.ascontiguousarray(tiles_u8, np.uint8)
    bv = _blur_values(tiles_u8, pm, blur_u8, blur)
    n = hyp.shape[0]
    out = np.zeros((n, pm.max_neighbors), dtype)
    f = lib().orc_eval_costs_f64 if dtype == np.float64 else lib().orc_eval_costs_f32
    f(_p(tiles_u8), _p(bv) if bv is not None else None,
      _p(np.ascontiguousarray(centers, np.float32)),
      _p(np.ascontiguousarray(nbr, np.int32)), _p(np.ascontiguousarray(cnt, np.int32)), V,
      C.byref(pm), _p(np.ascontiguousarray(hyp, np.int32)), n, _p(out),
      threads or nthreads())
    return out


def eval_margins(tiles_u8, centers, nbr, cnt, pm, hyp, threads=None, blur_u8=None, blur=None):
    """FP64 costs (n, K) and, per (hypothesis, view), where FP32 may legitimately disagree:
    geo (n, K), the smallest move in px of the anchor or a window sample's source point that
    changes its inside-the-crop test; gate (n, K), the distance of the window variances from
    the variance gate in units of their FP32 error estimate (inf where the evaluation ends
    before the gate); ferr (n, K), the first-order FP32 error estimate of the cost (0 where it
    is 2 by an early return)."""
    V = tiles_u8.shape[0]
    tiles_u8 = np.ascontiguousarray(tiles_u8, np.uint8)
    bv = _blur_values(tiles_u8, pm, blur_u8, blur)
    n = hyp.shape[0]
    K = pm.max_neighbors
    out = np.zeros((n, K), np.float64)
    mg = np.zeros((n, K, 3), np.float64)
    lib().orc_eval_margins_f64(
        _p(tiles_u8), _p(bv) if bv is not None else None,
        _p(np.ascontiguousarray(centers, np.float32)), _p(np.ascontiguousarray(nbr, np.int32)),
        _p(np.ascontiguousarray(cnt, np.int32)), V, C.byref(pm),
        _p(np.ascontiguousarray(hyp, np.int32)), n, _p(out), _p(mg), threads or nthreads())
    return out, mg[..., 0], mg[..., 1], mg[..., 2]


def filter_f32(gray, centers, nbr, cnt, crop, K, base, fp, plane, cost, disp, sel, threads=None):
    V = centers.shape[0]
    n = V * crop * crop
    po = np.zeros((n, 4), np.float32)
    co = np.zeros(n, np.float32)
    do = np.zeros((n, 4), np.float32)
    so = np.zeros(n, np.uint32)
    reason = np.zeros(n, np.uint8)
    lib().orc_filter_f32(
        _p(np.ascontiguousarray(gray, np.float32)), _p(np.ascontiguousarray(centers, np.float32)),
        _p(np.ascontiguousarray(nbr, np.int32)), _p(np.ascontiguousarray(cnt, np.int32)), V,
        crop, K, C.c_float(base), C.byref(fp), _p(np.ascontiguousarray(plane, np.float32)),
        _p(np.ascontiguousarray(cost, np.float32)), _p(np.ascontiguousarray(disp, np.float32)),
        _p(np.ascontiguousarray(sel, np.uint32)), _p(po), _p(co), _p(do), _p(so), _p(reason),
        threads or nthreads())
    return dict(plane=po, cost=co, disp=do, sel=so, reason=reason)


def filter_f32_inplace(gray, centers, nbr, cnt, crop, K, base, fp, plane, cost, disp, sel):
    """filter_f32 in the reference's order: MI after MI, each reading the disparities and
    costs the MIs before it left (its rejected pixels invalidated in place)."""
    V = centers.shape[0]
    n = V * crop * crop
    po = np.zeros((n, 4), np.float32)
    co = np.zeros(n, np.float32)
    do = np.zeros((n, 4), np.float32)
    so = np.zeros(n, np.uint32)
    reason = np.zeros(n, np.uint8)
    lib().orc_filter_f32_inplace(
        _p(np.ascontiguousarray(gray, np.float32)), _p(np.ascontiguousarray(centers, np.float32)),
        _p(np.ascontiguousarray(nbr, np.int32)), _p(np.ascontiguousarray(cnt, np.int32)), V,
        crop, K, C.c_float(base), C.byref(fp), _p(np.ascontiguousarray(plane, np.float32)),
        _p(np.ascontiguousarray(cost, np.float32)), _p(np.ascontiguousarray(disp, np.float32)),
        _p(np.ascontiguousarray(sel, np.uint32)), _p(po), _p(co), _p(do), _p(so), _p(reason))
    return dict(plane=po, cost=co, disp=do, sel=so, reason=reason)


def refocus(tiles_u8, centers, pitch, crop, out_w, out_h, S, vmin, vmax, threads=None):
    V = centers.shape[0]
    out = np.zeros((S, out_h, out_w), np.float64)
    lib().orc_refocus_f64(_p(np.ascontiguousarray(tiles_u8)),
                          _p(np.ascontiguousarray(centers, np.float32)), V, C.c_float(pitch),
                          crop, out_w, out_h, S, C.c_float(vmin), C.c_float(vmax), _p(out),
                          threads or nthreads())
    return out


def fit_v2r(aL, v):
    aL = np.ascontiguousarray(aL, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    c = np.zeros(3, np.float64)
    L = lib()
    L.orc_fit_v2r.restype = C.c_int
    rc = L.orc_fit_v2r(_p(aL), _p(v), aL.size, _p(c))
    if rc != 0:
        raise RuntimeError("oracle fit failed: %d" % rc)
    return c


def v2r_convert(v, c):
    v = np.ascontiguousarray(v, np.float32)
    z = np.zeros(v.size, np.float64)
    lib().orc_v2r_convert(_p(v), v.size, _p(np.ascontiguousarray(c, np.float64)), _p(z))
    return z


def topk(cv, top_k):
    cv = np.ascontiguousarray(cv, np.float64)
    mean = C.c_double()
    best = C.c_int()
    L = lib()
    L.orc_topk_f64.restype = C.c_uint
    m = L.orc_topk_f64(_p(cv), cv.size, top_k, C.byref(mean), C.byref(best))
    return m, mean.value, best.value


def weighted(cv, vw):
    cv = np.ascontiguousarray(cv, np.float64)
    vw = np.ascontiguousarray(vw, np.float64)
    am = C.c_int()
    L = lib()
    L.orc_weighted_f64.restype = C.c_double
    c = L.orc_weighted_f64(_p(cv), _p(vw), cv.size, C.byref(am))
    return c, am.value


def reencode_plane(hn, q, du, dv, p):
    hn = np.ascontiguousarray(hn, np.float64)
    out = np.zeros(4)
    lib().orc_reencode_plane_f64(_p(hn), int(q[0]), int(q[1]), C.c_double(du), C.c_double(dv),
                                 int(p[0]), int(p[1]), _p(out))
    return out


def random_normals(seed, uid, step, n):
    out = np.zeros((n, 3))
    lib().orc_random_normals_f64(C.c_uint32(seed), C.c_uint32(uid), C.c_uint32(step), n, _p(out))
    return out


def disparity_mosaic(plane, cost, centers, crop, max_cost, raw_w, raw_h):
    """plane (V*crop^2, 4) float32, cost (V*crop^2,) float32 -> (mosaic (h, w) f32, range (2,))."""
    plane = np.ascontiguousarray(plane, np.float32)
    cost = np.ascontiguousarray(cost, np.float32)
    centers = np.ascontiguousarray(centers, np.float32)
    V = centers.shape[0]
    out = np.zeros((raw_h, raw_w), np.float32)
    rng = np.zeros(2, np.float32)
    lib().orc_disparity_mosaic(_p(plane), _p(cost), _p(centers), V, crop, C.c_float(max_cost),
                               raw_w, raw_h, _p(out), _p(rng))
    return out, rng


def vd_map(plane, centers, gray, crop, mla_base, raw_w, raw_h, color=True):
    """-> (vd (h, w) f32, colour (h, w) u8 or None, range (2,) f32)."""
    plane = np.ascontiguousarray(plane, np.float32)
    centers = np.ascontiguousarray(centers, np.float32)
    V = centers.shape[0]
    g = None if gray is None else np.ascontiguousarray(gray, np.float32)
    vd = np.zeros((raw_h, raw_w), np.float32)
    col = np.zeros((raw_h, raw_w), np.uint8) if color else None
    rng = np.zeros(2, np.float32)
    lib().orc_vd_map(_p(plane), _p(centers), _p(g) if g is not None else None, V, crop,
                     C.c_float(mla_base), raw_w, raw_h, _p(vd),
                     _p(col) if col is not None else None, _p(rng))
    return vd, col, rng


def vd_map_bgr(plane, centers, bgr_tiles, crop, mla_base, raw_w, raw_h):
    """VD map with the colour crops bgr_tiles (V, crop, crop, 3) B, G, R: -> (vd (h, w) f32,
    colour (h, w, 3) u8 B, G, R, range (2,) f32)."""
    plane = np.ascontiguousarray(plane, np.float32)
    centers = np.ascontiguousarray(centers, np.float32)
    V = centers.shape[0]
    b = np.ascontiguousarray(bgr_tiles, np.uint8)
    assert b.shape == (V, crop, crop, 3)
    vd = np.zeros((raw_h, raw_w), np.float32)
    col = np.zeros((raw_h, raw_w, 3), np.uint8)
    rng = np.zeros(2, np.float32)
    lib().orc_vd_map_bgr(_p(plane), _p(centers), _p(b), V, crop, C.c_float(mla_base), raw_w,
                         raw_h, _p(vd), _p(col), _p(rng))
    return vd, col, rng


def aif_composite(plane, centers, bgr_tiles, crop, baseline, B, bl0, f_main, raw_w, raw_h,
                  rich=None, blur=None, garbage=None, L=None):
    """All-in-focus composite (AIFImageCompositeForMIA): dict vd, inv_vd, real_mm (h, w) f32,
    aif, filled, denoised (h, w, 3) u8, stats (8,) int32."""
    plane = np.ascontiguousarray(plane, np.float32)
    centers = np.ascontiguousarray(centers, np.float32)
    V = centers.shape[0]
    b = np.ascontiguousarray(bgr_tiles, np.uint8)
    assert b.shape == (V, crop, crop, 3)
    r = dict(vd=np.zeros((raw_h, raw_w), np.float32), inv_vd=np.zeros((raw_h, raw_w), np.float32),
             real_mm=np.zeros((raw_h, raw_w), np.float32),
             aif=np.zeros((raw_h, raw_w, 3), np.uint8), filled=np.zeros((raw_h, raw_w, 3), np.uint8),
             denoised=np.zeros((raw_h, raw_w, 3), np.uint8), stats=np.zeros(8, np.int32))
    ri = np.ascontiguousarray(rich, np.uint8) if rich is not None else None
    bl = np.ascontiguousarray(blur, np.uint8) if blur is not None else None
    ga = np.ascontiguousarray(garbage, np.uint8) if garbage is not None else None
    (L or lib()).orc_aif_composite(_p(plane), _p(centers), _p(b), _p(ri) if ri is not None else None,
                            _p(bl) if bl is not None else None, _p(ga) if ga is not None else None,
                            V, crop, C.c_float(baseline), C.c_float(B), C.c_float(bl0),
                            C.c_float(f_main), raw_w, raw_h, _p(r["vd"]), _p(r["inv_vd"]),
                            _p(r["real_mm"]), _p(r["aif"]), _p(r["filled"]), _p(r["denoised"]),
                            _p(r["stats"]))
    return r


def fuse_backproject(plane, centers, bgr_tiles, nbr, cnt, crop, baseline, B, bl0, raw_w, raw_h,
                     garbage=None, L=None):
    """Back-projection consistency fusion (FuseVirtualDepth_BackProject): dict vd (h, w) f32,
    aif (h, w, 3) u8, conf (h, w) u8, points (n, 9) f32, nfused."""
    plane = np.ascontiguousarray(plane, np.float32)
    centers = np.ascontiguousarray(centers, np.float32)
    V = centers.shape[0]
    b = np.ascontiguousarray(bgr_tiles, np.uint8)
    nbr = np.ascontiguousarray(nbr, np.int32)
    cnt = np.ascontiguousarray(cnt, np.int32)
    K = nbr.size // V
    ga = np.ascontiguousarray(garbage, np.uint8) if garbage is not None else None
    r = dict(vd=np.zeros((raw_h, raw_w), np.float32), aif=np.zeros((raw_h, raw_w, 3), np.uint8),
             conf=np.zeros((raw_h, raw_w), np.uint8))
    pts = np.zeros((V * crop * crop, 9), np.float32)
    npts = C.c_longlong()
    nf = C.c_longlong()
    (L or lib()).orc_fuse_backproject(_p(plane), _p(centers), _p(b), _p(nbr), _p(cnt), K,
                               _p(ga) if ga is not None else None, V, crop, C.c_float(baseline),
                               C.c_float(B), C.c_float(bl0), raw_w, raw_h, _p(r["vd"]),
                               _p(r["aif"]), _p(r["conf"]), _p(pts), C.byref(npts),
                               C.byref(nf))
    r["points"] = pts[:npts.value].copy()
    r["nfused"] = nf.value
    return r


def slice_bgr(bgr, origins, crop):
    """Colour crops (V, crop, crop, 3) of a (H, W, 3) B, G, R frame at the MI origins (the
    problem.m_Image_rgb crops of DepthSolver::Slice_RawMLAImage, DepthSolver.cpp:1951-1997)."""
    bgr = np.asarray(bgr, np.uint8)
    return np.stack([bgr[y:y + crop, x:x + crop] for x, y in np.asarray(origins)])


def solve_acmm(tiles_u8, centers, nbr, cnt, pm, threads=None, blur_u8=None, blur=None):
    """ACMM multi-scale solve (FP64) of the final level: dict plane/cost/disp/sel.  With
    pm.blur_lambda > 0, blur maps as in solve (a coarse level reads them area-resized)."""
    bv = _blur_values(tiles_u8, pm, blur_u8, blur)
    V, W, _ = tiles_u8.shape
    n = V * W * W
    plane = np.zeros((n, 4), np.float64)
    cost = np.zeros(n, np.float64)
    disp = np.zeros((n, 4), np.float64)
    sel = np.zeros(n, np.uint32)
    lib().orc_solve_acmm_f64(_p(np.ascontiguousarray(tiles_u8)),
                             _p(bv) if bv is not None else None,
                             _p(np.ascontiguousarray(centers, np.float32)),
                             _p(np.ascontiguousarray(nbr, np.int32)),
                             _p(np.ascontiguousarray(cnt, np.int32)), V, C.byref(pm), _p(plane),
                             _p(cost), _p(disp), _p(sel), threads or nthreads())
    return dict(plane=plane, cost=cost, disp=disp, sel=sel)


def area_resize(tile, W0):
    tile = np.ascontiguousarray(tile, np.float64)
    W = tile.shape[0]
    out = np.zeros((W0, W0), np.float64)
    lib().orc_area_resize_f64(_p(tile), W, W0, _p(out))
    return out


def acmm_levels(W):
    width = np.zeros(2, np.int32)
    scale = np.zeros(2, np.float32)
    radius = np.zeros(2, np.int32)
    iters = np.zeros(2, np.int32)
    topk = np.zeros(2, np.int32)
    L = lib().orc_acmm_levels(W, _p(width), _p(scale), _p(radius), _p(iters), _p(topk))
    return [dict(width=int(width[i]), scale=float(scale[i]), radius=int(radius[i]),
                 iterations=int(iters[i]), top_k=int(topk[i])) for i in range(L)]


# --- pieces of the checkerboard update (FP64), fed with explicit stream words ---------------
KMAX = 32
SLOTS = 64


def words_from_uniforms(slot_uniforms):
    """Raw stream words whose uniform ((x >> 8) + 0.5) * 2^-24 is the given value, rounded down
    to the 2^-24 grid: dict slot -> u in (0, 1).  Unset slots hold 0x80000000 (u = 0.5 + 2^-25)."""
    w = np.full(SLOTS, 0x80000000, np.uint32)
    for s, u in slot_uniforms.items():
        k = int(np.floor(float(u) * 2 ** 24 - 0.5))
        k = min(max(k, 0), 2 ** 24 - 1)
        w[s] = np.uint32(k << 8)
    return w


def uniform_of_word(x):
    return ((int(x) >> 8) + 0.5) * 2.0 ** -24


def propagation_candidates(W, x, y, cost, L=None):
    """(flag[8], pos[8], dup[8]) of pixel (x, y); cost = the MI's W*W previous costs."""
    L = L or lib()
    cost = np.ascontiguousarray(cost, np.float64).reshape(-1)
    assert cost.size == W * W
    flag = np.zeros(8, np.int32)
    pos = np.zeros(8, np.int32)
    dup = np.zeros(8, np.int32)
    L.orc_propagation_candidates_f64(W, x, y, _p(cost), _p(flag), _p(pos), _p(dup))
    return flag.astype(bool), pos, dup


def view_priors(W, x, y, sel, cnt, L=None):
    L = L or lib()
    sel = np.ascontiguousarray(sel, np.uint32).reshape(-1)
    assert sel.size == W * W
    prior = np.zeros(KMAX)
    L.orc_view_priors_f64(W, x, y, _p(sel), cnt, _p(prior))
    return prior[:cnt]


def view_select(ca, flag, prior, iter_, words, L=None):
    """ca (8, cnt) candidate costs, flag (8,), prior (cnt,): returns prob, cdf, vw, tsel."""
    L = L or lib()
    ca = np.asarray(ca, np.float64)
    cnt = ca.shape[1]
    c = np.full((8, KMAX), 2.0)
    c[:, :cnt] = ca
    fl = np.ascontiguousarray(np.asarray(flag, bool).astype(np.int32))
    pr = np.zeros(KMAX)
    pr[:cnt] = prior
    prob = np.zeros(KMAX)
    cdf = np.zeros(KMAX)
    vw = np.zeros(KMAX)
    tsel = C.c_uint32()
    L.orc_view_select_f64(_p(c), _p(fl), _p(pr), cnt, int(iter_),
                          _p(np.ascontiguousarray(words, np.uint32)), _p(prob), _p(cdf), _p(vw),
                          C.byref(tsel))
    return prob[:cnt], cdf[:cnt], vw[:cnt], tsel.value


def perturbed_normal(h, px, py, perturbation, words, L=None):
    L = L or lib()
    out = np.zeros(4)
    L.orc_perturbed_normal_f64(_p(np.ascontiguousarray(h, np.float64)), int(px), int(py),
                               C.c_double(perturbation),
                               _p(np.ascontiguousarray(words, np.uint32)), _p(out))
    return out


def refine_candidates(pm, x, y, best, bcost, words, L=None):
    L = L or lib()
    out = np.zeros((5, 4))
    L.orc_refine_candidates_f64(C.byref(pm), int(x), int(y),
                                _p(np.ascontiguousarray(best, np.float64)), C.c_double(bcost),
                                _p(np.ascontiguousarray(words, np.uint32)), _p(out))
    return out


def ring_neighbors(tilekeys, centers, rings, K):
    """Matching neighbour table of a lens table by brute force over all lens pairs: the lenses at
    hex cube distance 1..rings of each lens (keys (col, row), odd rows half a pitch right),
    ordered by the float centre distance sqrt(double dx^2 + double dy^2), then angle in [0, 2 pi);
    capped at K, -1 padded (SelectNeighbors::CollectMIANeighImagesForMatch with the score of
    ComputeNeighborScoresForMatch, MINeighbor/SelectNeighbors.cpp:113-193,770-830, and
    MLA_Problem::SortNeighScoreForMatch, MVStereo/LFDepthInfo.cpp:214-241)."""
    keys = np.asarray(tilekeys, np.int64)
    c = np.asarray(centers, np.float32)
    V = keys.shape[0]
    X, Y = keys[:, 0], keys[:, 1]
    q = X - (Y - (Y & 1)) // 2
    nbr = np.full((V, K), -1, np.int32)
    cnt = np.zeros(V, np.int32)
    for v in range(V):
        dq = q - q[v]
        dr = Y - Y[v]
        d = np.maximum(np.maximum(np.abs(dq), np.abs(dr)), np.abs(dq + dr))
        u = np.nonzero((d >= 1) & (d <= rings))[0]
        dx = (c[u, 0] - c[v, 0]).astype(np.float32).astype(np.float64)
        dy = (c[u, 1] - c[v, 1]).astype(np.float32).astype(np.float64)
        b = np.sqrt(dx * dx + dy * dy).astype(np.float32)
        a = np.arctan2(dy, dx)
        a = np.where(a < 0, a + 2 * np.pi, a)
        order = np.lexsort((a, b))
        sel = u[order][:K]
        nbr[v, :len(sel)] = sel
        cnt[v] = len(sel)
    return nbr, cnt
