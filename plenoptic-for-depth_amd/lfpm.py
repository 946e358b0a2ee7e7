"""Python mirror of the lfpm C ABI (include/lfpm.h) over ctypes.

Host entry points take numpy arrays; device entry points take torch CUDA tensors (PyTorch is
only the device-memory and stream plumbing here) and run on the current torch stream.  The
HIP library must be built (``make -C plenoptic-for-depth_amd``); there is no CPU fallback:
every device call goes through liblfpm.so and raises LfpmError on a non-zero status.
"""
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblfpm.so")


class LfpmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("lfpm error %d: %s" % (code, msg))
        self.code = code


class LayoutParams(C.Structure):
    _fields_ = [("pitch", C.c_float), ("crop", C.c_int), ("raw_w", C.c_int),
                ("raw_h", C.c_int), ("max_neighbors", C.c_int), ("lens_types", C.c_int),
                ("neighbors_per_type", C.c_int * 3)]


class LayoutInfo(C.Structure):
    _fields_ = [("num_mis", C.c_int), ("crop", C.c_int), ("lattice_cols", C.c_int),
                ("lattice_rows", C.c_int), ("row_step", C.c_float)]


class SynthParams(C.Structure):
    _fields_ = [("raw_w", C.c_int), ("raw_h", C.c_int), ("pitch", C.c_float),
                ("num_planes", C.c_int), ("vmin", C.c_float), ("vmax", C.c_float),
                ("scene_seed", C.c_uint32), ("noise_seed", C.c_uint32),
                ("noise_sigma", C.c_float)]


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


_lib = None


def lib():
    """Load liblfpm.so; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LfpmError(-4, "liblfpm.so not built (run make -C plenoptic-for-depth_amd)")
        L = C.CDLL(LIB_PATH)
        L.lfpm_last_error.restype = C.c_char_p
        L.lfpm_patchmatch_workspace_size.restype = C.c_size_t
        L.lfpm_products_workspace_size.restype = C.c_size_t
        L.lfpm_patchmatch_acmm_workspace_size.restype = C.c_size_t
        L.lfpm_batch_workspace_size.restype = C.c_size_t
        L.lfpm_aif_workspace_size.restype = C.c_size_t
        L.lfpm_fuse_workspace_size.restype = C.c_size_t
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise LfpmError(rc, lib().lfpm_last_error().decode())


def _np(a):
    return a.ctypes.data_as(C.c_void_p)


def _dp(t):
    """Device pointer of a torch tensor (must be contiguous)."""
    assert t.is_contiguous()
    return C.c_void_p(t.data_ptr())


def _stream(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


# --------------------------------------------------------------------------------------
# Host entry points
# --------------------------------------------------------------------------------------
def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().lfpm_philox4x32_10(c, k, o)
    return list(o)


def layout_params(pitch, raw_w, raw_h, max_neighbors, crop=0, lens_types=1, per_type=None):
    per = per_type or (max_neighbors, max_neighbors, max_neighbors)
    return LayoutParams(float(pitch), int(crop), int(raw_w), int(raw_h), int(max_neighbors),
                        int(lens_types), (C.c_int * 3)(*per))


def layout(lp):
    info = LayoutInfo()
    _check(lib().lfpm_layout_query(C.byref(lp), C.byref(info)))
    V, K = info.num_mis, lp.max_neighbors
    out = dict(V=V, crop=info.crop, cols=info.lattice_cols, rows=info.lattice_rows,
               row_step=info.row_step, pitch=lp.pitch, K=K,
               centers=np.zeros((V, 2), np.float32), origins=np.zeros((V, 2), np.int32),
               tilekeys=np.zeros((V, 2), np.int32), lens_type=np.zeros(V, np.int32),
               nbr=np.zeros((V, K), np.int32), cnt=np.zeros(V, np.int32),
               lattice_to_mi=np.zeros((info.lattice_rows, info.lattice_cols), np.int32))
    _check(lib().lfpm_layout_build(C.byref(lp), _np(out["centers"]), _np(out["origins"]),
                                   _np(out["tilekeys"]), _np(out["lens_type"]), _np(out["nbr"]),
                                   _np(out["cnt"]), _np(out["lattice_to_mi"])))
    return out


def fixed_neighbors(lp):
    """FixedPosition neighbour table (lfpm_layout_fixed_neighbors): nbr (V, K), cnt, garbage."""
    info = LayoutInfo()
    _check(lib().lfpm_layout_query(C.byref(lp), C.byref(info)))
    V, K = info.num_mis, lp.max_neighbors
    nbr = np.zeros((V, K), np.int32)
    cnt = np.zeros(V, np.int32)
    garbage = np.zeros(V, np.int32)
    _check(lib().lfpm_layout_fixed_neighbors(C.byref(lp), _np(nbr), _np(cnt), _np(garbage)))
    return nbr, cnt, garbage


def synth(sp, want_gt=True):
    raw = np.zeros((sp.raw_h, sp.raw_w), np.uint8)
    gt = np.zeros((sp.raw_h, sp.raw_w), np.float32) if want_gt else None
    _check(lib().lfpm_synth_frame(C.byref(sp), _np(raw), _np(gt) if want_gt else None))
    return raw, gt


def pm_default(crop, base):
    p = PmParams()
    lib().lfpm_pm_params_default(C.byref(p), int(crop), C.c_float(base))
    return p


def filter_default():
    f = FilterParams()
    lib().lfpm_filter_params_default(C.byref(f))
    return f


def workspace_size(V, pm):
    n = lib().lfpm_patchmatch_workspace_size(int(V), C.byref(pm))
    if n == 0:
        _check(-1)
    return n


# --------------------------------------------------------------------------------------
# Workloads (synthetic focused-plenoptic frames)
# --------------------------------------------------------------------------------------
class Workload:
    """One synthetic frame with its lens layout and solve/filter/refocus settings."""

    def __init__(self, raw_w, raw_h, pitch, K, radius, sigma_s, sigma_c, iterations, planes,
                 vmin, vmax, seed, scene_seed, noise_seed, noise_sigma=2.0, lens_types=1,
                 per_type=None, slices=32, frame_uid=0):
        self.lp = layout_params(pitch, raw_w, raw_h, K, lens_types=lens_types, per_type=per_type)
        self.sp = SynthParams(raw_w, raw_h, float(pitch), planes, float(vmin), float(vmax),
                              scene_seed, noise_seed, float(noise_sigma))
        self.L = layout(self.lp)
        pm = pm_default(self.L["crop"], float(pitch))
        pm.radius = radius
        pm.sigma_spatial = float(sigma_s)
        pm.sigma_color = float(sigma_c)
        pm.iterations = iterations
        pm.max_neighbors = K
        pm.seed = seed
        pm.frame_uid = frame_uid
        self.pm = pm
        self.fp = filter_default()
        self.slices = slices
        self.vmin, self.vmax = float(vmin), float(vmax)
        self._raw = None
        self._gt = None

    @property
    def V(self):
        return self.L["V"]

    @property
    def crop(self):
        return self.L["crop"]

    def raw(self):
        if self._raw is None:
            self._raw, self._gt = synth(self.sp)
        return self._raw

    def gt_virtual_depth(self):
        self.raw()
        return self._gt

    def blur_map(self):
        """Synthetic raw-size uint8 blur map: defocus grows with the distance of the scene's
        virtual depth from v = 4.5 (0 outside the MI discs), so MIs that see the same point
        carry similar blur."""
        v = self.gt_virtual_depth()
        b = np.where(v > 0, np.clip(np.abs(v - 4.5) / 6.0, 0.0, 1.0) * 255.0, 0.0)
        return np.floor(b + 0.5).astype(np.uint8)


def workload_c1(**kw):
    """1024x800 raw, pitch 23, 6 ring-1 neighbours, 11x11 window (s_s=3, s_c=0.1), 3 iterations,
    4 planes with virtual depth in [2.5, 6], Philox seed 1234."""
    args = dict(raw_w=1024, raw_h=800, pitch=23.0, K=6, radius=5, sigma_s=3.0, sigma_c=0.1,
                iterations=3, planes=4, vmin=2.5, vmax=6.0, seed=1234, scene_seed=1234,
                noise_seed=1235)
    args.update(kw)
    return Workload(**args)


def workload_c2(**kw):
    """4080x3068 raw, pitch 23 (~27k MIs), 18 ring-1+ring-2 neighbours, 15x15 window (s_s=5,
    s_c=0.12), 6 iterations, 12 planes with virtual depth in [2, 8], cycle check 0.5 px,
    32 refocus slices."""
    args = dict(raw_w=4080, raw_h=3068, pitch=23.0, K=18, radius=7, sigma_s=5.0, sigma_c=0.12,
                iterations=6, planes=12, vmin=2.0, vmax=8.0, seed=4321, scene_seed=4321,
                noise_seed=4322)
    args.update(kw)
    return Workload(**args)


def mixed_workloads(n, seed=11):
    """The uneven batch: raw size from {1024x800, 2048x1536, 4080x3068} and lens pitch from
    {15, 23, 31} per frame, multi-focus 3-type hex grid with 6/12/18 neighbours by lens type,
    6-12 seeded planes with virtual depth in [2, 8], 6 iterations."""
    rng = np.random.default_rng(seed)
    sizes = [(1024, 800), (2048, 1536), (4080, 3068)]
    radius = {15.0: 5, 23.0: 7, 31.0: 9}
    wls = []
    for i in range(n):
        w, h = sizes[int(rng.integers(3))]
        p = [15.0, 23.0, 31.0][int(rng.integers(3))]
        wls.append(workload_c2(raw_w=w, raw_h=h, pitch=p, radius=radius[p],
                               planes=int(rng.integers(6, 13)), lens_types=3,
                               per_type=(6, 12, 18), scene_seed=1000 + i,
                               noise_seed=2000 + i, seed=1000 + i, frame_uid=i))
    return wls


# --------------------------------------------------------------------------------------
# Device entry points (torch tensors on the current device)
# --------------------------------------------------------------------------------------
class DeviceFrame:
    """Device-resident inputs of one frame: raw, layout tables and tiles.  With blur_raw the
    tiles carry the MIs' blur values after the intensities, the layout the blur term of the
    cost vector reads (lfpm_pm_params.blur_lambda > 0): a raw-size uint8 blur map (value / 255),
    or "laplacian" for the reference's Laplacian blur score of the frame itself
    (lfpm_slice_blur_tiles)."""

    def __init__(self, wl, raw=None, device="cuda", blur_raw=None):
        import torch
        self.wl = wl
        L = wl.L
        raw = wl.raw() if raw is None else raw
        self.raw = torch.from_numpy(np.ascontiguousarray(raw)).to(device)
        self.centers = torch.from_numpy(L["centers"]).to(device)
        self.origins = torch.from_numpy(L["origins"]).to(device)
        self.nbr = torch.from_numpy(L["nbr"]).to(device)
        self.cnt = torch.from_numpy(L["cnt"]).to(device)
        self.l2m = torch.from_numpy(L["lattice_to_mi"]).to(device)
        V, W = L["V"], L["crop"]
        self.blur_lap = isinstance(blur_raw, str) and blur_raw == "laplacian"
        self.blur_raw = None if blur_raw is None or self.blur_lap else \
            torch.from_numpy(np.ascontiguousarray(blur_raw, np.uint8)).to(device)
        nt = V if blur_raw is None else 2 * V
        self.tiles = torch.empty((nt, W, W), dtype=torch.float32, device=device)
        self.gray = torch.empty((V, W, W), dtype=torch.float32, device=device)
        self.slice()

    def slice(self, stream=None):
        L = self.wl.L
        h, w = self.raw.shape
        _check(lib().lfpm_slice_tiles(_dp(self.raw), w, h, C.c_size_t(w), _dp(self.origins),
                                      L["V"], L["crop"], _dp(self.tiles), _stream(stream)))
        _check(lib().lfpm_slice_gray(_dp(self.raw), w, h, C.c_size_t(w), _dp(self.origins),
                                     L["V"], L["crop"], _dp(self.gray), _stream(stream)))
        if self.blur_raw is not None:
            _check(lib().lfpm_slice_tiles(_dp(self.blur_raw), w, h, C.c_size_t(w),
                                          _dp(self.origins), L["V"], L["crop"],
                                          _dp(self.tiles[L["V"]:]), _stream(stream)))
        if self.blur_lap:
            _check(lib().lfpm_slice_blur_tiles(_dp(self.raw), w, h, C.c_size_t(w),
                                               _dp(self.origins), L["V"], L["crop"],
                                               _dp(self.tiles[L["V"]:]), _stream(stream)))


def bgr_to_gray(bgr, out=None, stream=None):
    """Colour raw frame (H, W, 3) uint8 B, G, R on the device -> gray frame (H, W) uint8
    (lfpm_bgr_to_gray: cv::cvtColor(..., COLOR_BGR2GRAY) of DepthSolver.cpp:1985-1990).  Row
    pitches come from the tensors' strides."""
    import torch
    if bgr.dim() != 3 or bgr.shape[2] != 3 or bgr.dtype != torch.uint8 or \
            bgr.stride(2) != 1 or bgr.stride(1) != 3:
        raise ValueError("bgr_to_gray: expects (H, W, 3) uint8 with packed pixels")
    h, w = bgr.shape[0], bgr.shape[1]
    if out is None:
        out = torch.empty((h, w), dtype=torch.uint8, device=bgr.device)
    if out.shape != (h, w) or out.dtype != torch.uint8 or out.stride(1) != 1:
        raise ValueError("bgr_to_gray: out must be (H, W) uint8 with unit column stride")
    # row-pitched views are allowed: only the pixels of a row must be packed
    _check(lib().lfpm_bgr_to_gray(C.c_void_p(bgr.data_ptr()), w, h, C.c_size_t(bgr.stride(0)),
                                  C.c_void_p(out.data_ptr()), C.c_size_t(out.stride(0)),
                                  _stream(stream)))
    return out


def vd_point_cloud(vd, color=None, origin=(8.0, 8.0), pixel_size=1.0, main_lens_f=1.0, B=1.0,
                   bl0=0.0, ws=None, stream=None):
    """VD point cloud on the device (lfpm_vd_point_cloud): (records uint8 (n, 27), n).  vd and
    color are outputs of virtual_depth_map (color (raw_h, raw_w) gray or (raw_h, raw_w, 3)
    B, G, R: lfpm_vd_point_cloud_bgr); origin is the reference's (crop / 2, crop / 2)."""
    import torch
    h, w = vd.shape
    recs = torch.empty((h * w, 27), dtype=torch.uint8, device=vd.device)
    cnt = torch.zeros(1, dtype=torch.int32, device=vd.device)
    if ws is None:
        n = h * w + 1
        ws = torch.empty(2 * ((n * 4 + 255) // 256 * 256) + (n // 4096 + 2) * 4 + 512,
                         dtype=torch.uint8, device=vd.device)
    bgr = color is not None and color.dim() == 3
    fn = lib().lfpm_vd_point_cloud_bgr if bgr else lib().lfpm_vd_point_cloud
    _check(fn(
        _dp(vd), _dp(color) if color is not None else None, w, h, C.c_float(origin[0]),
        C.c_float(origin[1]), C.c_float(pixel_size), C.c_float(main_lens_f), C.c_float(B),
        C.c_float(bl0), _dp(recs), C.c_size_t(recs.numel()), _dp(cnt), _dp(ws),
        C.c_size_t(ws.numel()), _stream(stream)))
    n = int(cnt.item())
    return recs[:n], n


def write_ply_points(path, records):
    """lfpm_write_ply_points: the reference's PLY header + the (n, 27) records (host bytes)."""
    r = np.ascontiguousarray(records, np.uint8)
    _check(lib().lfpm_write_ply_points(str(path).encode(), r.ctypes.data_as(C.c_void_p),
                                       C.c_uint64(r.shape[0] if r.ndim == 2 else 0)))


def write_dmb(path, data):
    """lfpm_write_dmb: int32 {1, h, w, 1} + h*w float32 (writeDepthDmb)."""
    a = np.ascontiguousarray(data, np.float32)
    _check(lib().lfpm_write_dmb(str(path).encode(), a.ctypes.data_as(C.c_void_p),
                                a.shape[0], a.shape[1]))


def read_dmb(path):
    """lfpm_read_dmb: the (h, w) float32 map of a type-1 dmb file (readDepthDmb)."""
    h, w = C.c_int(), C.c_int()
    _check(lib().lfpm_read_dmb(str(path).encode(), None, C.c_size_t(0), C.byref(h), C.byref(w)))
    out = np.empty((h.value, w.value), np.float32)
    _check(lib().lfpm_read_dmb(str(path).encode(), out.ctypes.data_as(C.c_void_p),
                               C.c_size_t(out.size), C.byref(h), C.byref(w)))
    return out


class Intrinsics(C.Structure):
    _fields_ = [("D", C.c_float), ("d", C.c_float), ("sensor_pixel_size", C.c_float),
                ("main_lens_f", C.c_float), ("mla_u_size", C.c_int), ("mla_v_size", C.c_int),
                ("baseline", C.c_float), ("sensor_rotation", C.c_double * 9),
                ("sensor_translation", C.c_double * 3), ("distortion_depth", C.c_double * 3),
                ("distortion_radial", C.c_double * 3), ("distortion_tangential", C.c_double * 2)]


def read_intrinsics_xml(path):
    """lfpm_read_intrinsics_xml (DepthSolver::ReadIntrinsicsFromXML): an Intrinsics struct."""
    out = Intrinsics()
    _check(lib().lfpm_read_intrinsics_xml(str(path).encode(), C.byref(out)))
    return out


def read_calib_xml(path):
    """lfpm_read_calib_xml (DepthSolver::ReadCalibrationParamsFromXML): dict of the lens table in
    (row, col) order — tilekeys (n, 2) (col, row), centers (n, 2), lens_type, area — and
    blur_radius (3,) per lens type."""
    n = C.c_int()
    rad = np.zeros(3, np.float32)
    _check(lib().lfpm_read_calib_xml(str(path).encode(), 0, None, None, None, None,
                                     rad.ctypes.data_as(C.c_void_p), C.byref(n)))
    k = np.zeros((n.value, 2), np.int32)
    c = np.zeros((n.value, 2), np.float32)
    t = np.zeros(n.value, np.int32)
    a = np.zeros(n.value, np.float32)
    p = lambda x: x.ctypes.data_as(C.c_void_p)
    _check(lib().lfpm_read_calib_xml(str(path).encode(), n.value, p(k), p(c), p(t), p(a),
                                     p(rad), C.byref(n)))
    return dict(tilekeys=k, centers=c, lens_type=t, area=a, blur_radius=rad)


def table_origins(centers, crop, mla_u_size, mla_v_size, baseline, raw_w, raw_h):
    """lfpm_table_origins: (origins (n, 2) int32, kept (n,) int32) of a read lens table."""
    c = np.ascontiguousarray(centers, np.float32)
    n = c.shape[0]
    o = np.zeros((n, 2), np.int32)
    k = np.zeros(n, np.int32)
    p = lambda x: x.ctypes.data_as(C.c_void_p)
    _check(lib().lfpm_table_origins(p(c), n, crop, mla_u_size, mla_v_size, C.c_float(baseline),
                                    raw_w, raw_h, p(o), p(k)))
    return o, k


def table_ring_neighbors(tilekeys, centers, cols, rows, rings, max_neighbors):
    """Matching neighbour table of a read lens table (lfpm_table_ring_neighbors): the lenses on
    hex rings 1..rings ordered by centre distance, then angle: (nbr (V, K), cnt (V,))."""
    keys = np.ascontiguousarray(tilekeys, np.int32)
    cen = np.ascontiguousarray(centers, np.float32)
    V = keys.shape[0]
    nbr = np.zeros((V, max_neighbors), np.int32)
    cnt = np.zeros(V, np.int32)
    _check(lib().lfpm_table_ring_neighbors(_np(keys), _np(cen), V, int(cols), int(rows),
                                           int(rings), int(max_neighbors), _np(nbr), _np(cnt)))
    return nbr, cnt


def table_fixed_neighbors(tilekeys, centers, cols, rows, max_neighbors):
    """lfpm_table_fixed_neighbors: FixedPosition tables of a lens table -> (nbr, cnt, garbage)."""
    k = np.ascontiguousarray(tilekeys, np.int32)
    c = np.ascontiguousarray(centers, np.float32)
    V = k.shape[0]
    nbr = np.zeros((V, max_neighbors), np.int32)
    cnt = np.zeros(V, np.int32)
    g = np.zeros(V, np.int32)
    p = lambda x: x.ctypes.data_as(C.c_void_p)
    _check(lib().lfpm_table_fixed_neighbors(p(k), p(c), V, cols, rows, max_neighbors, p(nbr),
                                            p(cnt), p(g)))
    return nbr, cnt, g


class SolveBuffers:
    def __init__(self, V, W, pm, device="cuda"):
        import torch
        n = V * W * W
        self.plane = torch.empty((n, 4), dtype=torch.float32, device=device)
        self.cost = torch.empty(n, dtype=torch.float32, device=device)
        self.disp = torch.empty((n, 4), dtype=torch.float32, device=device)
        self.sel = torch.empty(n, dtype=torch.int32, device=device)
        self.ws_bytes = workspace_size(V, pm)
        self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device=device)


def patchmatch_frame(df, pm=None, bufs=None, sync_each_phase=False, stream=None):
    pm = pm or df.wl.pm
    V, W = df.wl.V, df.wl.crop
    bufs = bufs or SolveBuffers(V, W, pm, df.tiles.device)
    _check(lib().lfpm_patchmatch_frame(
        _dp(df.tiles), _dp(df.centers), _dp(df.nbr), _dp(df.cnt), V, C.byref(pm),
        _dp(bufs.plane), _dp(bufs.cost), _dp(bufs.disp), _dp(bufs.sel), _dp(bufs.ws),
        C.c_size_t(bufs.ws_bytes), 1 if sync_each_phase else 0, _stream(stream)))
    return bufs


def patchmatch_frame_acmm(df, pm=None, stream=None):
    """ACMM multi-scale solve (AdaptMIPMFrameACMM): SolveBuffers with plane/cost/disp/sel."""
    import torch
    pm = pm or df.wl.pm
    V, W = df.wl.V, df.wl.crop
    bufs = SolveBuffers(V, W, pm, df.tiles.device)
    n = lib().lfpm_patchmatch_acmm_workspace_size(V, C.byref(pm))
    if n == 0:
        _check(-1)
    ws = torch.empty(n, dtype=torch.uint8, device=df.tiles.device)
    _check(lib().lfpm_patchmatch_frame_acmm(
        _dp(df.tiles), _dp(df.centers), _dp(df.nbr), _dp(df.cnt), V, C.byref(pm),
        _dp(bufs.plane), _dp(bufs.cost), _dp(bufs.disp), _dp(bufs.sel), _dp(ws), C.c_size_t(n),
        _stream(stream)))
    return bufs


def eval_costs(df, hyp, pm=None, stream=None):
    """hyp: torch int32 (n, 6) on device, sorted by MI.  Returns (n, K) float32 costs."""
    import torch
    pm = pm or df.wl.pm
    n = hyp.shape[0]
    out = torch.empty((n, pm.max_neighbors), dtype=torch.float32, device=hyp.device)
    _check(lib().lfpm_eval_costs(_dp(df.tiles), _dp(df.centers), _dp(df.nbr), _dp(df.cnt),
                                 df.wl.V, C.byref(pm), _dp(hyp), n, _dp(out), _stream(stream)))
    return out


def filter_frame(df, bufs, fp=None, stream=None, order="snapshot"):
    """CheckModule pass over a solve (lfpm_filter_frame).  order="reference" runs
    lfpm_filter_frame_ordered: MIs in the reference's map order, rejections cascading to later
    MIs (MIDisparityFilterCPU::FilterGlobal, CheckModule/MIDisparityFilter.cpp:973-985), on the
    device frame's level schedule (built once, see filter_levels)."""
    import torch
    if order not in ("snapshot", "reference"):
        raise ValueError("filter_frame: order must be 'snapshot' or 'reference'")
    fp = fp or df.wl.fp
    V, W = df.wl.V, df.wl.crop
    n = V * W * W
    dev = bufs.plane.device
    out = dict(plane=torch.empty((n, 4), dtype=torch.float32, device=dev),
               cost=torch.empty(n, dtype=torch.float32, device=dev),
               disp=torch.empty((n, 4), dtype=torch.float32, device=dev),
               sel=torch.empty(n, dtype=torch.int32, device=dev),
               reason=torch.empty(n, dtype=torch.uint8, device=dev))
    args = (_dp(df.gray), _dp(df.centers), _dp(df.nbr), _dp(df.cnt), V, W,
            df.wl.pm.max_neighbors, C.c_float(df.wl.pm.base), C.byref(fp), _dp(bufs.plane),
            _dp(bufs.cost), _dp(bufs.disp), _dp(bufs.sel), _dp(out["plane"]), _dp(out["cost"]),
            _dp(out["disp"]), _dp(out["sel"]), _dp(out["reason"]))
    if order == "snapshot":
        _check(lib().lfpm_filter_frame(*args, _stream(stream)))
        return out
    order_d, start_h = _df_levels(df)
    _check(lib().lfpm_filter_frame_ordered(
        *args, _dp(order_d), start_h.ctypes.data_as(C.POINTER(C.c_int)), len(start_h) - 1,
        _stream(stream)))
    return out


def _df_levels(df):
    """The ordered filter's level schedule of a device frame's layout, built once: (MI order on
    the device, level boundaries on the host)."""
    import torch
    sched = getattr(df, "_filter_levels", None)
    if sched is None:
        order_h, start_h = filter_levels(df.wl.L["nbr"], df.wl.L["cnt"], df.wl.pm.max_neighbors)
        sched = (torch.from_numpy(order_h).to(df.nbr.device), start_h)
        df._filter_levels = sched
    return sched


def filter_levels(nbr, cnt, K):
    """Level schedule of the ordered filter (lfpm_filter_order_levels, host only): the MIs level
    by level (int32 [V]) and the level boundaries (int32 [levels + 1])."""
    nbr = np.ascontiguousarray(nbr, np.int32).reshape(-1)
    cnt = np.ascontiguousarray(cnt, np.int32)
    V = cnt.shape[0]
    if nbr.shape[0] != V * K:
        raise ValueError("filter_levels: neighbour table is not V x K")
    order = np.zeros(V, np.int32)
    start = np.zeros(V + 1, np.int32)
    nl = C.c_int(0)
    P = C.POINTER(C.c_int)
    _check(lib().lfpm_filter_order_levels(nbr.ctypes.data_as(P), cnt.ctypes.data_as(P), V, K,
                                          order.ctypes.data_as(P), start.ctypes.data_as(P),
                                          C.byref(nl)))
    return order, start[:nl.value + 1].copy()


def refocus(df, slices=None, vmin=None, vmax=None, out=None, stream=None):
    import torch
    wl = df.wl
    S = slices or wl.slices
    vmin = wl.vmin if vmin is None else vmin
    vmax = wl.vmax if vmax is None else vmax
    h, w = df.raw.shape
    if out is None:
        out = torch.empty((S, h, w), dtype=torch.float32, device=df.tiles.device)
    _check(lib().lfpm_refocus(_dp(df.tiles), _dp(df.centers), _dp(df.l2m), wl.L["cols"],
                              wl.L["rows"], C.c_float(wl.L["pitch"]), wl.crop, w, h, S,
                              C.c_float(vmin), C.c_float(vmax), _dp(out), _stream(stream)))
    return out


def products_workspace(df):
    import torch
    wl = df.wl
    h, w = df.raw.shape
    n = lib().lfpm_products_workspace_size(wl.V, wl.crop, w, h)
    return torch.empty(n, dtype=torch.uint8, device=df.tiles.device)


def disparity_mosaic(df, plane, cost, max_cost=0.2, ws=None, stream=None):
    """Raw-size disparity mosaic of a solved frame (WriteDisMap_TileKey_new_AccuSequence,
    DepthSolver.cpp:6716-6897).  Returns (mosaic float32 (h, w), range float32 (2,))."""
    import torch
    wl = df.wl
    h, w = df.raw.shape
    ws = products_workspace(df) if ws is None else ws
    out = torch.empty((h, w), dtype=torch.float32, device=plane.device)
    rng = torch.empty(2, dtype=torch.float32, device=plane.device)
    _check(lib().lfpm_disparity_mosaic(_dp(plane), _dp(cost), _dp(df.centers), wl.V, wl.crop,
                                       C.c_float(max_cost), w, h, _dp(out), _dp(rng), _dp(ws),
                                       C.c_size_t(ws.numel()), _stream(stream)))
    return out, rng


def slice_bgr_tiles(df, bgr, stream=None):
    """Colour crops (V, crop, crop, 3) B, G, R of a device colour frame bgr (h, w, 3) uint8 at
    the frame's MI origins (lfpm_slice_bgr_tiles)."""
    import torch
    wl = df.wl
    h, w = bgr.shape[:2]
    assert bgr.is_contiguous() and bgr.shape[2] == 3
    out = torch.empty((wl.V, wl.crop, wl.crop, 3), dtype=torch.uint8, device=bgr.device)
    _check(lib().lfpm_slice_bgr_tiles(_dp(bgr), w, h, C.c_size_t(bgr.stride(0)),
                                      _dp(df.origins), wl.V, wl.crop, _dp(out), _stream(stream)))
    return out


def virtual_depth_map(df, plane, mla_base=None, color=True, ws=None, stream=None,
                      bgr_tiles=None):
    """Virtual-depth median splat (Virtual_depth_map_TileKeySequence, DepthSolver.cpp:7832-8000).
    Returns (vd float32 (h, w), colour or None, range float32 (2,)); the colour image is uint8
    (h, w) from the gray crops, or (h, w, 3) B, G, R from the colour crops bgr_tiles
    (slice_bgr_tiles; lfpm_virtual_depth_map_bgr)."""
    import torch
    wl = df.wl
    h, w = df.raw.shape
    mla_base = wl.L["pitch"] if mla_base is None else mla_base
    ws = products_workspace(df) if ws is None else ws
    vd = torch.empty((h, w), dtype=torch.float32, device=plane.device)
    rng = torch.empty(2, dtype=torch.float32, device=plane.device)
    if bgr_tiles is not None:
        col = torch.empty((h, w, 3), dtype=torch.uint8, device=plane.device)
        _check(lib().lfpm_virtual_depth_map_bgr(_dp(plane), _dp(df.centers), _dp(bgr_tiles),
                                                wl.V, wl.crop, C.c_float(mla_base), w, h,
                                                _dp(vd), _dp(col), _dp(rng), _dp(ws),
                                                C.c_size_t(ws.numel()), _stream(stream)))
        return vd, col, rng
    col = torch.empty((h, w), dtype=torch.uint8, device=plane.device) if color else None
    _check(lib().lfpm_virtual_depth_map(_dp(plane), _dp(df.centers), _dp(df.gray), wl.V,
                                        wl.crop, C.c_float(mla_base), w, h, _dp(vd),
                                        _dp(col) if col is not None else None, _dp(rng), _dp(ws),
                                        C.c_size_t(ws.numel()), _stream(stream)))
    return vd, col, rng


def aif_composite(df, plane, bgr_tiles, richness=None, blur=None, garbage=None, baseline=None,
                  B=1.0, bl0=0.0, main_lens_f=105.0, stream=None):
    """All-in-focus composite of a solved frame (lfpm_aif_composite: AIFImageCompositeForMIA,
    FocalStack/LFRefocus.cpp:183-425).  bgr_tiles: colour crops (V, crop, crop, 3); richness /
    blur: optional uint8 crops (V, crop, crop) of the binarised richness and the blur score;
    garbage: optional uint8 (V,).  baseline defaults to the lens pitch.  Returns a dict of
    device tensors vd, inv_vd, real_mm (h, w) float32, aif, filled, denoised (h, w, 3) uint8 and
    stats (8,) int32."""
    import torch
    wl = df.wl
    h, w = df.raw.shape
    dev = plane.device
    baseline = wl.L["pitch"] if baseline is None else baseline
    n = lib().lfpm_aif_workspace_size(wl.V, wl.crop, w, h)
    ws = torch.empty(n, dtype=torch.uint8, device=dev)
    out = dict(vd=torch.empty((h, w), dtype=torch.float32, device=dev),
               inv_vd=torch.empty((h, w), dtype=torch.float32, device=dev),
               real_mm=torch.empty((h, w), dtype=torch.float32, device=dev),
               aif=torch.empty((h, w, 3), dtype=torch.uint8, device=dev),
               filled=torch.empty((h, w, 3), dtype=torch.uint8, device=dev),
               denoised=torch.empty((h, w, 3), dtype=torch.uint8, device=dev),
               stats=torch.empty(8, dtype=torch.int32, device=dev))
    opt = lambda t: _dp(t) if t is not None else None  # noqa: E731
    _check(lib().lfpm_aif_composite(
        _dp(plane), _dp(df.centers), _dp(bgr_tiles), opt(richness), opt(blur), opt(garbage),
        wl.V, wl.crop, C.c_float(baseline), C.c_float(B), C.c_float(bl0),
        C.c_float(main_lens_f), w, h, _dp(out["vd"]), _dp(out["inv_vd"]), _dp(out["real_mm"]),
        _dp(out["aif"]), _dp(out["filled"]), _dp(out["denoised"]), _dp(out["stats"]), _dp(ws),
        C.c_size_t(n), _stream(stream)))
    return out


def fuse_schedule(nbr, cnt, garbage=None):
    """Level schedule of a neighbour table (lfpm_fuse_schedule, host): (order int32 (V,),
    level_start int32 (levels + 1,), levels)."""
    nbr = np.ascontiguousarray(nbr, np.int32)
    cnt = np.ascontiguousarray(cnt, np.int32)
    V = cnt.shape[0]
    K = nbr.size // V
    g = None if garbage is None else np.ascontiguousarray(garbage, np.uint8)
    order = np.zeros(V, np.int32)
    start = np.zeros(V + 1, np.int32)
    nl = C.c_int()
    _check(lib().lfpm_fuse_schedule(_np(nbr), _np(cnt), _np(g) if g is not None else None, V, K,
                                    _np(order), _np(start), C.byref(nl)))
    return order, start[:nl.value + 1].copy(), nl.value


def fuse_backproject(df, plane, bgr_tiles, garbage=None, baseline=None, B=1.0, bl0=0.0,
                     points=True, schedule=None, stream=None):
    """Back-projection consistency fusion of a solved frame (lfpm_fuse_backproject:
    FuseVirtualDepth_BackProject, FocalStack/LFRefocus.cpp:1553-1793) over the frame's matching
    table.  Returns a dict of device tensors vd (h, w) float32, aif (h, w, 3) uint8, conf (h, w)
    uint8, points (n, 9) float32 (or None), and ints npoints, nfused, levels."""
    import torch
    wl = df.wl
    h, w = df.raw.shape
    dev = plane.device
    baseline = wl.L["pitch"] if baseline is None else baseline
    K = df.nbr.numel() // wl.V
    gh = None if garbage is None else np.ascontiguousarray(
        garbage.cpu().numpy() if hasattr(garbage, "cpu") else garbage, np.uint8)
    if schedule is None:
        schedule = fuse_schedule(df.nbr.cpu().numpy(), df.cnt.cpu().numpy(), gh)
    order, start, nl = schedule
    order_d = torch.from_numpy(order).to(dev)
    gd = None if gh is None else torch.from_numpy(gh).to(dev)
    n = lib().lfpm_fuse_workspace_size(wl.V, wl.crop, w, h)
    ws = torch.empty(n, dtype=torch.uint8, device=dev)
    M = wl.V * wl.crop * wl.crop
    pts = torch.empty((M, 9), dtype=torch.float32, device=dev) if points else None
    out = dict(vd=torch.empty((h, w), dtype=torch.float32, device=dev),
               aif=torch.empty((h, w, 3), dtype=torch.uint8, device=dev),
               conf=torch.empty((h, w), dtype=torch.uint8, device=dev))
    counts = torch.zeros(2, dtype=torch.int64, device=dev)
    _check(lib().lfpm_fuse_backproject(
        _dp(plane), _dp(df.centers), _dp(bgr_tiles), _dp(df.nbr), _dp(df.cnt), K,
        _dp(gd) if gd is not None else None, _dp(order_d), _np(start), nl, wl.V, wl.crop,
        C.c_float(baseline), C.c_float(B), C.c_float(bl0), w, h, _dp(out["vd"]),
        _dp(out["aif"]), _dp(out["conf"]), _dp(pts) if pts is not None else None,
        C.c_size_t(pts.numel() * 4 if pts is not None else 0), _dp(counts), _dp(ws),
        C.c_size_t(n), _stream(stream)))
    c = counts.cpu().numpy()
    out.update(npoints=int(c[0]), nfused=int(c[1]), levels=nl,
               points=pts[:int(c[0])] if pts is not None else None)
    return out


class Context:
    """Host-level context (lfpm_create / lfpm_ctx_*): host frames in, host results out; device
    buffers stay resident across calls of the same layout."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        _check(lib().lfpm_create(device, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().lfpm_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    PLAIN, ACMM = 0, 1

    def set_variant(self, variant):
        """Solver used by solve_frame / solve_tiles: Context.PLAIN or Context.ACMM."""
        _check(lib().lfpm_ctx_set_variant(self.h, int(variant)))

    def set_filter_order(self, order):
        """'snapshot' (LFPM_FILTER_SNAPSHOT, the default) or 'reference'
        (LFPM_FILTER_REFERENCE_ORDER): the order of the filter pass of every later solve."""
        codes = {"snapshot": 0, "reference": 1}
        if order not in codes:
            raise ValueError("set_filter_order: order must be 'snapshot' or 'reference'")
        _check(lib().lfpm_ctx_set_filter_order(self.h, codes[order]))

    @staticmethod
    def _outs(n):
        return dict(plane=np.zeros((n, 4), np.float32), cost=np.zeros(n, np.float32),
                    disp=np.zeros((n, 4), np.float32), sel=np.zeros(n, np.uint32),
                    reason=np.zeros(n, np.uint8))

    def solve_frame(self, wl, raw=None, filt=True, pm=None):
        raw = np.ascontiguousarray(wl.raw() if raw is None else raw)
        o = self._outs(wl.V * wl.crop ** 2)
        _check(lib().lfpm_ctx_solve_frame(self.h, C.byref(wl.lp), _np(raw), C.byref(pm or wl.pm),
                                          C.byref(wl.fp), _np(o["plane"]), _np(o["cost"]),
                                          _np(o["disp"]), _np(o["sel"]),
                                          _np(o["reason"]) if filt else None))
        return o

    def solve_frames(self, wl, raws, filt=True, pm=None):
        """Frames of wl's layout in one batched solve; frame i uses frame_uid + i."""
        n = len(raws)
        raws = [np.ascontiguousarray(r) for r in raws]
        outs = [self._outs(wl.V * wl.crop ** 2) for _ in range(n)]
        VP = C.c_void_p * n

        def arr(key):
            return VP(*[o[key].ctypes.data for o in outs])
        _check(lib().lfpm_ctx_solve_frames(
            self.h, C.byref(wl.lp), n, VP(*[r.ctypes.data for r in raws]), C.byref(pm or wl.pm),
            C.byref(wl.fp), arr("plane"), arr("cost"), arr("disp"), arr("sel"),
            arr("reason") if filt else None))
        return outs

    def solve_tiles(self, tiles_u8, centers, nbr, cnt, pm, fp=None):
        V = centers.shape[0]
        o = self._outs(V * pm.crop ** 2)
        tiles_u8 = np.ascontiguousarray(tiles_u8, np.uint8)
        centers = np.ascontiguousarray(centers, np.float32)
        nbr = np.ascontiguousarray(nbr, np.int32)
        cnt = np.ascontiguousarray(cnt, np.int32)
        _check(lib().lfpm_ctx_solve_tiles(self.h, _np(tiles_u8), _np(centers), _np(nbr), _np(cnt),
                                          V, C.byref(pm), C.byref(fp) if fp else None,
                                          _np(o["plane"]), _np(o["cost"]), _np(o["disp"]),
                                          _np(o["sel"]), _np(o["reason"]) if fp else None))
        return o

    def frame_products(self, raw_w, raw_h, mla_base, max_cost=0.2, frame=0):
        mosaic = np.zeros((raw_h, raw_w), np.float32)
        vd = np.zeros((raw_h, raw_w), np.float32)
        col = np.zeros((raw_h, raw_w), np.uint8)
        mr = np.zeros(2, np.float32)
        vr = np.zeros(2, np.float32)
        _check(lib().lfpm_ctx_frame_products(self.h, frame, raw_w, raw_h, C.c_float(max_cost),
                                             C.c_float(mla_base), _np(mosaic), _np(mr), _np(vd),
                                             _np(col), _np(vr)))
        return dict(mosaic=mosaic, mosaic_range=mr, vd=vd, color=col, vd_range=vr)


def fit_v2r(a_l, v, stream=None):
    c = (C.c_double * 3)()
    _check(lib().lfpm_fit_v2r(_dp(a_l), _dp(v), a_l.numel(), c, _stream(stream)))
    return np.array(list(c))


def v2r_convert(v, c, stream=None):
    import torch
    z = torch.empty(v.numel(), dtype=torch.float64, device=v.device)
    cc = (C.c_double * 3)(*[float(x) for x in c])
    _check(lib().lfpm_v2r_convert(_dp(v), v.numel(), cc, _dp(z), _stream(stream)))
    return z


def ncc_sample_counts(pm):
    """FLOP accounting constants: 22 FLOP per evaluated window sample, 16 per finalize."""
    return 22, 16


def hypotheses_from_arrays(mi, pix, planes):
    """Pack (n,) mi, (n,) pixel, (n,4) float32 planes into the (n,6) int32 record layout."""
    n = mi.shape[0]
    rec = np.zeros((n, 6), np.int32)
    rec[:, 0] = mi
    rec[:, 1] = pix
    rec[:, 2:] = np.ascontiguousarray(planes, np.float32).view(np.int32)
    return rec


def cone_hypotheses(wl, n, seed, vmin=2.0, vmax=8.0, cone_deg=60.0):
    """n seeded (MI, pixel, plane) records sorted by MI: normals uniform on the spherical cap of
    `cone_deg` about +z, virtual depth uniform in [vmin, vmax] -> disparity pitch / v."""
    rng = np.random.default_rng(seed)
    V, W = wl.V, wl.crop
    mi = np.sort(rng.integers(0, V, n)).astype(np.int32)
    pix = rng.integers(0, W * W, n).astype(np.int32)
    cz = 1.0 - rng.random(n) * (1.0 - np.cos(np.radians(cone_deg)))
    phi = 2 * np.pi * rng.random(n)
    sz = np.sqrt(np.maximum(0.0, 1.0 - cz * cz))
    v = vmin + (vmax - vmin) * rng.random(n)
    planes = np.stack([sz * np.cos(phi), sz * np.sin(phi), cz, wl.L["pitch"] / v], 1)
    return hypotheses_from_arrays(mi, pix, planes.astype(np.float32))


def version():
    return "lfpm-0.1"


_ = math


# --------------------------------------------------------------------------------------
# Batched / served forms
# --------------------------------------------------------------------------------------
STAGE_IDLE, STAGE_INIT, STAGE_PHASE, STAGE_FINAL = 0, 1, 2, 3


class BatchFrame(C.Structure):
    _fields_ = [("tiles_dev", C.c_void_p), ("gray_dev", C.c_void_p), ("centers_dev", C.c_void_p),
                ("neighbor_ids_dev", C.c_void_p), ("neighbor_counts_dev", C.c_void_p),
                ("lattice_to_mi_dev", C.c_void_p), ("num_mis", C.c_int),
                ("lattice_cols", C.c_int), ("lattice_rows", C.c_int), ("raw_w", C.c_int),
                ("raw_h", C.c_int), ("params", C.POINTER(PmParams)),
                ("filter", C.POINTER(FilterParams)),
                ("plane_dev", C.c_void_p), ("cost_dev", C.c_void_p), ("sel_dev", C.c_void_p),
                ("disp_dev", C.c_void_p), ("plane_b_dev", C.c_void_p), ("cost_b_dev", C.c_void_p),
                ("sel_b_dev", C.c_void_p), ("f_plane_dev", C.c_void_p), ("f_cost_dev", C.c_void_p),
                ("f_disp_dev", C.c_void_p), ("f_sel_dev", C.c_void_p), ("reason_dev", C.c_void_p),
                ("refocus_slices", C.c_int), ("refocus_vmin", C.c_float),
                ("refocus_vmax", C.c_float), ("refocus_dev", C.c_void_p),
                ("filter_mi_order_dev", C.c_void_p),
                ("filter_level_start", C.POINTER(C.c_int)), ("filter_num_levels", C.c_int)]


class FrameSlot:
    """Device-resident inputs + state + products of one frame in a batch (one serving slot)."""

    def __init__(self, wl, device="cuda", filt=True, refocus=True, blur_raw=None, pm=None,
                 filter_order="snapshot"):
        import torch
        if filter_order not in ("snapshot", "reference"):
            raise ValueError("FrameSlot: filter_order must be 'snapshot' or 'reference'")
        self.wl = wl
        self.filter_order = filter_order
        self.df = DeviceFrame(wl, device=device, blur_raw=blur_raw)
        V, W = wl.V, wl.crop
        n = V * W * W
        f32 = dict(dtype=torch.float32, device=device)
        self.plane = torch.empty((n, 4), **f32)
        self.cost = torch.empty(n, **f32)
        self.sel = torch.empty(n, dtype=torch.int32, device=device)
        self.disp = torch.empty((n, 4), **f32)
        self.plane_b = torch.empty((n, 4), **f32)
        self.cost_b = torch.empty(n, **f32)
        self.sel_b = torch.empty(n, dtype=torch.int32, device=device)
        self.filt = filt
        if filt:
            self.f_plane = torch.empty((n, 4), **f32)
            self.f_cost = torch.empty(n, **f32)
            self.f_disp = torch.empty((n, 4), **f32)
            self.f_sel = torch.empty(n, dtype=torch.int32, device=device)
            self.reason = torch.empty(n, dtype=torch.uint8, device=device)
        self.slices = wl.slices if refocus else 0
        h, w = self.df.raw.shape
        self.refocus = torch.empty((self.slices, h, w), **f32) if refocus else None
        self.pm = pm if pm is not None else wl.pm
        self.fp = wl.fp

    def nbytes(self):
        ts = [self.df.raw, self.df.tiles, self.df.gray, self.plane, self.cost, self.sel, self.disp,
              self.plane_b, self.cost_b, self.sel_b]
        if self.filt:
            ts += [self.f_plane, self.f_cost, self.f_disp, self.f_sel, self.reason]
        if self.refocus is not None:
            ts.append(self.refocus)
        return sum(t.numel() * t.element_size() for t in ts)

    def desc(self):
        d = BatchFrame()
        df, L = self.df, self.wl.L
        P = lambda t: t.data_ptr() if t is not None else None  # noqa: E731
        d.tiles_dev, d.gray_dev, d.centers_dev = P(df.tiles), P(df.gray), P(df.centers)
        d.neighbor_ids_dev, d.neighbor_counts_dev = P(df.nbr), P(df.cnt)
        d.lattice_to_mi_dev = P(df.l2m)
        d.num_mis, d.lattice_cols, d.lattice_rows = self.wl.V, L["cols"], L["rows"]
        d.raw_h, d.raw_w = df.raw.shape
        d.params = C.pointer(self.pm)
        d.filter = C.pointer(self.fp) if self.filt else None
        d.plane_dev, d.cost_dev, d.sel_dev, d.disp_dev = (P(self.plane), P(self.cost),
                                                          P(self.sel), P(self.disp))
        d.plane_b_dev, d.cost_b_dev, d.sel_b_dev = P(self.plane_b), P(self.cost_b), P(self.sel_b)
        if self.filt:
            d.f_plane_dev, d.f_cost_dev, d.f_disp_dev = (P(self.f_plane), P(self.f_cost),
                                                         P(self.f_disp))
            d.f_sel_dev, d.reason_dev = P(self.f_sel), P(self.reason)
            if self.filter_order == "reference":
                order_d, start_h = _df_levels(df)
                d.filter_mi_order_dev = P(order_d)
                d.filter_level_start = start_h.ctypes.data_as(C.POINTER(C.c_int))
                d.filter_num_levels = len(start_h) - 1
        d.refocus_slices = self.slices
        d.refocus_vmin, d.refocus_vmax = self.wl.vmin, self.wl.vmax
        d.refocus_dev = P(self.refocus)
        return d


def _frames_array(slots):
    arr = (BatchFrame * len(slots))()
    for i, s in enumerate(slots):
        arr[i] = s.desc()
    return arr


class Batch:
    """A fixed set of frames solved together (eagerly or through one captured hipGraph)."""

    def __init__(self, slots, device="cuda"):
        import torch
        self.slots = list(slots)
        self.arr = _frames_array(self.slots)
        n = lib().lfpm_batch_workspace_size(self.arr, len(self.slots))
        if n == 0:
            _check(-1)
        self.ws_bytes = n
        self.ws = torch.empty(n, dtype=torch.uint8, device=device)
        self.graph = None

    def run(self, stream=None):
        _check(lib().lfpm_batch_run(self.arr, len(self.slots), _dp(self.ws),
                                    C.c_size_t(self.ws_bytes), _stream(stream)))

    def capture(self):
        g = C.c_void_p()
        _check(lib().lfpm_batch_graph_create(self.arr, len(self.slots), _dp(self.ws),
                                             C.c_size_t(self.ws_bytes), C.byref(g)))
        self.graph = g
        return g

    def replay(self, stream=None):
        if self.graph is None:
            self.capture()
        _check(lib().lfpm_graph_launch(self.graph, _stream(stream)))

    def close(self):
        if self.graph is not None:
            lib().lfpm_graph_destroy(self.graph)
            self.graph = None

    def step(self, stages, iters, phases, stream=None):
        n = len(self.slots)
        st = (C.c_int * n)(*stages)
        it = (C.c_int * n)(*iters)
        ph = (C.c_int * n)(*phases)
        _check(lib().lfpm_batch_step(self.arr, n, st, it, ph, _dp(self.ws),
                                     C.c_size_t(self.ws_bytes), _stream(stream)))

    def autotune(self, i=0):
        best = C.c_int()
        ms = C.c_float()
        a1 = (BatchFrame * 1)(self.arr[i])
        _check(lib().lfpm_autotune_class(a1, _dp(self.ws), C.c_size_t(self.ws_bytes),
                                         C.byref(best), C.byref(ms)))
        return best.value, ms.value


def autotune_get(pm):
    """Launch shape the solver uses for pm's parameter class: (workgroups per CU, wavefronts
    per workgroup) (lfpm_autotune_get)."""
    per_cu, waves = C.c_int(), C.c_int()
    _check(lib().lfpm_autotune_get(C.byref(pm), C.byref(per_cu), C.byref(waves)))
    return per_cu.value, waves.value
