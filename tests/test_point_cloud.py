"""VD point cloud and the product file formats.

lfpm_vd_point_cloud builds the records of DepthSolver::Virtual_depth_map_TileKeySequence
(DepthSolver.cpp:8019-8053) in the binary layout of StoreColorPlyFileBinaryPointCloud
(MVStereo/ACMP.cpp:451-500); lfpm_write_ply_points / lfpm_write_dmb / lfpm_read_dmb are that
writer and writeDepthDmb / readDepthDmb (ACMP.cpp:338-390).

The oracle's point map is pinned by the thin-lens relation it encodes (1/z + 1/tz = 1/f, lateral
magnification x/tx = y/ty = -z/tz), an exactly representable case (including the sign of a
zero), the writer's non-finite rule, and row-major order.  The GPU records are compared with the
oracle's byte for byte; the files with the format written out field by field."""
import ctypes as C
import struct

import numpy as np
import pytest


def _f(rec):
    return rec[:, :24].copy().view("<f4").reshape(-1, 6)


def test_thin_lens_relation(orc):
    rng = np.random.default_rng(0)
    vd = rng.uniform(2.0, 9.0, size=(40, 50)).astype(np.float32)
    f, B, bl0, pix = 35.0, 0.8, 2.5, 0.0055
    rec = orc.vd_point_cloud(vd, None, (8.0, 8.0), pix, f, B, bl0)
    p = _f(rec).astype(np.float64)
    ys, xs = np.nonzero(vd > 0)
    tz = vd[ys, xs].astype(np.float64) * B + bl0
    tx = (xs - 8.0) * np.float32(pix)
    ty = (ys - 8.0) * np.float32(pix)
    assert np.allclose(1.0 / p[:, 2] + 1.0 / tz, 1.0 / f, rtol=1e-5, atol=0)
    m = -p[:, 2] / tz
    ok = np.abs(tx) > 1e-3
    assert np.allclose(p[ok, 0] / tx[ok], m[ok], rtol=1e-5)
    ok = np.abs(ty) > 1e-3
    assert np.allclose(p[ok, 1] / ty[ok], m[ok], rtol=1e-5)
    assert (p[:, 3:] == 0).all()


def test_exact_case_and_signed_zero(orc):
    vd = np.zeros((2, 4), np.float32)
    vd[1, 3] = 2.0
    # tz = 2 * 0.5 + 1 = 2, f = 1: z = 1 / 1 * 2 = 2, x = -(3 - 1) * 0.25 = -0.5, y = -0 * 0 = -0
    rec = orc.vd_point_cloud(vd, np.full((2, 4), 77, np.uint8), (1.0, 1.0), 0.25, 1.0, 0.5, 1.0)
    assert rec.shape == (1, 27)
    x, y, z, nx, ny, nz = struct.unpack("<6f", rec[0, :24].tobytes())
    assert (x, z, nx, ny, nz) == (-0.5, 2.0, 0.0, 0.0, 0.0)
    assert y == 0.0 and np.signbit(y)
    assert rec[0, 24:].tolist() == [77, 77, 77]


def test_non_finite_points_written_as_zero(orc):
    vd = np.array([[1.0, 3.0]], np.float32)
    # tz = vd * 1 + 0 = f = 1 at the first pixel: division by zero
    rec = orc.vd_point_cloud(vd, None, (0.0, 0.0), 1.0, 1.0, 1.0, 0.0)
    p = _f(rec)
    assert (p[0, :3] == 0).all() and not np.signbit(p[0, :3]).any()
    assert np.isfinite(p[1]).all() and p[1, 2] != 0


def test_row_major_order_and_empty(orc):
    vd = np.zeros((3, 3), np.float32)
    vd[2, 0], vd[0, 2], vd[1, 1] = 4.0, 5.0, 6.0
    vd[0, 0] = -1.0  # no sample
    rec = orc.vd_point_cloud(vd, None, (0.0, 0.0), 1.0, 1.0, 1.0, 5.0)
    tz = _f(rec)[:, 2]
    # tz = vd + 5 -> z = tz / (tz - 1): in the order (0,2), (1,1), (2,0) of row-major pixels
    want = np.array([10.0 / 9.0, 11.0 / 10.0, 9.0 / 8.0])
    assert np.allclose(tz, want, rtol=1e-6)
    assert orc.vd_point_cloud(np.zeros((4, 4), np.float32), None, (0, 0), 1, 1, 1, 0).shape == (0, 27)


def test_ply_header_and_payload(lfpm, tmp_path):
    rec = np.arange(2 * 27, dtype=np.uint8).reshape(2, 27)
    path = tmp_path / "VD.ply"
    lfpm.write_ply_points(path, rec)
    data = path.read_bytes()
    head = ("ply\nformat binary_little_endian 1.0\nelement vertex 2\n"
            "property float x\nproperty float y\nproperty float z\n"
            "property float nx\nproperty float ny\nproperty float nz\n"
            "property uchar red\nproperty uchar green\nproperty uchar blue\nend_header\n")
    assert data[:len(head)].decode() == head
    assert data[len(head):] == rec.tobytes()


def test_dmb_round_trip_and_type_check(lfpm, tmp_path):
    a = np.random.default_rng(1).normal(size=(7, 5)).astype(np.float32)
    path = tmp_path / "d.dmb"
    lfpm.write_dmb(path, a)
    raw = path.read_bytes()
    assert struct.unpack("<4i", raw[:16]) == (1, 7, 5, 1)
    assert raw[16:] == a.tobytes()
    assert np.array_equal(lfpm.read_dmb(path), a)
    bad = tmp_path / "bad.dmb"
    bad.write_bytes(struct.pack("<4i", 2, 1, 1, 1) + b"\0" * 4)
    with pytest.raises(Exception):
        lfpm.read_dmb(bad)
    short = tmp_path / "short.dmb"
    short.write_bytes(raw[:30])
    with pytest.raises(Exception):
        lfpm.read_dmb(short)


def test_point_cloud_rejects_bad_arguments(lfpm):
    L = lfpm.lib()
    one = C.c_void_p(1)
    args = (C.c_float(8), C.c_float(8), C.c_float(1), C.c_float(1), C.c_float(1), C.c_float(0))
    assert L.lfpm_vd_point_cloud(None, None, 4, 4, *args, one, C.c_size_t(432), one, one,
                                 C.c_size_t(1 << 20), None) != 0
    assert L.lfpm_vd_point_cloud(one, None, 4, 4, *args, one, C.c_size_t(431), one, one,
                                 C.c_size_t(1 << 20), None) != 0
    assert L.lfpm_vd_point_cloud(one, None, 4, 4, *args, one, C.c_size_t(432), one, one,
                                 C.c_size_t(16), None) != 0


def _say(*a):
    print(*a, flush=True)


@pytest.mark.gpu
@pytest.mark.parametrize("w,h", [(4080, 3068), (333, 77), (1, 1)])
def test_point_cloud_matches_oracle(lfpm, orc, w, h):
    import torch
    _say("point cloud", w, h)
    rng = np.random.default_rng(w + h)
    vd = rng.uniform(1.5, 9.0, size=(h, w)).astype(np.float32)
    vd[rng.random((h, w)) < 0.4] = 0.0
    if w > 1:
        vd[0, 0] = 40.625  # tz = 40.625 * 0.8 + 2.5 == f: the non-finite rule
    col = rng.integers(0, 256, size=(h, w), dtype=np.uint8)
    # B and bl0 away from 1 and 0: a fused multiply-add in tz would change the results
    prm = dict(origin=(8.0, 8.0), pixel_size=0.0055, main_lens_f=35.0, B=0.8, bl0=2.5)
    rec, n = lfpm.vd_point_cloud(torch.from_numpy(vd).cuda(), torch.from_numpy(col).cuda(), **prm)
    want = orc.vd_point_cloud(vd, col, prm["origin"], prm["pixel_size"], prm["main_lens_f"],
                              prm["B"], prm["bl0"])
    assert n == want.shape[0]
    assert np.array_equal(rec.cpu().numpy(), want)


@pytest.mark.gpu
def test_point_cloud_of_solved_frame_writes_ply(lfpm, orc, tmp_path):
    """The chain of the reference's VD product: solve, VD map, point cloud, PLY file."""
    import torch
    wl = lfpm.workload_c1()
    df = lfpm.DeviceFrame(wl)
    b = lfpm.patchmatch_frame(df)
    vd, col, _ = lfpm.virtual_depth_map(df, b.plane)
    prm = dict(origin=(wl.crop * 0.5, wl.crop * 0.5), pixel_size=0.0055, main_lens_f=35.0,
               B=0.8, bl0=2.5)
    rec, n = lfpm.vd_point_cloud(vd, col, **prm)
    torch.cuda.synchronize()
    want = orc.vd_point_cloud(vd.cpu().numpy(), col.cpu().numpy(), prm["origin"],
                              prm["pixel_size"], prm["main_lens_f"], prm["B"], prm["bl0"])
    _say("solved frame points", n)
    assert n > 0 and np.array_equal(rec.cpu().numpy(), want)
    path = tmp_path / "VD.ply"
    lfpm.write_ply_points(path, rec.cpu().numpy())
    data = path.read_bytes()
    assert data.endswith(want.tobytes()) and b"element vertex %d\n" % n in data
