"""Calibration ingest: intrinsics.xml and CalibInfo.xml in the structure that
DepthSolver::ReadIntrinsicsFromXML / ReadCalibrationParamsFromXML read (DepthSolver.cpp:
872-1110, 1112-1251), and the FixedPosition neighbour table of a lens table read that way.

The reference ships no calibration files, so the documents below are written from the readers'
element walk (tag names, "_" wrappers, the order of values, sscanf formats); the checks pin the
walk's rules: the rotation's "_" groups are columns, the last value under mia/mesh/pitch wins,
keys are "row,col" read with %i (a leading 0 is octal), a blank before a comma ends a %f,%f
read, lenses come back in (row, col) order
with the last of a repeated key kept (the map assignment of DepthSolver.cpp:1242).  The table-driven neighbour builder must reproduce the
layout-driven one (lfpm_layout_fixed_neighbors) on the same lenses."""
import numpy as np
import pytest

INTRINSICS = """<?xml version="1.0" encoding="UTF-8"?>
<!-- written from the reader's element walk -->
<_>
  <D>12.5</D>
  <d>0.75</d>
  <sensor>
    <pose>
      <rotation>
        <_><_>1</_><_>2</_><_>3</_></_>
        <_><_>4</_><_>5</_><_>6</_></_>
        <_><_>7</_><_>8</_><_>9</_></_>
      </rotation>
      <translation><_><_>0.1</_><_>-0.2</_><_>0.3</_></_></translation>
    </pose>
    <scale>0.0055</scale>
  </sensor>
  <main_lens><f>35.0</f></main_lens>
  <mia>
    <mesh>
      <width>177</width>
      <height>154</height>
      <pitch><_><_>22.9</_><_>23.1</_></_></pitch>
    </mesh>
  </mia>
  <distortions>
    <depth><_><_>1e-3</_><_>2e-3</_><_>3e-3</_></_></depth>
    <radial><_><_>0.1</_><_>0.2</_><_>0.3</_></_></radial>
    <tangential><_><_>0.01</_><_>0.02</_></_></tangential>
  </distortions>
</_>
"""

CALIB = """<?xml version="1.0"?>
<Calib>
  <BlurInfo>
    <pose><r>1,0,0,1</r><t>0,0</t></pose>
    <MLAType id="0">2.5</MLAType>
    <MLAType id="2">4.25</MLAType>
  </BlurInfo>
  <MLAInfo>
    <MLA><key>1,2</key><center>100.5,200.25</center><MLAType>1</MLAType><aera>415.5</aera></MLA>
    <MLA><key>0,7</key><center>30,40</center><MLAType>2</MLAType><aera>400</aera></MLA>
    <MLA><key>010,3</key><center>5.5,6.5</center><MLAType>0</MLAType><aera>1e2</aera></MLA>
    <MLA><key>1,2</key><center>-1,-1</center><MLAType>0</MLAType><aera>0</aera></MLA>
    <MLA><key>0,1</key><center>7 , 8</center><MLAType>0</MLAType><aera>3</aera></MLA>
  </MLAInfo>
</Calib>
"""


def test_intrinsics_walk(lfpm, tmp_path):
    p = tmp_path / "intrinsics.xml"
    p.write_text(INTRINSICS)
    i = lfpm.read_intrinsics_xml(p)
    assert (i.D, i.d, i.main_lens_f) == (np.float32(12.5), np.float32(0.75), np.float32(35.0))
    assert i.sensor_pixel_size == np.float32(0.0055)
    assert (i.mla_u_size, i.mla_v_size) == (177, 154)
    assert i.baseline == np.float32(23.1)  # the last value under pitch
    # "_" groups are columns: column j holds R[0][j], R[1][j], R[2][j]
    assert list(i.sensor_rotation) == [1, 4, 7, 2, 5, 8, 3, 6, 9]
    assert list(i.sensor_translation) == [0.1, -0.2, 0.3]
    assert list(i.distortion_depth) == [1e-3, 2e-3, 3e-3]
    assert list(i.distortion_radial) == [0.1, 0.2, 0.3]
    assert list(i.distortion_tangential) == [0.01, 0.02]


def test_intrinsics_errors(lfpm, tmp_path):
    p = tmp_path / "bad.xml"
    p.write_text("<root><D>1</D></root>")
    with pytest.raises(Exception):
        lfpm.read_intrinsics_xml(p)
    p.write_text("<_><D>1</D>")  # unterminated
    with pytest.raises(Exception):
        lfpm.read_intrinsics_xml(p)
    with pytest.raises(Exception):
        lfpm.read_intrinsics_xml(tmp_path / "missing.xml")


def test_calib_table(lfpm, tmp_path):
    p = tmp_path / "CalibInfo.xml"
    p.write_text(CALIB)
    c = lfpm.read_calib_xml(p)
    # (row, col) order: (0,1), (0,7), (1,2) [last of the repeated key], (8,3) [octal 010]
    assert c["tilekeys"].tolist() == [[1, 0], [7, 0], [2, 1], [3, 8]]
    # "7 , 8": %f,%f stops at the blank before the comma, y keeps the initial 0
    assert c["centers"].tolist() == [[7, 0], [30, 40], [-1, -1], [5.5, 6.5]]
    assert c["lens_type"].tolist() == [0, 2, 0, 0]
    assert c["area"].tolist() == [3, 400, 0, 100]
    assert c["blur_radius"].tolist() == [2.5, 0.0, 4.25]


@pytest.mark.parametrize("pitch,w,h,K", [(23.0, 1024, 800, 18), (15.0, 640, 480, 12),
                                         (31.0, 900, 700, 8)])
def test_table_neighbours_equal_layout_neighbours(lfpm, pitch, w, h, K):
    lp = lfpm.layout_params(pitch, w, h, K)
    L = lfpm.layout(lp)
    want = lfpm.fixed_neighbors(lp)
    # the table in a shuffled order: ids are table indices, so map them back
    rng = np.random.default_rng(int(pitch))
    perm = rng.permutation(L["V"])
    inv = np.argsort(perm)
    nbr, cnt, gar = lfpm.table_fixed_neighbors(L["tilekeys"][perm], L["centers"][perm],
                                                L["cols"], L["rows"], K)
    back = np.where(nbr >= 0, perm[np.maximum(nbr, 0)], -1)
    assert np.array_equal(back[inv], want[0])
    assert np.array_equal(cnt[inv], want[1]) and np.array_equal(gar[inv], want[2])


def test_table_neighbours_reject_bad_keys(lfpm):
    k = np.array([[0, 0], [0, 0]], np.int32)
    c = np.zeros((2, 2), np.float32)
    with pytest.raises(Exception):
        lfpm.table_fixed_neighbors(k, c, 4, 4, 8)  # repeated key
    with pytest.raises(Exception):
        lfpm.table_fixed_neighbors(np.array([[4, 0]], np.int32), c[:1], 4, 4, 8)


def test_table_origins_corner_rule(lfpm):
    # crop 16: half = 7.5.  Lens area 10 x 4 lenses of pitch 23 -> 230 x 92 px; raw 240 x 100.
    cen = np.array([[7.5, 7.5],      # corner (0, 0): kept
                    [7.4, 50.0],     # left corner -0.1 < 0: abandoned
                    [222.5, 84.5],   # right / top corner exactly 230, 92: kept (not > limit)
                    [222.6, 50.0],   # right corner 230.1 > 230: abandoned
                    [100.9, 20.2]],  # corner (93.4, 12.7) truncated to (93, 12)
                   np.float32)
    o, k = lfpm.table_origins(cen, 16, 10, 4, 23.0, 240, 100)
    assert k.tolist() == [1, 0, 1, 0, 1]
    assert o.tolist() == [[0, 0], [-1, -1], [215, 77], [-1, -1], [93, 12]]
    # the same lens inside the lens area but past a smaller raw frame is dropped
    o, k = lfpm.table_origins(cen[2:3], 16, 10, 4, 23.0, 228, 100)
    assert k.tolist() == [0]


def test_table_origins_equal_layout_origins(lfpm):
    lp = lfpm.layout_params(23.0, 1024, 800, 18)
    L = lfpm.layout(lp)
    o, k = lfpm.table_origins(L["centers"], L["crop"], 10 ** 6, 10 ** 6, 1.0, 1024, 800)
    assert k.all() and np.array_equal(o, L["origins"])


@pytest.mark.gpu
def test_calib_file_drives_the_solver(lfpm, tmp_path):
    """A lens table written as CalibInfo.xml and read back (keys, centres, crop origins and the
    FixedPosition table from the reader's path) solves a frame exactly as the layout's own
    FixedPosition tables do."""
    import torch
    wl = lfpm.workload_c1()
    L = wl.L
    print("calib chain V", wl.V, flush=True)
    lines = ["<?xml version=\"1.0\"?>", "<Calib>", "<MLAInfo>"]
    for (X, Y), (cx, cy) in zip(L["tilekeys"], L["centers"]):
        lines.append("<MLA><key>%d,%d</key><center>%.9g,%.9g</center><MLAType>0</MLAType>"
                     "<aera>1</aera></MLA>" % (Y, X, cx, cy))
    lines += ["</MLAInfo>", "</Calib>"]
    path = tmp_path / "CalibInfo.xml"
    path.write_text("\n".join(lines))
    c = lfpm.read_calib_xml(path)
    assert np.array_equal(c["tilekeys"], L["tilekeys"]) and np.array_equal(c["centers"], L["centers"])
    h, w = wl.raw().shape
    o, k = lfpm.table_origins(c["centers"], L["crop"], L["cols"], L["rows"], 1e6, w, h)
    assert k.all() and np.array_equal(o, L["origins"])
    K = wl.pm.max_neighbors
    nbr_t, cnt_t, _ = lfpm.table_fixed_neighbors(c["tilekeys"], c["centers"], L["cols"],
                                                  L["rows"], K)
    nbr_l, cnt_l, _ = lfpm.fixed_neighbors(wl.lp)
    outs = []
    for nbr, cnt in ((nbr_t, cnt_t), (nbr_l, cnt_l)):
        df = lfpm.DeviceFrame(wl)
        df.nbr = torch.from_numpy(np.ascontiguousarray(nbr, np.int32)).cuda()
        df.cnt = torch.from_numpy(np.ascontiguousarray(cnt, np.int32)).cuda()
        b = lfpm.patchmatch_frame(df)
        torch.cuda.synchronize()
        outs.append((b.plane.cpu().numpy(), b.cost.cpu().numpy()))
    assert np.array_equal(outs[0][0].view(np.uint32), outs[1][0].view(np.uint32))
    assert np.array_equal(outs[0][1].view(np.uint32), outs[1][1].view(np.uint32))


@pytest.mark.parametrize("pitch,w,h,K,rings", [(23.0, 1024, 800, 18, 2), (15.0, 640, 480, 6, 1),
                                               (31.0, 900, 700, 24, 3)])
def test_table_ring_neighbours_equal_brute_force(lfpm, orc, pitch, w, h, K, rings):
    """lfpm_table_ring_neighbors (ring walk over a key -> lens map) equals the oracle's brute
    force over all lens pairs, on the layout's centres and on calibrated-like centres (jittered by
    up to 0.5 px, lenses in shuffled order); interior lenses get the 6 / 18 / 36 lenses of rings
    1 / 2 / 3 (capped at K)."""
    L = lfpm.layout(lfpm.layout_params(pitch, w, h, 18))
    keys, cen = L["tilekeys"], L["centers"]
    rng = np.random.default_rng(int(pitch))
    for jitter in (False, True):
        c = cen.copy()
        k = keys.copy()
        if jitter:
            c = (c + rng.uniform(-0.5, 0.5, c.shape)).astype(np.float32)
            perm = rng.permutation(len(k))
            k, c = k[perm], c[perm]
        nb, ct = lfpm.table_ring_neighbors(k, c, L["cols"], L["rows"], rings, K)
        onb, oct_ = orc.ring_neighbors(k, c, rings, K)
        assert np.array_equal(nb, onb) and np.array_equal(ct, oct_)
        ring_size = 3 * rings * (rings + 1)
        assert ct.max() == min(K, ring_size)
        # nearest first: distances never decrease along a row
        V = len(k)
        for v in range(0, V, max(1, V // 50)):
            d = np.hypot(*(c[nb[v, :ct[v]]] - c[v]).T)
            assert np.all(np.diff(d) >= -1e-4)


@pytest.mark.gpu
def test_calib_file_ring_table_solve_vs_fp64(lfpm, orc, tmp_path):
    """A frame driven by its CalibInfo.xml (keys and centres read back, crop origins and the
    ring-walk matching table of rings 1-2 built from the read table) solved on the GPU: within
    0.05 px of the FP64 oracle solving the same tables on >= 99.5 % of valid pixels."""
    import torch
    from _helpers import pm_to_oracle
    wl = lfpm.workload_c2(raw_w=640, raw_h=512)
    L = wl.L
    print("calib ring chain V", wl.V, flush=True)
    lines = ["<?xml version=\"1.0\"?>", "<Calib>", "<MLAInfo>"]
    for (X, Y), (cx, cy) in zip(L["tilekeys"], L["centers"]):
        lines.append("<MLA><key>%d,%d</key><center>%.9g,%.9g</center><MLAType>0</MLAType>"
                     "<aera>1</aera></MLA>" % (Y, X, cx, cy))
    lines += ["</MLAInfo>", "</Calib>"]
    path = tmp_path / "CalibInfo.xml"
    path.write_text("\n".join(lines))
    c = lfpm.read_calib_xml(path)
    h, w = wl.raw().shape
    o, k = lfpm.table_origins(c["centers"], L["crop"], L["cols"], L["rows"], 1e6, w, h)
    assert k.all()
    K = wl.pm.max_neighbors
    nbr, cnt = lfpm.table_ring_neighbors(c["tilekeys"], c["centers"], L["cols"], L["rows"], 2, K)
    assert cnt.max() == 18
    df = lfpm.DeviceFrame(wl)
    df.nbr = torch.from_numpy(nbr).cuda()
    df.cnt = torch.from_numpy(cnt).cuda()
    b = lfpm.patchmatch_frame(df)
    torch.cuda.synchronize()
    d = b.plane.cpu().numpy()[:, 3].astype(np.float64)
    tiles = orc.slice_u8(wl.raw(), o, L["crop"])
    ores = orc.solve(tiles, c["centers"], nbr, cnt, pm_to_oracle(orc, wl.pm))
    valid = ores["cost"] <= 1.8
    frac = float((np.abs(d - ores["plane"][:, 3])[valid] <= 0.05).mean())
    print("calib ring solve frac", frac, flush=True)
    assert frac >= 0.995
