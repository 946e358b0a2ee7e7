"""Pins of the oracle's refocus-side composites: the all-in-focus composite with hole filling and
denoising (LFRefocus::AIFImageCompositeForMIA, FocalStack/LFRefocus.cpp:183-425, 2283-2498) and
the back-projection consistency fusion (LFRefocus::FuseVirtualDepth_BackProject, :1553-1793).

Each check takes an oracle library handle, so the mutation test can run the same checks on
builds with one rule changed.  The checks are
  * closed forms: an identity and a 2x projection of one MI (positions, depth maps, colours), a
    half-way position (round half away from zero), the score order with its integer quotients
    and transposed reads, a three-MI row whose fused points are exact (vd = 1, integer pitch),
    two MIs (never two consistent neighbours);
  * hand-derived hole-filling rows (the fill attributed to the MI of the last collected
    candidate, a candidate on the frame's far edge included);
  * scipy's median filter for the denoising passes;
  * independent straightforward restatements below (plain Python over numpy float32 scalars,
    dictionaries and lists, written from the reference text) on small random frames, compared
    bit for bit.
"""
import math

import numpy as np

import oracle as O

F = np.float32


# ------------------------------------------------------------------------------------------
# independent restatements (small inputs only)

def _round_half_away(v):
    return int(math.copysign(math.floor(abs(float(v)) + 0.5), float(v)))


def py_aif(plane, centers, bgr, rich, blur, garbage, W, baseline, B, bl0, f, cols, rows):
    """AIFImageCompositeForMIA steps 1-4 as the reference text reads (dict of lists)."""
    V = len(centers)
    co = {}
    last = None
    for v in range(V):
        if garbage is not None and garbage[v]:
            continue
        for y in range(W):
            for x in range(W):
                d = F(plane[v, y, x, 3])
                if not d > 0:
                    continue
                vd = F(F(baseline) / d)
                cen = F(F(W - 1) * F(0.5))
                X = _round_half_away(F(F(F(x) - cen) * vd) + F(centers[v][0]))
                Y = _round_half_away(F(F(F(y) - cen) * vd) + F(centers[v][1]))
                if 0 <= X <= cols and 0 <= Y <= rows:
                    co.setdefault((Y, X), []).append((v, x, y, vd))
                    last = (v, x, y, vd)
    chosen = {}
    for (Y, X), lst in co.items():
        if Y >= rows or X >= cols:
            continue
        best = None
        for (v, x, y, vd) in lst:
            s = (int(rich[v, x, y]) // 255) if rich is not None else 0
            b = (int(blur[v, x, y]) // 255) if blur is not None else 0
            sc = 2 * s + b
            if best is None or sc > best[0]:
                best = (sc, v, x, y, vd)
        chosen[(Y, X)] = list(best[1:])
    vdm = np.zeros((rows, cols), np.float32)
    inv = np.zeros((rows, cols), np.float32)
    real = np.zeros((rows, cols), np.float32)
    img = np.zeros((rows, cols, 3), np.uint8)
    for (Y, X), (v, x, y, vd) in chosen.items():
        t = F(F(vd * F(B)) + F(bl0))
        real[Y, X] = F(F(F(f) * t) / F(t - F(f)))
        inv[Y, X] = F(F(1) / vd)
        vdm[Y, X] = vd
        img[Y, X] = bgr[v, y, x]
    aif = img.copy()
    kstar = last[0] if last is not None else -1
    pm = {k: (v, x, y) for k, (v, x, y, _) in chosen.items()}

    def hole(y, x):
        return not img[y, x].any()

    count = int(np.all(img >= 1, axis=2).sum())
    counter = 0
    passes = 0
    while True:
        passes += 1
        filled = 0
        for y in range(rows):
            for x in range(cols):
                if not hole(y, x):
                    continue
                if x + 1 < cols and not hole(y, x + 1):
                    q, dx, dy = (y, x + 1), -1, 0
                elif x - 1 >= 0 and not hole(y, x - 1):
                    q, dx, dy = (y, x - 1), 1, 0
                elif y + 1 < rows and not hole(y + 1, x):
                    q, dx, dy = (y + 1, x), 0, -1
                elif y - 1 >= 0 and not hole(y - 1, x):
                    q, dx, dy = (y - 1, x), 0, 1
                else:
                    continue
                if q not in pm:
                    continue
                v, sx, sy = pm[q]
                nx, ny = sx + dx, sy + dy
                if not (0 <= nx < W and 0 <= ny < W):
                    continue
                img[y, x] = bgr[v, ny, nx]
                pm[(y, x)] = (kstar, nx, ny)
                counter += 1
                filled += 1
                count += 1
        ratio = F(F(count) / F(rows * cols))
        if ratio >= F(0.98) or filled == 0 or counter > 30000:
            break
    return dict(vd=vdm, inv_vd=inv, real_mm=real, aif=aif, filled=img, passes=passes,
                fills=counter, kstar=kstar)


def py_fuse(plane, centers, bgr, nbr, cnt, garbage, W, baseline, B, bl0, cols, rows):
    """FuseVirtualDepth_BackProject step 2 as the reference text reads."""
    V = len(centers)
    mla = F(F(F(baseline) * F(bl0)) / F(F(bl0) + F(B)))
    cmi = F(F(F(W - 1) * F(0.5)) + F(1))
    masks = np.zeros((V, W, W), np.uint8)

    def vdof(d):
        if float(d) <= 0:
            return F(-9999.0)
        return F(float(mla) / float(d))

    def bad(v):
        return abs(F(v - F(-9999.0))) < F(1e-6) or not np.isfinite(v)

    def v3d(vd, c, r, v):
        return [F(F(F(F(c) - cmi) * vd) + F(centers[v][0])),
                F(F(F(F(r) - cmi) * vd) + F(centers[v][1])), vd]

    def back(P, v):
        return (F(F(F(P[0] - F(centers[v][0])) / P[2]) + cmi),
                F(F(F(P[1] - F(centers[v][1])) / P[2]) + cmi))

    vd_out = np.zeros((rows, cols), np.float32)
    aif = np.zeros((rows, cols, 3), np.uint8)
    conf = np.zeros((rows, cols), np.uint8)
    pts = []
    nfused = 0
    for a in range(V):
        if garbage is not None and garbage[a]:
            continue
        for r in range(W):
            for c in range(W):
                if masks[a, r, c] == 1:
                    continue
                d = F(plane[a, r, c, 3])
                if d < F(1):
                    continue
                vr = vdof(d)
                if bad(vr):
                    continue
                P = v3d(vr, c, r, a)
                Pr = list(P)
                nrm = [F(plane[a, r, c, k]) for k in range(3)]
                col = [F(bgr[a, r, c, k]) for k in range(3)]
                used = []
                for j in range(cnt[a]):
                    b = int(nbr[a][j])
                    if b < 0 or (garbage is not None and garbage[b]):
                        continue
                    q = back(Pr, b)
                    sr, sc = int(F(q[1] + F(0.5))), int(F(q[0] + F(0.5)))
                    if sc < 0 or sc >= W or sr < 0 or sr >= W:
                        continue
                    if masks[b, sr, sc] == 1:
                        continue
                    ds = F(plane[b, sr, sc, 3])
                    if ds <= F(0):
                        continue
                    vs = vdof(ds)
                    if bad(vs):
                        continue
                    Ps = v3d(vs, sc, sr, b)
                    qr = back(Ps, a)
                    e0, e1 = F(F(c) - qr[0]), F(F(r) - qr[1])
                    err = F(math.sqrt(float(F(F(e0 * e0) + F(e1 * e1)))))
                    rel = F(abs(F(vs - vr)) / max(vr, F(1e-6)))
                    if err < F(1) and rel < F(0.05):
                        P = [F(P[k] + Ps[k]) for k in range(3)]
                        nrm = [F(nrm[k] + F(plane[b, sr, sc, k])) for k in range(3)]
                        col = [F(col[k] + F(bgr[b, sr, sc, k])) for k in range(3)]
                        used.append((b, sr, sc))
                if len(used) >= 2:
                    ial = F(F(1) / F(F(len(used)) + F(1)))
                    P = [F(x * ial) for x in P]
                    nrm = [F(x * ial) for x in nrm]
                    col = [F(x * ial) for x in col]
                    nfused += 1
                    vx, vy = int(np.rint(P[0])), int(np.rint(P[1]))
                    if 0 <= vy < rows and 0 <= vx < cols:
                        vd_out[vy, vx] = P[2]
                        conf[vy, vx] = min(255, (len(used) + 1) * 60)
                        aif[vy, vx] = [min(max(int(np.rint(x)), 0), 255) for x in col]
                        pts.append(P + nrm + col)
                    for (b, sr, sc) in used:
                        masks[b, sr, sc] = 1
    return dict(vd=vd_out, aif=aif, conf=conf, points=np.array(pts, np.float32).reshape(-1, 9),
                nfused=nfused)


# ------------------------------------------------------------------------------------------
# small synthetic inputs

def _hex_layout(cols_mi, rows_mi, pitch, W, x0, y0):
    """Centres of a hex grid (odd rows shifted by half a pitch) and its ring-1 neighbour table."""
    cen, idx = [], {}
    for r in range(rows_mi):
        for c in range(cols_mi):
            idx[(r, c)] = len(cen)
            cen.append((x0 + c * pitch + (pitch / 2 if r % 2 else 0.0),
                        y0 + r * pitch * math.sqrt(3) / 2))
    V = len(cen)
    nbr = -np.ones((V, 6), np.int32)
    cnt = np.zeros(V, np.int32)
    for (r, c), v in idx.items():
        odd = r % 2
        cand = [(r, c - 1), (r, c + 1), (r - 1, c - 1 + odd), (r - 1, c + odd),
                (r + 1, c - 1 + odd), (r + 1, c + odd)]
        for q in cand:
            if q in idx:
                nbr[v, cnt[v]] = idx[q]
                cnt[v] += 1
    return np.array(cen, np.float32), nbr, cnt


def random_frame(seed, cols_mi=4, rows_mi=3, pitch=6.0, W=7, noise=0.03, invalid=0.1):
    """A small hex frame whose disparities follow one tilted plane of virtual depth (so
    neighbours agree), with noise and invalid pixels; random colours and normals."""
    rng = np.random.default_rng(seed)
    cen, nbr, cnt = _hex_layout(cols_mi, rows_mi, pitch, W, 6.0, 6.0)
    V = len(cen)
    cols = int(math.ceil(cen[:, 0].max() + 8))
    rows = int(math.ceil(cen[:, 1].max() + 8))
    plane = np.zeros((V, W, W, 4), np.float32)
    for v in range(V):
        for y in range(W):
            for x in range(W):
                px = cen[v, 0] + x - W // 2
                vd = 2.0 + 0.02 * px
                d = pitch / vd * (1.0 + noise * rng.standard_normal())
                if rng.random() < invalid:
                    d = -1.0 if rng.random() < 0.5 else 0.5
                n = rng.standard_normal(3)
                n /= np.linalg.norm(n)
                plane[v, y, x] = (n[0], n[1], n[2], d)
    bgr = rng.integers(1, 256, (V, W, W, 3)).astype(np.uint8)
    return dict(plane=plane, centers=cen, nbr=nbr, cnt=cnt, bgr=bgr, W=W, cols=cols, rows=rows,
                pitch=pitch, V=V, rng=rng)


def _plane_const(V, W, d):
    p = np.zeros((V, W, W, 4), np.float32)
    p[..., 2] = 1.0
    p[..., 3] = d
    return p


def _colours(V, W, seed):
    return np.random.default_rng(seed).integers(1, 256, (V, W, W, 3)).astype(np.uint8)


def _aif(L, plane, centers, bgr, W, cols, rows, baseline=1.0, B=0.8, bl0=40.0, f=25.0, **kw):
    return O.aif_composite(plane.reshape(-1, 4), centers, bgr, W, baseline, B, bl0, f, cols,
                           rows, L=L, **kw)


# ------------------------------------------------------------------------------------------
# AIF checks

def check_aif_identity(L):
    """One MI, vd = 1: pixel (x, y) lands at (x, y) - (W - 1) / 2 + centre; depth maps 1, 1 and
    the thin-lens real depth f t / (t - f), t = B + bl0."""
    W, cols, rows = 5, 16, 12
    cen = np.array([[7.0, 6.0]], np.float32)
    bgr = _colours(1, W, 3)
    r = _aif(L, _plane_const(1, W, 1.0), cen, bgr, W, cols, rows)
    exp = np.zeros((rows, cols, 3), np.uint8)
    exp[4:9, 5:10] = bgr[0]
    assert np.array_equal(r["aif"], exp)
    m = exp.any(axis=2)
    assert np.all(r["vd"][m] == 1.0) and np.all(r["inv_vd"][m] == 1.0)
    t = 0.8 + 40.0
    assert np.allclose(r["real_mm"][m], 25.0 * t / (t - 25.0), rtol=1e-6)
    assert not r["vd"][~m].any() and not r["real_mm"][~m].any()
    assert r["stats"][0] == 25 and r["stats"][1] == 25


def check_aif_magnify(L):
    """vd = 2 (d = baseline / 2): (x, y) lands at 2 (x, y) - 4 + centre."""
    W, cols, rows = 5, 20, 20
    cen = np.array([[9.0, 8.0]], np.float32)
    bgr = _colours(1, W, 4)
    r = _aif(L, _plane_const(1, W, 0.5), cen, bgr, W, cols, rows)
    exp = np.zeros((rows, cols, 3), np.uint8)
    for y in range(W):
        for x in range(W):
            exp[2 * y - 4 + 8, 2 * x - 4 + 9] = bgr[0, y, x]
    assert np.array_equal(r["aif"], exp)
    assert np.all(r["vd"][exp.any(axis=2)] == 2.0)


def check_aif_half_away(L):
    """A centre at x.5 puts every column half-way: round half away from zero (std::round)."""
    W, cols, rows = 5, 16, 12
    cen = np.array([[7.5, 6.0]], np.float32)
    bgr = _colours(1, W, 5)
    r = _aif(L, _plane_const(1, W, 1.0), cen, bgr, W, cols, rows)
    exp = np.zeros((rows, cols, 3), np.uint8)
    exp[4:9, 6:11] = bgr[0]  # x - 2 + 7.5 -> x + 6
    assert np.array_equal(r["aif"], exp)


def check_aif_score_order(L):
    """Two MIs on one centre: every position has two candidates, MI 0 first.  MI 1 wins where its
    score class (rich 255 -> 0.9, blur 255 -> 0.1) beats MI 0's; 254 counts as 0 (integer
    quotients); the maps are read at [source column][source row]."""
    W, cols, rows = 5, 16, 12
    cen = np.array([[7.0, 6.0], [7.0, 6.0]], np.float32)
    bgr = _colours(2, W, 6)
    rich = np.zeros((2, W, W), np.uint8)
    blur = np.zeros((2, W, W), np.uint8)
    # source pixel (x, y) of MI 1 is scored at rich[1][x][y]
    rich[1, 1, 3] = 255      # MI 1 pixel (x=1, y=3): rich -> wins
    blur[1, 4, 0] = 255      # MI 1 pixel (x=4, y=0): blur -> wins
    rich[1, 2, 2] = 254      # not saturated: tie, MI 0 stays
    rich[0, 0, 4] = 255      # MI 0 pixel (0, 4) rich ...
    rich[1, 0, 4] = 255      # ... MI 1 too, and blurred: wins
    blur[1, 0, 4] = 255
    rich[0, 3, 1] = 255      # MI 0 rich, MI 1 only blurred: MI 0 stays
    blur[1, 3, 1] = 255
    r = _aif(L, _plane_const(2, W, 1.0), cen, bgr, W, cols, rows, rich=rich, blur=blur)
    win1 = {(1, 3), (4, 0), (0, 4)}
    for y in range(W):
        for x in range(W):
            src = bgr[1, y, x] if (x, y) in win1 else bgr[0, y, x]
            assert np.array_equal(r["aif"][y + 4, x + 5], src), (x, y)
    assert r["stats"][4] == 2


def check_hole_fill_attribution(L):
    """One-row frame, two MIs (W = 5, vd = 1).  MI 0 covers columns 0..4 with its row 2, columns
    3 and 4 invalid; MI 1 (centre column 9) has only its row 3 valid, which lands on row 1 = the
    frame height: collected, never kept, but the last candidate collected.  Filling (raster
    order): column 3 takes MI 0's pixel (3, 2) from its left neighbour and is attributed to MI 1;
    column 4 then steps right of that source inside MI 1 and takes MI 1's pixel (4, 2); column 5
    would step out of the crop, the rest have no non-hole neighbour."""
    W, cols, rows = 5, 12, 1
    cen = np.array([[2.0, 0.0], [9.0, 0.0]], np.float32)
    plane = _plane_const(2, W, 1.0)
    plane[0, :, 3:, 3] = -1.0
    plane[1, :, :, 3] = -1.0
    plane[1, 3, :, 3] = 1.0
    bgr = _colours(2, W, 7)
    r = _aif(L, plane, cen, bgr, W, cols, rows)
    exp = np.zeros((1, cols, 3), np.uint8)
    exp[0, 0:3] = bgr[0, 2, 0:3]
    assert np.array_equal(r["aif"], exp)
    exp[0, 3] = bgr[0, 2, 3]
    exp[0, 4] = bgr[1, 2, 4]
    assert np.array_equal(r["filled"], exp)
    assert r["stats"][5] == 1 and r["stats"][3] == 2 and r["stats"][2] == 2


def check_hole_fill_preference(L):
    """One MI (W = 5, vd = 1) covering a 5x5 frame, rows 0-1 and the first column of rows 2-3
    invalid.  Every fill steps its source one pixel away from the neighbour it came from, so the
    holes fill with the crop's own pixels: pass 1 fills row 1 from below ((1, 1)) and then from
    the left, and (0, 2), (0, 3) from the right; pass 2 fills row 0 from below / left and (0, 1)
    from the right; pass 3 fills (0, 0) and reaches a full frame.  3 passes, 12 fills."""
    W, cols, rows = 5, 5, 5
    cen = np.array([[2.0, 2.0]], np.float32)
    plane = _plane_const(1, W, 1.0)
    plane[0, 0:2, :, 3] = -1.0
    plane[0, 2:4, 0, 3] = -1.0
    bgr = _colours(1, W, 8)
    r = _aif(L, plane, cen, bgr, W, cols, rows)
    assert np.array_equal(r["filled"], bgr[0])
    assert r["stats"][2] == 3 and r["stats"][3] == 12 and r["stats"][7] == 13


def check_aif_vs_restatement(L):
    """Random small frames, with richness / blur maps and a garbage MI: every output bit for
    bit against the independent restatement above."""
    for seed in range(3):
        fr = random_frame(seed)
        V, W = fr["V"], fr["W"]
        rng = fr["rng"]
        rich = np.where(rng.random((V, W, W)) < 0.3, 255, 0).astype(np.uint8)
        rich[rng.random((V, W, W)) < 0.05] = 254
        blur = rng.integers(0, 256, (V, W, W)).astype(np.uint8)
        blur[rng.random((V, W, W)) < 0.3] = 255
        garbage = np.zeros(V, np.uint8)
        garbage[seed % V] = 1
        args = (fr["plane"], fr["centers"], fr["bgr"], rich, blur, garbage, W, fr["pitch"], 0.8,
                40.0, 25.0, fr["cols"], fr["rows"])
        ref = py_aif(*args)
        got = _aif(L, fr["plane"], fr["centers"], fr["bgr"], W, fr["cols"], fr["rows"],
                   baseline=fr["pitch"], rich=rich, blur=blur, garbage=garbage)
        for k in ("vd", "inv_vd", "real_mm", "aif", "filled"):
            assert np.array_equal(got[k].view(np.uint8), ref[k].view(np.uint8)), (seed, k)
        assert got["stats"][2] == ref["passes"] and got["stats"][3] == ref["fills"]
        assert got["stats"][5] == ref["kstar"]


def check_denoise_scipy(L):
    """Denoising = five 5x5 median passes with replicated borders, per channel (scipy)."""
    from scipy.ndimage import median_filter
    fr = random_frame(11)
    r = _aif(L, fr["plane"], fr["centers"], fr["bgr"], fr["W"], fr["cols"], fr["rows"],
             baseline=fr["pitch"])
    exp = r["aif"].copy()
    for _ in range(5):
        exp = np.stack([median_filter(exp[..., c], size=5, mode="nearest") for c in range(3)],
                       axis=2)
    assert np.array_equal(r["denoised"], exp)


# ------------------------------------------------------------------------------------------
# back-projection fusion checks

def _fuse(L, plane, centers, bgr, nbr, cnt, W, cols, rows, baseline=2.0, B=1.0, bl0=1.0, **kw):
    return O.fuse_backproject(plane.reshape(-1, 4), centers, bgr, nbr, cnt, W, baseline, B, bl0,
                              cols, rows, L=L, **kw)


def _row_layout(n, pitch, W):
    cen = np.array([[10.0 + pitch * i, 10.0] for i in range(n)], np.float32)
    nbr = -np.ones((n, 2), np.int32)
    cnt = np.zeros(n, np.int32)
    for i in range(n):
        for j in (i - 1, i + 1):
            if 0 <= j < n:
                nbr[i, cnt[i]] = j
                cnt[i] += 1
    return cen, nbr, cnt


def check_fuse_two_mis(L):
    """Two MIs: no pixel has two consistent neighbours, nothing is fused."""
    W = 9
    cen, nbr, cnt = _row_layout(2, 2, W)
    r = _fuse(L, _plane_const(2, W, 1.0), cen, _colours(2, W, 9), nbr, cnt, W, 32, 24)
    assert r["nfused"] == 0 and not r["vd"].any() and not r["conf"].any()


def check_fuse_row_closed_form(L):
    """Four MIs in a row, pitch 2, W = 9, vd = 1 (mla_base = 2 * 1 / (1 + 1) = 1 = d): pixel c of
    MI i back-projects to pixel c + 2 of MI i - 1 and c - 2 of MI i + 1, exactly, and those
    re-project onto it.  MI 1 fuses columns 2..6 of every row (its neighbours' pixels in range),
    masking MI 2's columns 0..4; MI 2 then fuses columns 5 and 6 only; MIs 0 and 3 have one
    neighbour.  Fused point = the pixel's own (exact), colour = round of the three-colour mean,
    confidence 180."""
    W, n = 9, 4
    cen, nbr, cnt = _row_layout(n, 2, W)
    bgr = _colours(n, W, 10)
    cols, rows = 32, 24
    r = _fuse(L, _plane_const(n, W, 1.0), cen, bgr, nbr, cnt, W, cols, rows)
    exp_vd = np.zeros((rows, cols), np.float32)
    exp_aif = np.zeros((rows, cols, 3), np.uint8)
    fused = [(1, c) for c in range(2, 7)] + [(2, c) for c in (5, 6)]
    for (i, c) in fused:
        for row in range(W):
            X, Y = int(cen[i, 0]) + c - 5, int(cen[i, 1]) + row - 5
            s = bgr[i, row, c].astype(int) + bgr[i - 1, row, c + 2] + bgr[i + 1, row, c - 2]
            exp_vd[Y, X] = 1.0
            exp_aif[Y, X] = [int(round(v / 3.0)) for v in s]
    assert r["nfused"] == len(fused) * W
    assert np.array_equal(r["vd"], exp_vd)
    assert np.array_equal(r["aif"], exp_aif)
    assert np.all(r["conf"][exp_vd > 0] == 180) and not r["conf"][exp_vd == 0].any()


def check_fuse_reference_disparity_floor(L):
    """The row of the closed form with d = 0.5 and mla_base = 0.5 (the same geometry, vd = 1):
    reference pixels need d >= 1, so nothing is fused, although every source pixel (d > 0) would
    be consistent."""
    W, n = 9, 3
    cen, nbr, cnt = _row_layout(n, 2, W)
    r = _fuse(L, _plane_const(n, W, 0.5), cen, _colours(n, W, 12), nbr, cnt, W, 32, 24,
              baseline=1.0, B=1.0, bl0=1.0)
    assert r["nfused"] == 0 and not r["vd"].any()
    r = _fuse(L, _plane_const(n, W, 1.0), cen, _colours(n, W, 12), nbr, cnt, W, 32, 24,
              baseline=2.0, B=1.0, bl0=1.0)
    assert r["nfused"] == 5 * W


def check_fuse_vs_restatement(L):
    """Random small hex frames (noisy consistent disparities, invalid pixels, a garbage MI):
    maps, points and count bit for bit against the independent restatement above."""
    for seed in range(3):
        fr = random_frame(100 + seed)
        V, W = fr["V"], fr["W"]
        garbage = np.zeros(V, np.uint8)
        if seed:
            garbage[(3 * seed) % V] = 1
        # mla_base = baseline * bl0 / (bl0 + B) = pitch: the disparities' own baseline
        args = dict(baseline=2 * fr["pitch"], B=3.0, bl0=3.0)
        ref = py_fuse(fr["plane"], fr["centers"], fr["bgr"], fr["nbr"], fr["cnt"], garbage, W,
                      args["baseline"], args["B"], args["bl0"], fr["cols"], fr["rows"])
        got = _fuse(L, fr["plane"], fr["centers"], fr["bgr"], fr["nbr"], fr["cnt"], W,
                    fr["cols"], fr["rows"], garbage=garbage, **args)
        assert ref["nfused"] > 0, seed
        for k in ("vd", "aif", "conf", "points"):
            assert got[k].shape == ref[k].shape, (seed, k)
            assert np.array_equal(got[k].view(np.uint8), ref[k].view(np.uint8)), (seed, k)
        assert got["nfused"] == ref["nfused"]


def all_checks():
    return [(f.__name__, f) for f in (
        check_aif_identity, check_aif_magnify, check_aif_half_away, check_aif_score_order,
        check_hole_fill_attribution, check_hole_fill_preference, check_aif_vs_restatement,
        check_denoise_scipy, check_fuse_two_mis, check_fuse_row_closed_form,
        check_fuse_reference_disparity_floor, check_fuse_vs_restatement)]
