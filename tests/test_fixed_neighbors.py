"""FixedPosition neighbour tables (SelectNeighbors::FixedPositionAlgorithm, mode 1).

* The oracle's per-key offset lists equal the reference's literal CollectValidNeighborKey lists
  (tests/golden/fixed_position_offsets.json, transcribed from MINeighbor/SelectNeighbors.cpp:
  1062-1286) for a key in each place of the lattice, including the reference's top-row test,
  which sends every row-0 key (corners included) to the top-edge list.
* The product's table (table-driven C++) equals the oracle's (if/else chain) on layouts from
  one lens row up to 4080x3068.
* Properties from the definition: each listed neighbour is a kept MI at one of the MI's
  offsets, the list is sorted by centre distance, and the garbage flag is the first key
  row/column or an empty list.
"""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "fixed_position_offsets.json")))["places"]


def _offs(place):
    return [tuple(o) for o in FIX[place]["offsets"]]


# (X, Y) keys in a u x v = 12 x 9 lattice and the place whose list the reference uses there
U, V = 12, 9
PLACES = [((5, 4), "interior"), ((2, 1), "interior"), ((U - 3, V - 2), "interior"),
          ((0, 0), "top_left"), ((1, 0), "top"), ((6, 0), "top"),
          ((U - 2, 0), "top"), ((U - 1, 0), "top"),   # the always-true top-row test
          ((0, V - 1), "bottom_left"), ((1, V - 1), "bottom_1"),
          ((U - 1, V - 1), "bottom_right"), ((U - 2, V - 1), "bottom_right_2"),
          ((0, 4), "left"), ((1, 4), "left_1"), ((5, V - 1), "bottom"),
          ((U - 1, 4), "right"), ((U - 2, 4), "right_2")]


@pytest.mark.parametrize("key,place", PLACES)
def test_oracle_offsets_equal_reference_lists(orc, key, place):
    assert orc.fixed_offsets(key[0], key[1], U, V) == _offs(place)


def test_top_right_lists_unreachable(orc):
    tr, tr2 = _offs("top_right"), _offs("top_right_2")
    for Y in range(V):
        for X in range(U):
            got = orc.fixed_offsets(X, Y, U, V)
            assert got != tr and got != tr2, (X, Y)


LAYOUTS = [(23.0, 1024, 800, 16), (23.0, 4080, 3068, 16), (15.0, 640, 480, 8),
           (31.0, 1024, 800, 12), (23.0, 160, 60, 16), (23.0, 60, 200, 16)]


@pytest.mark.parametrize("cfg", LAYOUTS)
def test_fixed_table_product_matches_oracle(lfpm, orc, cfg):
    pitch, w, h, K = cfg
    nbr, cnt, gar = lfpm.fixed_neighbors(lfpm.layout_params(pitch, w, h, K))
    onbr, ocnt, ogar = orc.fixed_neighbors(orc.LayoutParams(pitch, 0, w, h, K, 1, (K, K, K)))
    assert np.array_equal(cnt, ocnt)
    assert np.array_equal(nbr, onbr)
    assert np.array_equal(gar, ogar)


@pytest.mark.parametrize("cfg", LAYOUTS[:4])
def test_fixed_table_properties(lfpm, orc, cfg):
    pitch, w, h, K = cfg
    lp = lfpm.layout_params(pitch, w, h, K)
    L = lfpm.layout(lp)
    nbr, cnt, gar = lfpm.fixed_neighbors(lp)
    keys, cen, l2m = L["tilekeys"], L["centers"], L["lattice_to_mi"]
    u, v = L["cols"], L["rows"]
    for m in range(L["V"]):
        X, Y = keys[m]
        offs = orc.fixed_offsets(X, Y, u, v)
        kept = [l2m[Y + dy, X + dx] for dx, dy in offs
                if 0 <= X + dx < u and 0 <= Y + dy < v and l2m[Y + dy, X + dx] >= 0]
        n = cnt[m]
        assert n == min(len(kept), K)
        got = nbr[m, :n].tolist()
        assert set(got) <= set(kept) and np.all(nbr[m, n:] == -1)
        d = np.hypot(*(cen[got].astype(np.float64) - cen[m]).T)
        assert np.all(np.diff(d) >= -1e-4)
        # the cap keeps the nearest ones
        if len(kept) > K:
            dk = np.sort(np.hypot(*(cen[kept].astype(np.float64) - cen[m]).T))
            assert d[-1] <= dk[K] + 1e-4
        assert gar[m] == int(len(kept) == 0 or X < 1 or Y < 1)
    # an interior lens of a large layout has all 12 mode-1 neighbours (capped at K)
    # (the last lattice columns/rows hold lenses cut by the frame edge, which are not kept)
    inner = [m for m in range(L["V"]) if 2 < keys[m][0] < u - 5 and 1 < keys[m][1] < v - 4]
    assert inner and all(cnt[m] == min(12, K) for m in inner)


def test_facade_select_neighbors_fixed_position(lfpm, tmp_path):
    """LFMVS::SelectNeighbors::SelectNeighborsForMIA(eSNT_FixedPosition) fills every problem's
    m_Res_Image_KeyVec and m_bGarbage with the C ABI's FixedPosition table (host only)."""
    import subprocess
    root = os.path.dirname(HERE)
    drv = os.path.join(root, "tests", "cpp", "_build", "select_neighbors_driver")
    out = tmp_path / "sn.bin"
    pitch, w, h, K = 23.0, 1024, 800, 8
    subprocess.check_call([drv, str(pitch), str(w), str(h), str(K), str(out)])
    a = np.fromfile(str(out), np.int32)
    V, K2 = a[0], a[1]
    assert K2 == K
    recs = a[2:2 + V * (4 + 2 * K)].reshape(V, 4 + 2 * K)
    assert a[-1] == 1  # the feature-based type is not provided and throws
    lp = lfpm.layout_params(pitch, w, h, K)
    L = lfpm.layout(lp)
    nbr, cnt, gar = lfpm.fixed_neighbors(lp)
    assert V == L["V"]
    assert np.array_equal(recs[:, :2], L["tilekeys"])
    assert np.array_equal(recs[:, 2], gar)
    assert np.array_equal(recs[:, 3], cnt)
    keys = recs[:, 4:].reshape(V, K, 2)
    want = np.where(nbr[..., None] >= 0, L["tilekeys"][np.maximum(nbr, 0)], -1)
    assert np.array_equal(keys, want)
