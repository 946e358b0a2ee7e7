"""Pins of the oracle's all-in-focus composite and back-projection fusion (tests/_fusion_pins.py:
closed forms, hand-derived hole-filling rows, scipy's median filter, independent restatements on
small random frames), a mutation check (builds of the oracle with one rule of the composites
changed each fail at least one check), and the product's level schedule for the fusion walk
(lfpm_fuse_schedule, host code) against a brute-force conflict test."""
import concurrent.futures as cf
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O
from _fusion_pins import all_checks, random_frame
from _helpers import build_oracle_mutant


@pytest.mark.parametrize("name", [n for n, _ in all_checks()])
def test_fusion_piece(name):
    dict(all_checks())[name](O.lib())


# (label, [(old, new), ...]) against oracle/lfpm_oracle.cpp
MUTATIONS = [
    ("aif centre W / 2", [("const float centerX = (W - 1) * 0.5f, centerY = (W - 1) * 0.5f;",
                           "const float centerX = W * 0.5f, centerY = W * 0.5f;")]),
    ("aif round half even", [("const int X = (int)std::round(offX * vdp + cx);",
                              "const int X = (int)std::nearbyint(offX * vdp + cx);")]),
    ("aif bound X < cols", [("if (X <= cols && X >= 0 && Y <= rows && Y >= 0) {",
                             "if (X < cols && X >= 0 && Y < rows && Y >= 0) {")]),
    ("aif score not transposed", [("const size_t t = (size_t)q.mi * WW + (size_t)q.sx * W + q.sy;",
                                   "const size_t t = (size_t)q.mi * WW + (size_t)q.sy * W + q.sx;")]),
    ("aif score float quotient", [("(float)((rich ? rich[t] : 0) / 255);",
                                   "(float)((rich ? rich[t] : 0) / 255.0f);")]),
    ("aif score weights swapped", [("const float a = 0.90f, b = 0.10f;",
                                    "const float a = 0.10f, b = 0.90f;")]),
    ("aif colour transposed", [("const uint8_t* s = bgr + ((size_t)c.mi * WW + (size_t)c.sy * W + c.sx) * 3;",
                                "const uint8_t* s = bgr + ((size_t)c.mi * WW + (size_t)c.sx * W + c.sy) * 3;")]),
    ("aif real depth sign", [("const float real_d = (f_main * tmp) / (tmp - f_main);",
                              "const float real_d = (f_main * tmp) / (tmp + f_main);")]),
    ("fill right dropped", [("if (x + 1 < cols && !hole(p + 1)) {", "if (false) {")]),
    ("fill down before left", [("} else if (x - 1 >= 0 && !hole(p - 1)) {\n                    q = p - 1;",
                                "} else if (x - 1 >= 0 && !hole(p - 1) && !(y + 1 < rows && !hole(p + cols))) {\n                    q = p - 1;")]),
    ("fill keeps the neighbour's MI", [("ch[p] = Chosen{kstar, nsx, nsy, ch[p].vd, true};",
                                        "ch[p] = Chosen{ch[q].mi, nsx, nsy, ch[p].vd, true};")]),
    ("fill up steps the wrong way", [("nsx = ch[q].sx; nsy = ch[q].sy + 1;",
                                      "nsx = ch[q].sx; nsy = ch[q].sy - 1;")]),
    ("fill ratio 0.9", [("const float targetFillRatio = 0.98f;", "const float targetFillRatio = 0.9f;")]),
    ("denoise four passes", [("median5_passes(dn, rows, cols, 3, 5);",
                              "median5_passes(dn, rows, cols, 3, 4);")]),
    ("fuse one neighbour enough", [("if (num >= 2) {", "if (num >= 1) {")]),
    ("fuse reprojection 0.5 px", [("if (err < 1.0f && rel < 0.05f) {",
                                   "if (err < 0.5f && rel < 0.05f) {")]),
    ("fuse depth 10 %", [("if (err < 1.0f && rel < 0.05f) {",
                          "if (err < 1.0f && rel < 0.10f) {")]),
    ("fuse own mask ignored", [("if (masks[(size_t)a * WW + r * W + c] == 1) continue;", "")]),
    ("fuse source mask ignored", [("if (masks[(size_t)b * WW + sr * W + sc] == 1) continue;", "")]),
    ("fuse image baseline", [("const float mla_base = baseline * bl0 / (bl0 + B);",
                              "const float mla_base = baseline;")]),
    ("fuse centre without +1", [("const float cmi = (W - 1) * 0.5f + 1.0f;",
                                 "const float cmi = (W - 1) * 0.5f;")]),
    ("fuse reference d > 0", [("if (disp_ref < 1.0f) continue;", "if (disp_ref <= 0.0f) continue;")]),
    ("fuse confidence n * 60", [("conf_out[o] = (uint8_t)std::min(255, (num + 1) * 60);",
                                 "conf_out[o] = (uint8_t)std::min(255, num * 60);")]),
    ("fuse source rounding floor", [("const int sr = static_cast<int>(q[1] + 0.5f);",
                                     "const int sr = static_cast<int>(std::floor(q[1] + 0.5f));")]),
]


def test_fusion_mutations_are_caught(tmp_path):
    checks = all_checks()
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        sos = list(ex.map(lambda a: build_oracle_mutant(a[0], a[1][1], str(tmp_path)),
                          enumerate(MUTATIONS)))
    survivors = []
    for (label, _), so in zip(MUTATIONS, sos):
        L = C.CDLL(so)
        failed = False
        for _, fn in checks:
            try:
                fn(L)
            except AssertionError:
                failed = True
                break
        if not failed:
            survivors.append(label)
    assert not survivors, "mutations no check rejects: %s" % survivors


def _conflicts(nbr, cnt, garbage, a, b):
    def N(x):
        return {int(m) for m in nbr[x][:cnt[x]] if m >= 0 and not garbage[m]}
    Na, Nb = N(a), N(b)
    return bool(Na & (Nb | {b})) or bool(Nb & (Na | {a}))


@pytest.mark.parametrize("case", ["hex", "hex_garbage", "c1_layout"])
def test_fuse_schedule_levels(lfpm, case):
    """Every pair of MIs whose walks touch a common mask runs in MI order (the later one in a
    later level); MIs of one level touch no common mask; every non-garbage MI appears once."""
    if case == "c1_layout":
        wl = lfpm.workload_c1(raw_w=320, raw_h=256)
        nbr, cnt = wl.L["nbr"].reshape(wl.V, -1), wl.L["cnt"]
    else:
        fr = random_frame(5, cols_mi=7, rows_mi=6)
        nbr, cnt = fr["nbr"], fr["cnt"]
    V = len(cnt)
    garbage = np.zeros(V, np.uint8)
    if case == "hex_garbage":
        garbage[[3, 10, 11]] = 1
    order, start, nl = lfpm.fuse_schedule(nbr, cnt, garbage)
    live = [v for v in range(V) if not garbage[v]]
    placed = np.concatenate([order[start[i]:start[i + 1]] for i in range(nl)])
    assert sorted(placed.tolist()) == live
    level = {int(v): i for i in range(nl) for v in order[start[i]:start[i + 1]]}
    for a in live:
        for b in live:
            if a < b and _conflicts(nbr, cnt, garbage, a, b):
                assert level[a] < level[b], (a, b)
    for i in range(nl):
        lv = order[start[i]:start[i + 1]].tolist()
        assert lv == sorted(lv)
        for x in range(len(lv)):
            for y in range(x + 1, len(lv)):
                assert not _conflicts(nbr, cnt, garbage, lv[x], lv[y])
