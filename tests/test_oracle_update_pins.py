"""Pins of the oracle's checkerboard update (view-selection priors, probabilities, CDF and
draws; the propagation slots; the perturbed normal; the refinement candidates) against
hand-computed vectors (tests/_update_pins.py), and a mutation check: builds of the oracle with
one constant, comparison or index changed each fail at least one of the same checks.  The
mutated builds are compiled from the oracle's source into pytest's temporary directory; nothing
of them is kept."""
import concurrent.futures as cf
import ctypes as C
import os

import pytest

import oracle as O
from _helpers import build_oracle_mutant as _build
from _update_pins import all_checks


@pytest.fixture(scope="module")
def pm(lfpm):
    import _helpers
    wl = lfpm.workload_c2(raw_w=256, raw_h=200)
    return _helpers.pm_to_oracle(O, wl.pm)


@pytest.mark.parametrize("name", [n for n, _ in all_checks(None)])
def test_update_piece(name, pm):
    dict(all_checks(pm))[name](O.lib())


# (label, [(old, new), ...]): each old text must occur exactly once in the oracle source.
MUTATIONS = [
    ("swap 0.18 / 0.32", [("(T)-0.18", "@A@"), ("(T)-0.32", "(T)-0.18"), ("@A@", "(T)-0.32")]),
    ("swap 0.9 / 0.1 priors", [("(T)0.9 : (T)0.1", "(T)0.1 : (T)0.9")]),
    ("threshold decay /90 -> /9", [("(T)(iter * iter) / (T)-90.0", "(T)(iter * iter) / (T)-9.0")]),
    ("threshold 0.8 -> 0.85", [("const T thr = (T)0.8 *", "const T thr = (T)0.85 *")]),
    ("count > 2 -> >= 2", [("if (count > (T)2 && cf < 3)", "if (count >= (T)2 && cf < 3)")]),
    ("fallback cf < 3 -> <= 3", [("else if (cf < 3)", "else if (cf <= 3)")]),
    ("count_false > 1.2 -> >= 1.2", [("if (c > (T)1.2) ++cf;", "if (c >= (T)1.2) ++cf;")]),
    ("draw without FLT_EPSILON", [(" - (T)1.1920928955078125e-07;", ";")]),
    ("fallback weight 0", [("vw[i] = (T)1; tsel |= 1u << i;", "vw[i] = (T)0.5; tsel |= 1u << i;")]),
    ("prior over present slots only", [("for (int ii = 0; ii < 4; ++ii) {\n        if (direct[ii] < 0) continue;",
                                        "for (int ii = 0; ii < 3; ++ii) {\n        if (direct[ii] < 0) continue;")]),
    ("prior up neighbour from y > 1", [("y > 0 ? center - W : -1", "y > 1 ? center - W : -1")]),
    ("rotation R[1] sign", [("R[1] = c3 * s1 * s2 - c1 * s3;", "R[1] = c3 * s1 * s2 + c1 * s3;")]),
    ("rotation transposed R[2]/R[6]", [("R[2] = s1 * s3 + c1 * c3 * s2;", "R[2] = -s2;"),
                                       ("R[6] = -s2;", "R[6] = s1 * s3 + c1 * c3 * s2;")]),
    ("revert test flipped", [("if (o.x * vx + o.y * vy + o.z * vz >= (T)0)",
                              "if (o.x * vx + o.y * vy + o.z * vz <= (T)0)")]),
    ("refine local range below 0.2", [("if (bcost < (T)0.1)\n        drand", "if (bcost < (T)0.2)\n        drand")]),
    ("refine perturbation 0.03", [("const T pert = (T)0.02;", "const T pert = (T)0.03;")]),
    ("refine local range +-0.5", [("- (T)0.5) * (T)2 + dnow;", "- (T)0.5) * (T)1 + dnow;")]),
    ("refine candidate 1 depth", [("cand[1] = P4<T>{nr[0], nr[1], nr[2], dnow};",
                                   "cand[1] = P4<T>{nr[0], nr[1], nr[2], drand};")]),
    ("refine candidate 4 normal", [("cand[4] = P4<T>{best.x, best.y, best.z, dpert};",
                                    "cand[4] = P4<T>{nr[0], nr[1], nr[2], dpert};")]),
    ("near-up diagonal mirrored", [("if (y > 1 + i && x > i) pick(b, cm, un - (1 + i) * W - i);",
                                    "if (y > 1 + i && x > i) pick(b, cm, un - (1 + i) * W + i);")]),
    ("far-up copy from y <= 1", [("if (y <= 2 && y > 0) { flag[1] = true;",
                                  "if (y <= 1 && y > 0) { flag[1] = true;")]),
    ("far extra step dropped", [("if (y > 3) pick(b, cm, center - 4 * W);", "")]),
]


def test_mutations_are_caught(tmp_path, pm):
    checks = all_checks(pm)
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        sos = list(ex.map(lambda a: _build(a[0], a[1][1], str(tmp_path)), enumerate(MUTATIONS)))
    survivors = []
    for (label, _), so in zip(MUTATIONS, sos):
        L = C.CDLL(so)
        failed = []
        for name, fn in checks:
            try:
                fn(L)
            except AssertionError:
                failed.append(name)
        if not failed:
            survivors.append(label)
    assert not survivors, "mutations no check rejects: %s" % survivors
