"""The quotients the wave engine computes as Markstein-corrected products (weighted view cost
tot / wn, view-sampling exponent x / -0.18) equal the IEEE quotients on their whole input
domains.  The CPU suite checks every 1031st float numerator of each domain; the exhaustive run
is `tests/cpp/_build/quotient_check 1` (about 3.4e10 quotients)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_markstein_quotients_match_ieee():
    exe = os.path.join(ROOT, "tests", "cpp", "_build", "quotient_check")
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp"),
                           "_build/quotient_check"])
    out = subprocess.run([exe, "1031"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "0 mismatches" in out.stdout


def test_quotient_check_detects_a_missing_correction(tmp_path):
    """The checker itself: the uncorrected product a * RN(1/b) must produce mismatches."""
    src = open(os.path.join(ROOT, "tests", "cpp", "quotient_check.c")).read()
    planted = src.replace("const float q = fmaf(fmaf(-q0, b, a), y, q0);",
                          "const float q = q0;")
    assert planted != src
    p = tmp_path / "planted.c"
    p.write_text(planted)
    exe = tmp_path / "planted"
    subprocess.check_call(["gcc", "-O2", "-mfma", "-ffp-contract=off", "-o", str(exe), str(p),
                           "-lm"])
    out = subprocess.run([str(exe), "65537"], capture_output=True, text=True, timeout=300)
    assert out.returncode != 0
