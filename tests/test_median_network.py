"""The median-of-25 selection network of the 5x5 median passes (k_median5, the ten
cv::medianBlur(ksize 5) passes of DepthSolver::Virtual_depth_map_TileKeySequence,
DepthSolver.cpp:8010-8012), read back from the generated header the kernel includes.

0-1 principle: a min/max network computes the median of every input iff it computes it for
every input of 0s and 1s, so all 2^25 binary inputs are run (bit-packed, 64 per word) and
compared with "at least 13 ones".  A plain random check on byte values follows, and the header
must match what its generator prints."""
import os
import re
import subprocess
import sys

import numpy as np

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                    "plenoptic-for-depth_amd", "csrc")


def _network():
    src = open(os.path.join(CSRC, "lfpm_median25.h")).read()
    ops = re.findall(r"const T (\w+) = (vmin|vmax)\((\w+), (\w+)\);", src)
    target = re.search(r"return (\w+);", src).group(1)
    return ops, target


def _run(ops, target, inputs, vmin, vmax):
    vals = {"v%d" % i: inputs[i] for i in range(25)}
    last = {}
    for k, (dst, f, a, b) in enumerate(ops):
        last[a] = k
        last[b] = k
    for k, (dst, f, a, b) in enumerate(ops):
        vals[dst] = (vmin if f == "vmin" else vmax)(vals[a], vals[b])
        for name in (a, b):
            if last.get(name) == k and name != target:
                vals.pop(name, None)  # free what no later operation reads
    return vals[target]


def test_header_matches_generator():
    out = subprocess.run([sys.executable, os.path.join(CSRC, "gen_median25.py")],
                         capture_output=True, text=True, check=True).stdout
    assert out == open(os.path.join(CSRC, "lfpm_median25.h")).read()


def test_zero_one_principle_all_inputs():
    ops, target = _network()
    nwords = (1 << 25) // 64
    w = np.arange(nwords, dtype=np.uint64)
    low = [0xAAAAAAAAAAAAAAAA, 0xCCCCCCCCCCCCCCCC, 0xF0F0F0F0F0F0F0F0, 0xFF00FF00FF00FF00,
           0xFFFF0000FFFF0000, 0xFFFFFFFF00000000]
    inputs = []
    for i in range(25):
        if i < 6:
            inputs.append(np.full(nwords, low[i], np.uint64))
        else:
            inputs.append(np.where((w >> np.uint64(i - 6)) & np.uint64(1),
                                   np.uint64(0xFFFFFFFFFFFFFFFF), np.uint64(0)))
    got = _run(ops, target, inputs, np.bitwise_and, np.bitwise_or)
    idx = np.arange(1 << 25, dtype=np.uint32)
    want = np.packbits((np.bitwise_count(idx) >= 13).astype(np.uint8), bitorder="little")
    assert np.array_equal(got.view(np.uint8), want)


def test_random_bytes():
    ops, target = _network()
    rng = np.random.default_rng(0)
    x = rng.integers(0, 256, size=(25, 20000)).astype(np.int32)
    x[:, :100] = rng.integers(0, 3, size=(25, 100))  # many ties
    got = _run(ops, target, list(x), np.minimum, np.maximum)
    assert np.array_equal(got, np.sort(x, axis=0)[12])
