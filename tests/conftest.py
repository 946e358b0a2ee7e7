"""Shared test setup: import paths, the `gpu` marker, and built-library checks."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "plenoptic-for-depth_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct MI355X (gfx950)")
    # build only what is missing (the build step proper is __graft_entry__.build())
    if not os.path.exists(os.path.join(ORACLE, "liblfpm_oracle.so")):
        subprocess.check_call(["make", "-s", "-C", ORACLE])
    if not (os.path.exists(os.path.join(PKG, "liblfpm.so")) and
            os.path.exists(os.path.join(PKG, "liblfmvs.so"))):
        subprocess.check_call(["make", "-s", "-C", PKG])
    if not all(os.path.exists(os.path.join(ROOT, "tests", "cpp", "_build", n))
               for n in ("facade_driver", "select_neighbors_driver")):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")])


def stats_dump(name, payload):
    """Append a JSON record of measured parity statistics under $LFPM_STATS_DIR (if set)."""
    import json
    d = os.environ.get("LFPM_STATS_DIR")
    if not d:
        return
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "parity_stats.jsonl"), "a") as f:
        f.write(json.dumps(dict(test=name, **payload)) + "\n")


@pytest.fixture(scope="session")
def lfpm():
    import lfpm as m
    return m


@pytest.fixture(scope="session")
def orc():
    import oracle as m
    return m
