"""The C-ABI library loads on a CPU host and exports every entry point include/lfpm.h declares;
host-side argument validation reports errors through status codes and lfpm_last_error()."""
import ctypes as C
import os
import re
import subprocess

import numpy as np

from conftest import PKG, ROOT


def _declared():
    src = open(os.path.join(ROOT, "include", "lfpm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(lfpm_[a-z0-9_]+)\s*\(", src)))


def test_exports_every_declared_symbol(lfpm):
    lib = lfpm.lib()
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.check_output(["nm", "-D", "--defined-only", os.path.join(PKG, "liblfpm.so")],
                                  text=True)
    exported = set(l.split()[-1] for l in out.splitlines() if " T " in l)
    assert set(names) <= exported


def test_library_holds_gfx950_code(tmp_path):
    """The embedded HIP fat binary of our own build carries a gfx950 code object."""
    fb = tmp_path / "fatbin"
    subprocess.check_call(["/opt/rocm/llvm/bin/llvm-objcopy", "--dump-section",
                           ".hip_fatbin=%s" % fb, os.path.join(PKG, "liblfpm.so"),
                           str(tmp_path / "copy.so")])
    out = subprocess.run(["/opt/rocm/llvm/bin/clang-offload-bundler", "--list", "--type=o",
                          "--input=%s" % fb], capture_output=True, text=True)
    assert "gfx950" in out.stdout, out.stdout + out.stderr


def test_layout_errors(lfpm):
    lp = lfpm.layout_params(2.0, 100, 100, 6)
    info = lfpm.LayoutInfo()
    rc = lfpm.lib().lfpm_layout_query(C.byref(lp), C.byref(info))
    assert rc == -1
    assert b"pitch" in lfpm.lib().lfpm_last_error()
    lp = lfpm.layout_params(23.0, 512, 400, 40)
    import pytest
    with pytest.raises(lfpm.LfpmError):
        lfpm.layout(lp)


def test_workspace_size_validates(lfpm):
    pm = lfpm.pm_default(16, 23.0)
    pm.max_neighbors = 18
    assert lfpm.workspace_size(100, pm) > 0
    pm.radius = 11
    import pytest
    with pytest.raises(lfpm.LfpmError):
        lfpm.workspace_size(100, pm)


def test_defaults_follow_reference(lfpm):
    pm = lfpm.pm_default(16, 23.0)
    # AdaptMIPMFrame::InitGPUParamsFromCPUParams (MIStereo/AdaptMIPMFrame.cpp:126-143)
    assert (pm.radius, pm.top_k, pm.iterations) == (2, 5, 3)
    assert pm.depth_min == 0.0 and abs(pm.depth_max - 11.5) < 1e-6
    f = lfpm.filter_default()
    # MIStereoMatch::FilterAndStatic (MIStereo/MIStereoMatch.cpp:759-778)
    assert (f.max_triplet, f.min_good_neighbors, f.spike_min_neighbors) == (4, 2, 3)
    assert abs(f.max_cost - 1.8) < 1e-6 and abs(f.max_geo_err_px - 0.5) < 1e-6


def test_device_calls_fail_loudly_without_device_pointers(lfpm):
    pm = lfpm.pm_default(16, 23.0)
    rc = lfpm.lib().lfpm_patchmatch_frame(None, None, None, None, 10, C.byref(pm), None, None,
                                          None, None, None, C.c_size_t(0), 0, None)
    assert rc != 0
    _ = np


def test_facade_library_exports_reference_entry_points():
    """liblfmvs.so (C++ facades over the C ABI) loads and defines the reference's pipeline
    entry points with their names (DepthSolverManager.h:23-26, AdaptMIPMFrame.h:22-27)."""
    import ctypes
    import os
    import subprocess
    pkg = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "plenoptic-for-depth_amd")
    so = os.path.join(pkg, "liblfmvs.so")
    ctypes.CDLL(os.path.join(pkg, "liblfpm.so"), mode=ctypes.RTLD_GLOBAL)
    ctypes.CDLL(so)
    syms = subprocess.run(["nm", "-DC", "--defined-only", so], capture_output=True, text=True,
                          check=True).stdout
    for name in ("LFMVS::DepthSolverManager::DepthSolverPipelineBySequence(LFMVS::eStereoType)",
                 "LFMVS::DepthSolverManager::DepthSolverPipelineByBatch()",
                 "LFMVS::DepthSolver::ProcessProblemsImpSequence(",
                 "LFMVS::AdaptMIPMFrame::Initialize(",
                 "LFMVS::AdaptMIPMFrame::RunPatchMatchCUDAForFrame()",
                 "LFMVS::AdaptMIPMFrame::WriteBackResults(",
                 "LFMVS::AdaptMIPMFrame::ReleaseMemory()",
                 "LFMVS::AdaptMIPM::RunPatchMatchCUDAForMI()",
                 "LFMVS::AdaptMIPM::Initialize("):
        assert name in syms, name


def test_filter_order_arguments_checked_on_the_host(lfpm):
    """The filter-order switches reject unknown orders before any device work."""
    import sys
    import pytest
    sys.path.insert(0, PKG)
    import serving
    for call in (lambda: lfpm.filter_frame(None, None, order="inplace"),
                 lambda: lfpm.FrameSlot(None, filter_order="inplace"),
                 lambda: serving.Server(filter_order="inplace")):
        with pytest.raises(ValueError):
            call()
    assert lfpm.lib().lfpm_ctx_set_filter_order(None, 1) == -1
