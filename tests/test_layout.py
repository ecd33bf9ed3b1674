"""The fused kernel's LDS layout arithmetic (rm_internal.h), compiled on the
host for every stage-buffer layout the build supports."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("swz", [0, 1, 2])
def test_topk_lds_layout_disjoint(tmp_path, swz):
    exe = tmp_path / f"layout{swz}"
    r = subprocess.run(["g++", "-std=c++17", "-D__HIP_PLATFORM_AMD__", f"-DRM_LDS_SWZ={swz}",
                        "-I/opt/rocm/include", os.path.join(ROOT, "tests", "layout_check.cpp"),
                        "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
