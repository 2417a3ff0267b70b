"""The 4-wide form of the SAH tree and the flat leaf list that the GPU walks are checked on the host
against the binary tree they come from (tests/cpp/test_bvh_wide.cpp, compiled with hipcc together
with nuray_amd/csrc/bvh_build.cpp): same leaves and boxes, the same set of leaves reached by
random, axis-aligned, nearly axis-aligned and vertex-aimed rays, the stack bound, the flat list."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not (os.path.exists(HIPCC) or shutil.which("hipcc")), reason="hipcc not installed")
def test_wide_tree_and_flat_list_match_binary_tree():
    out = os.path.join(HERE, "cpp", "_bin")
    os.makedirs(out, exist_ok=True)
    exe = os.path.join(out, "test_bvh_wide")
    cmd = [HIPCC if os.path.exists(HIPCC) else "hipcc", "-std=c++17", "-O2", "-ffp-contract=off", "-fopenmp",
           "-x", "hip", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I" + os.path.join(ROOT, "nuray_amd", "csrc"), os.path.join(HERE, "cpp", "test_bvh_wide.cpp"),
           os.path.join(ROOT, "nuray_amd", "csrc", "bvh_build.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-2000:]
    assert "all checks passed" in r.stdout
