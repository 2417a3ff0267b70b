"""The traversal kernels' filtered slab test (nr_device.h: slab_ray / slab_filter) against the
reference's slab test (bvhnode.cpp:35-48, as oracle/orc_common.h restates it): every decision the
filter reports as proven must be the reference's, with the band constants the kernels are
compiled with (read from nr_device.h).  A band below the error bound (2^-23 relative, 2^-25
absolute) must be caught, so the random and adversarial cases do reach the band's edge.
Host only (tests/cpp/slab_filter_check.cpp, built with g++)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _constants():
    src = open(os.path.join(ROOT, "nuray_amd", "csrc", "nr_device.h")).read()
    rel = re.search(r"constexpr float SLAB_BAND_REL = ([0-9.e+-]+)f;", src).group(1)
    ab = re.search(r"constexpr float SLAB_BAND_ABS = ([0-9.e+-]+)f;", src).group(1)
    return rel, ab


@pytest.fixture(scope="module")
def check(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("slab") / "slab_filter_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-o", exe,
                    os.path.join(ROOT, "tests", "cpp", "slab_filter_check.cpp")], check=True, timeout=300)
    return exe


def _run(exe, cases, seed, rel, ab):
    out = subprocess.run([exe, str(cases), str(seed), rel, ab], capture_output=True, text=True, timeout=600, check=True)
    n, clear, viol = (int(x) for x in out.stdout.split())
    return n, clear, viol


def test_proven_decisions_are_the_reference_s(check):
    rel, ab = _constants()
    n, clear, viol = _run(check, 6_000_000, 11, rel, ab)
    print(f"band rel {rel} abs {ab}: {clear} of {n} decisions proven, {viol} differ from the reference")
    assert clear > 0.3 * n and viol == 0


def test_a_band_below_the_bound_is_caught(check):
    n, clear, viol = _run(check, 6_000_000, 11, "1.1920928955078125e-07", "2.98023223876953125e-08")
    print(f"band 2^-23 / 2^-25: {viol} proven decisions differ from the reference")
    assert viol > 0
