"""Host code under the sanitizers (CPU only):

* the oracle (oracle/orc_*.cpp) with AddressSanitizer + UndefinedBehaviorSanitizer, driven by
  tests/cpp/sanitize_oracle.cpp over every material type, a texture, a sky, both integrators, FP32
  and FP64, sky sampling, the adjoint with texel gradients and the radiance-cache network;
* the product's host SAH builder (nuray_amd/csrc/bvh_build.cpp, OpenMP tasks) with ASan + UBSan,
  and with ThreadSanitizer through the LLVM OpenMP runtime's Archer tool, driven by
  tests/cpp/sanitize_bvh.cpp (1 and 8 threads must build the same tree).

The builder shares a header with the device code, so it is compiled by hipcc with the sanitizer
flags for the host side only.  Suppressed: one allocation libomp keeps until exit (leak) and
libomp's own suspend-mutex initialisation racing its lock (tests/cpp/*_suppressions.txt); no
frame of this repository's code is suppressed."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

CPP = os.path.join(ROOT, "tests", "cpp")
HIPCC = "/opt/rocm/bin/hipcc"
ARCHER = "/opt/rocm/llvm/lib/libarcher.so"


def _run(exe, env_extra, timeout=600):
    env = dict(os.environ, **env_extra)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=timeout, env=env)
    print(r.stdout[-2000:], r.stderr[-4000:])
    return r


def test_oracle_asan_ubsan(tmp_path):
    exe = str(tmp_path / "orc_san")
    srcs = [os.path.join(CPP, "sanitize_oracle.cpp")] + [os.path.join(ROOT, "oracle", f) for f in
                                                         ("orc_capi.cpp", "orc_nrc.cpp", "orc_adjoint.cpp")]
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fopenmp", "-ffp-contract=off", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", "-o", exe] + srcs, check=True, timeout=600)
    r = _run(exe, {"ASAN_OPTIONS": "detect_leaks=1:halt_on_error=1", "UBSAN_OPTIONS": "print_stacktrace=1",
                   "OMP_NUM_THREADS": "4"})
    assert r.returncode == 0 and r.stdout.startswith("OK") and "runtime error" not in r.stderr


def _hipcc_host(args, out):
    subprocess.run([HIPCC, "-std=c++17", "-O1", "-g", "-x", "hip", "--offload-arch=gfx950"] + args + ["-o", out],
                   check=True, timeout=900)


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="needs hipcc")
def test_bvh_builder_asan_ubsan(tmp_path):
    san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined"]
    b, d, exe = (str(tmp_path / x) for x in ("bvh.o", "drv.o", "bvh_asan"))
    _hipcc_host(["-fopenmp", "-c", os.path.join(ROOT, "nuray_amd", "csrc", "bvh_build.cpp")] + san, b)
    _hipcc_host(["-c", os.path.join(CPP, "sanitize_bvh.cpp")] + san, d)
    subprocess.run([HIPCC, "-fopenmp", b, d, "-o", exe] + san, check=True, timeout=600)
    r = _run(exe, {"ASAN_OPTIONS": "detect_leaks=1:halt_on_error=1",
                   "LSAN_OPTIONS": "suppressions=" + os.path.join(CPP, "lsan_suppressions.txt"),
                   "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1"})
    assert r.returncode == 0 and r.stdout.startswith("OK") and "runtime error" not in r.stderr


@pytest.mark.skipif(not (os.path.exists(HIPCC) and os.path.exists(ARCHER)), reason="needs hipcc and Archer")
def test_bvh_builder_tsan(tmp_path):
    san = ["-Xarch_host", "-fsanitize=thread"]
    b, d, exe = (str(tmp_path / x) for x in ("bvh.o", "drv.o", "bvh_tsan"))
    _hipcc_host(["-fopenmp", "-c", os.path.join(ROOT, "nuray_amd", "csrc", "bvh_build.cpp")] + san, b)
    _hipcc_host(["-c", os.path.join(CPP, "sanitize_bvh.cpp")] + san, d)
    subprocess.run([HIPCC, "-fopenmp", b, d, "-o", exe] + san, check=True, timeout=600)
    r = _run(exe, {"OMP_TOOL_LIBRARIES": ARCHER, "OMP_NUM_THREADS": "8",
                   "TSAN_OPTIONS": "halt_on_error=1:exitcode=66:suppressions=" +
                                   os.path.join(CPP, "tsan_suppressions.txt")})
    assert r.returncode == 0 and r.stdout.startswith("OK") and "ThreadSanitizer" not in r.stderr
