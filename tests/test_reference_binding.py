"""The reference-side binding (nuray_amd/binding/rendererptnee_mi355x.h: RendererPTNEE_MI355X, a
subclass of the reference's Renderer over libnuray.so) compiles against the reference's own headers
(src/renderer/renderer.h:15-35 and the material, texture, envmap and config headers it reads) and
Qt 5's headers.  Compile only: nothing of the reference is linked or run.  Skipped where the
reference tree or the Qt headers are absent."""
import os
import subprocess

import pytest

from conftest import ROOT

REF_SRC = "/root/reference/src"
QT = "/opt/conda/include/qt"


@pytest.mark.skipif(not (os.path.isdir(REF_SRC) and os.path.isdir(os.path.join(QT, "QtGui"))),
                    reason="needs the reference's headers and Qt 5 headers")
def test_binding_compiles_against_the_reference_headers(tmp_path):
    cmd = ["g++", "-std=c++17", "-fPIC", "-Wall", "-Werror=return-type", "-c",
           os.path.join(ROOT, "tests", "cpp", "binding_compile.cpp"), "-o", str(tmp_path / "binding.o"),
           "-I" + REF_SRC, "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "nuray_amd", "binding"),
           "-I" + QT, "-I" + os.path.join(QT, "QtCore"), "-I" + os.path.join(QT, "QtGui")]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    # the object defines the overrides of the reference's virtual interface
    syms = subprocess.run(["nm", "-C", str(tmp_path / "binding.o")], capture_output=True, text=True).stdout
    for name in ("RendererPTNEE_MI355X::prepare", "RendererPTNEE_MI355X::render", "RendererPTNEE_MI355X::setEnvmap"):
        assert name in syms, name
