"""The C ABI library loads without a device and exports every function include/nuray.h declares."""
import ctypes
import os
import re

from conftest import ROOT


def _declared():
    src = open(os.path.join(ROOT, "include", "nuray.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(nr_[a-z0-9_]+)\s*\(", src)) - {"nr_progress_fn"})


def test_library_exports_every_declared_symbol():
    import nuray_amd

    lib = ctypes.CDLL(nuray_amd.LIB_PATH)
    names = _declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(nuray_amd.EXPORTS) <= set(names)


def test_errors_cross_the_abi_as_codes():
    import nuray_amd as nr

    L = nr.lib()
    # a bad material id is rejected with a message, no exception escapes
    import numpy as np

    p = np.zeros((1, 9), np.float32)
    h = L.nr_scene_create(p, None, None, np.array([5], np.int32), 1, np.zeros((1, 8), np.float32), 1)
    assert not h
    assert L.nr_last_error()
