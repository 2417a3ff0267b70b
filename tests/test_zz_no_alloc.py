"""nr_intersect_noalloc, the entry point that allocates nothing (its ray records, hit records,
retry list and counters live in the library's static per-GPU scratch array), against nr_intersect
on the same rays: triangle ids and t / b1 / b2 bit for bit, on two calls in a row (the second one
smaller, with the shadow rule and a one-entry first-launch stack that sends most rays through
the retry list), then the size limit and a NaN segment length rejected with a clear error.

Device arrays come from the HIP runtime libnuray itself links (hipMalloc through ctypes)."""
import ctypes as C

import numpy as np
import pytest

from raybatch import make_batch

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    import nuray_amd

    nuray_amd.lib()
    path = next(line.split()[-1] for line in open("/proc/self/maps") if "libamdhip64" in line)
    H = C.CDLL(path)
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipFree.argtypes = [C.c_void_p]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    H.hipDeviceSynchronize.argtypes = []
    return H


class _Dev:
    """Device copies of the ray arrays and room for the results (freed by close())."""

    def __init__(self, H, o, d, tmax=None):
        self.H, self.n, self.bufs = H, len(o), []
        self.o = self._put(np.ascontiguousarray(o, np.float32))
        self.d = self._put(np.ascontiguousarray(d, np.float32))
        self.tmax = None if tmax is None else self._put(np.ascontiguousarray(tmax, np.float32))
        self.tri, self.t, self.b1, self.b2 = (self._alloc(4 * self.n) for _ in range(4))

    def _alloc(self, nbytes):
        p = C.c_void_p()
        assert self.H.hipMalloc(C.byref(p), max(nbytes, 4)) == 0
        self.bufs.append(p)
        return p

    def _put(self, a):
        p = self._alloc(a.nbytes)
        assert self.H.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def results(self):
        out = [np.zeros(self.n, np.int32)] + [np.zeros(self.n, np.float32) for _ in range(3)]
        for a, p in zip(out, (self.tri, self.t, self.b1, self.b2)):
            assert self.H.hipMemcpy(a.ctypes.data, p, a.nbytes, D2H) == 0
        return out

    def close(self):
        for p in self.bufs:
            self.H.hipFree(p)
        self.bufs = []


def _noalloc(nr, dev, buf, any_hit):
    return nr.lib().nr_intersect_noalloc(dev.h, buf.o, buf.d, buf.n, int(any_hit), buf.tmax, buf.tri, buf.t, buf.b1,
                                         buf.b2)


def _same(ref, got, label):
    assert np.array_equal(ref[0], got[0]), f"{label}: triangle ids differ"
    for a, b, nm in zip(ref[1:], got[1:], ("t", "b1", "b2")):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"{label}: {nm} differs"


def test_intersect_noalloc_equals_intersect(nr, scenes):
    H = _hip()
    sc = scenes.terrain_spheres(n_cells=128, n_spheres=300, width=160, height=90)
    dev = nr.Scene(sc)
    try:
        # call 1: 200,000 closest-hit rays
        o, d = make_batch(sc, 200000, seed=17)
        print("noalloc call 1: 200000 closest-hit rays", flush=True)
        buf = _Dev(H, o, d)
        assert _noalloc(nr, dev, buf, False) == 0, nr.lib().nr_last_error()
        got = buf.results()
        buf.close()
        ref = dev.intersect(o, d)
        _same(ref, got, "call 1")
        assert (ref[0] >= 0).mean() > 0.3
        # call 2: fewer rays, shadow rule, a 1-entry first-launch stack (most rays retried)
        o2, d2 = o[:5000], d[:5000]
        tmax = np.where(ref[0][:5000] >= 0, ref[1][:5000] * np.float32(1.5), np.float32(1e3)).astype(np.float32)
        dev.set_short_stack(1)
        print("noalloc call 2: 5000 shadow rays, 1-entry short stack", flush=True)
        buf = _Dev(H, o2, d2, tmax)
        assert _noalloc(nr, dev, buf, True) == 0, nr.lib().nr_last_error()
        got = buf.results()
        buf.close()
        ref2 = dev.intersect(o2, d2, any_hit=True, tmax=tmax)
        _same(ref2, got, "call 2")
        assert 0 < (ref2[0] >= 0).sum() < 5000
        dev.set_short_stack(16)
        # the size limit is checked before anything runs
        L = nr.lib()
        rc = L.nr_intersect_noalloc(dev.h, None, None, 262145, 0, None, None, None, None, None)
        assert rc < 0 and b"maximum" in L.nr_last_error()
        # a NaN segment length is reported (the ray is traced as a zero-length segment)
        tnan = tmax[:64].copy()
        tnan[3] = np.nan
        print("noalloc call 3: NaN segment length", flush=True)
        buf = _Dev(H, o2[:64], d2[:64], tnan)
        rc = _noalloc(nr, dev, buf, True)
        buf.close()
        assert rc < 0 and b"NaN" in L.nr_last_error()
    finally:
        dev.close()
