"""Python binding of libnuray.so (the C ABI in include/nuray.h), used by tests and bench.py.

The library is built in-tree by __graft_entry__.build() (nuray_amd/Makefile).  There is no
Python or CPU fallback for any entry point: if libnuray.so is missing, importing `lib()` fails.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnuray.so")

MAT_LAMBERT, MAT_LIGHT, MAT_GGXREFL, MAT_GGX = 0, 1, 2, 3

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")

EXPORTS = [
    "nr_last_error", "nr_init", "nr_synchronize", "nr_scene_create", "nr_scene_destroy",
    "nr_scene_set_texture_rgba8", "nr_scene_set_envmap", "nr_bvh_build_sah", "nr_bvh_info", "nr_bvh_export",
    "nr_bvh_build_host", "nr_intersect", "nr_render_ptnee", "nr_render_fetch", "nr_render_stats",
    "nr_set_profiling", "nr_bench_copy", "nr_bench_atomics", "nr_scene_set_params", "nr_loss_grad_l2", "nr_adjoint_ptnee",
    "nr_nrc_create", "nr_nrc_destroy", "nr_nrc_set_weights", "nr_nrc_get_weights", "nr_nrc_infer",
    "nr_nrc_train_step", "nr_nrc_gradient", "nr_render_nrc", "nr_nrc_frame_records", "nr_nrc_bench", "nr_scene_set_texture_rgb32f",
    "nr_adjoint_ptnee_tex", "nr_scene_set_material_ext", "nr_scene_set_short_stack", "nr_scene_set_trace_order", "nr_shutdown",
    "nr_intersect_noalloc",
]

NRC_NPARAM = 5 * 4096 + 16 * 64


class nr_camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("euler", C.c_float * 3), ("fov", C.c_float), ("aspect", C.c_float)]


class nr_params(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32), ("spp_offset", C.c_int32),
                ("seed", C.c_uint32), ("prr", C.c_float), ("mis", C.c_int32), ("maxdepth", C.c_int32),
                ("rrdepth", C.c_int32), ("batch_spp", C.c_int32), ("accumulate", C.c_int32),
                ("integrator", C.c_int32), ("env_is", C.c_int32), ("path_kernel", C.c_int32)]


class nr_nrc_params(C.Structure):
    _fields_ = [("train_paths", C.c_int32), ("train_maxlen", C.c_int32), ("spread", C.c_float),
                ("use_cache", C.c_int32), ("train", C.c_int32), ("train_steps", C.c_int32),
                ("min_depth", C.c_int32)]


PROGRESS = C.CFUNCTYPE(None, C.c_float, C.c_void_p)

_lib = None


class NuRayError(RuntimeError):
    pass


def lib():
    """Load libnuray.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NuRayError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.nr_last_error.restype = C.c_char_p
        L.nr_init.argtypes = [C.c_int32]
        L.nr_scene_create.restype = vp
        L.nr_scene_create.argtypes = [f32p, vp, vp, i32p, C.c_uint32, f32p, C.c_uint32]
        L.nr_scene_destroy.argtypes = [vp]
        L.nr_scene_set_texture_rgba8.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, u8p]
        L.nr_scene_set_envmap.argtypes = [vp, C.c_int32, C.c_int32, f32p, C.c_float]
        L.nr_bvh_build_sah.argtypes = [vp, C.c_int32]
        L.nr_bvh_info.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_double)]
        L.nr_bvh_export.argtypes = [vp, i32p, i32p, f32p, i32p, i32p, i32p]
        L.nr_bvh_build_host.argtypes = [f32p, C.c_uint32, C.c_int32, C.POINTER(C.c_int32), i32p, i32p, f32p,
                                        i32p, i32p, i32p]
        L.nr_intersect.argtypes = [vp, f32p, f32p, C.c_uint32, C.c_int32, vp, i32p, f32p, f32p, f32p]
        L.nr_intersect_noalloc.argtypes = [vp, vp, vp, C.c_uint32, C.c_int32, vp, vp, vp, vp, vp]
        L.nr_shutdown.argtypes = []
        L.nr_render_ptnee.argtypes = [vp, C.POINTER(nr_camera), C.POINTER(nr_params), vp, vp, vp, vp, vp, vp, vp]
        L.nr_render_fetch.argtypes = [vp, vp, vp, vp, vp]
        L.nr_render_stats.argtypes = [vp, f64p]
        L.nr_set_profiling.argtypes = [vp, C.c_int32]
        L.nr_scene_set_short_stack.argtypes = [vp, C.c_int32]
        L.nr_scene_set_trace_order.argtypes = [vp, C.c_int32]
        L.nr_bench_copy.argtypes = [C.c_uint64, C.c_int32, C.POINTER(C.c_double)]
        L.nr_bench_atomics.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_double)]
        L.nr_scene_set_params.argtypes = [vp, f32p, C.c_uint32]
        L.nr_loss_grad_l2.argtypes = [vp, f32p, C.POINTER(C.c_double), vp]
        L.nr_adjoint_ptnee.argtypes = [vp, C.POINTER(nr_camera), C.POINTER(nr_params), vp, f64p]
        L.nr_scene_set_texture_rgb32f.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, f32p]
        L.nr_scene_set_material_ext.argtypes = [vp, C.c_int32, f32p]
        L.nr_adjoint_ptnee_tex.argtypes = [vp, C.POINTER(nr_camera), C.POINTER(nr_params), vp, f64p, C.c_int32, vp]
        L.nr_nrc_create.restype = vp
        L.nr_nrc_create.argtypes = [C.c_uint32, C.c_float, C.c_int32]
        L.nr_nrc_destroy.argtypes = [vp]
        L.nr_nrc_set_weights.argtypes = [vp, f32p, C.c_uint32]
        L.nr_nrc_get_weights.argtypes = [vp, f32p, C.c_uint32]
        L.nr_nrc_infer.argtypes = [vp, f32p, C.c_uint32, f32p]
        L.nr_nrc_train_step.argtypes = [vp, f32p, f32p, C.c_uint32, C.POINTER(C.c_double)]
        L.nr_nrc_gradient.argtypes = [vp, f32p, f32p, C.c_uint32, f32p, C.POINTER(C.c_double)]
        L.nr_nrc_bench.argtypes = [vp, C.c_uint32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.nr_render_nrc.argtypes = [vp, vp, C.POINTER(nr_camera), C.POINTER(nr_params), C.POINTER(nr_nrc_params), vp]
        L.nr_nrc_frame_records.argtypes = [vp, C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]
        _lib = L
    return _lib


def check(rc):
    if rc < 0:
        raise NuRayError(lib().nr_last_error().decode())
    return rc


def init(device=0):
    check(lib().nr_init(device))


def synchronize():
    check(lib().nr_synchronize())


def bench_atomics(counters, stride_ints, iters=90):
    """Million device-scope atomic adds per second in the wavefront's append pattern (block b adds
    to counter b % counters, the counters stride_ints ints apart)."""
    m = C.c_double()
    check(lib().nr_bench_atomics(counters, stride_ints, iters, C.byref(m)))
    return m.value


def bench_copy(nbytes, reps=5):
    """GB/s (read + write) of a plain device copy of nbytes."""
    g = C.c_double()
    check(lib().nr_bench_copy(nbytes, reps, C.byref(g)))
    return g.value


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def bvh_build_host(p, nthreads=0):
    """Host-only SAH build with the reference rule; returns the canonical preorder export."""
    p = _f32(p)
    n = len(p)
    cap = max(2 * n - 1, 1)
    left = np.zeros(cap, np.int32); right = np.zeros(cap, np.int32); box = np.zeros((cap, 6), np.float32)
    lf = np.zeros(cap, np.int32); lc = np.zeros(cap, np.int32); order = np.zeros(max(n, 1), np.int32)
    depth = C.c_int32(0)
    nn = check(lib().nr_bvh_build_host(p, n, nthreads, C.byref(depth), left, right, box, lf, lc, order))
    return dict(left=left[:nn], right=right[:nn], box=box[:nn], leaf_first=lf[:nn], leaf_count=lc[:nn],
                prim_order=order[:n], depth=depth.value)


def make_camera(cam):
    c = nr_camera()
    c.pos[:] = [float(x) for x in cam["pos"]]
    c.euler[:] = [float(x) for x in cam["euler"]]
    c.fov = float(cam["fov"])
    c.aspect = float(cam["aspect"])
    return c


class Scene:
    """A device scene (nr_scene) built from a nuray_amd.scenes dict."""

    def __init__(self, sc, nthreads=0, build=True):
        L = lib()
        p = _f32(sc["p"])
        n = None if sc.get("n") is None else _f32(sc["n"])
        uv = None if sc.get("uv") is None else _f32(sc["uv"])
        self._keep = (p, n, uv)
        self.n_tris = len(p)
        self.h = L.nr_scene_create(p, None if n is None else n.ctypes.data, None if uv is None else uv.ctypes.data,
                                   np.ascontiguousarray(sc["mat"], np.int32), len(p), _f32(sc["mats"]),
                                   len(sc["mats"]))
        if not self.h:
            raise NuRayError(L.nr_last_error().decode())
        for mid, ext in (sc.get("mats_ext") or {}).items():   # Blinn-Phong Ks rows
            check(L.nr_scene_set_material_ext(self.h, int(mid), _f32(ext)))
        for tid, tex in enumerate(sc.get("textures") or []):
            if np.asarray(tex).dtype == np.float32:   # linear RGB float texture
                t = _f32(tex)
                check(L.nr_scene_set_texture_rgb32f(self.h, tid, t.shape[1], t.shape[0], t))
            else:
                t = np.ascontiguousarray(tex, np.uint8)
                check(L.nr_scene_set_texture_rgba8(self.h, tid, t.shape[1], t.shape[0], t))
        env = sc.get("envmap")
        if env is not None:
            rgb = _f32(env["rgb"])
            check(L.nr_scene_set_envmap(self.h, rgb.shape[1], rgb.shape[0], rgb, float(env.get("gain", 1.0))))
        if build:
            self.build(nthreads)

    def build(self, nthreads=0):
        check(lib().nr_bvh_build_sah(self.h, nthreads))

    def close(self):
        if getattr(self, "h", None):
            lib().nr_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        nn, depth, ms = C.c_int32(), C.c_int32(), C.c_double()
        check(lib().nr_bvh_info(self.h, C.byref(nn), C.byref(depth), C.byref(ms)))
        return dict(n_nodes=nn.value, depth=depth.value, build_ms=ms.value)

    def export(self):
        nn = self.info()["n_nodes"]
        left = np.zeros(nn, np.int32); right = np.zeros(nn, np.int32); box = np.zeros((nn, 6), np.float32)
        lf = np.zeros(nn, np.int32); lc = np.zeros(nn, np.int32); order = np.zeros(self.n_tris, np.int32)
        check(lib().nr_bvh_export(self.h, left, right, box, lf, lc, order))
        return dict(left=left, right=right, box=box, leaf_first=lf, leaf_count=lc, prim_order=order)

    def intersect(self, o, d, any_hit=False, tmax=None):
        o = _f32(o); d = _f32(d)
        n = len(o)
        tri = np.zeros(n, np.int32); t = np.zeros(n, np.float32)
        b1 = np.zeros(n, np.float32); b2 = np.zeros(n, np.float32)
        tm = None if tmax is None else _f32(tmax)
        check(lib().nr_intersect(self.h, o, d, n, int(any_hit), None if tm is None else tm.ctypes.data, tri, t, b1, b2))
        return tri, t, b1, b2

    def render(self, cam, W, H, spp, seed=0, prr=0.8, mis=1, maxdepth=0, rrdepth=0, spp_offset=0, batch_spp=0,
               accumulate=False, fetch=True, progress=None, cancel=None, integrator=0, env_is=False,
               path_kernel=False):
        """path_kernel: False (wavefront), True (one lane per path) or "auto" (nr_params.path_kernel = 2)."""
        pk = 2 if path_kernel == "auto" else int(bool(path_kernel))
        p = nr_params(W, H, spp, spp_offset, seed, prr, mis, maxdepth, rrdepth, batch_spp, int(accumulate), integrator,
                      int(env_is), pk)
        c = make_camera(cam)
        cb = PROGRESS(lambda f, u: progress(f)) if progress else None
        rc = check(lib().nr_render_ptnee(self.h, C.byref(c), C.byref(p), None, None, None, None,
                                         C.cast(cb, C.c_void_p) if cb else None, None,
                                         None if cancel is None else cancel.ctypes.data))
        if not fetch:
            return rc
        return self.fetch(W, H)

    def fetch(self, W, H):
        rgb = np.zeros((H, W, 3), np.float32)
        s = np.zeros((H, W, 3), np.float64)
        s2 = np.zeros((H, W, 3), np.float64)
        rgb8 = np.zeros((H, W, 3), np.uint8)
        check(lib().nr_render_fetch(self.h, rgb.ctypes.data, s.ctypes.data, s2.ctypes.data, rgb8.ctypes.data))
        return dict(rgb=rgb, sum=s, sum2=s2, rgb8=rgb8)

    # ---- differentiable rendering
    def set_params(self, theta):
        """theta: (materials, 4) float32 rows {r, g, b, alpha}."""
        theta = _f32(theta)
        check(lib().nr_scene_set_params(self.h, theta.reshape(-1), len(theta)))

    def loss_l2(self, target, want_grad=False):
        t = _f32(target).reshape(-1)
        loss = C.c_double()
        g = np.zeros(len(t), np.float32) if want_grad else None
        check(lib().nr_loss_grad_l2(self.h, t, C.byref(loss), None if g is None else g.ctypes.data))
        return (loss.value, g) if want_grad else loss.value

    def adjoint(self, cam, W, H, spp, n_mats, seed=0, prr=0.8, mis=1, maxdepth=0, rrdepth=0, spp_offset=0,
                dl_dimg=None, tex_id=-1, tex_shape=None, integrator=0, env_is=False, path_kernel=False):
        """Material gradient (n_mats, 4); with tex_id >= 0 also the texel gradient (h, w, 3) of that
        texture (tex_shape = (h, w)).  path_kernel (as in render) picks the form of the forward
        recompute when the last render's radiance is not these samples'."""
        pk = 2 if path_kernel == "auto" else int(bool(path_kernel))
        p = nr_params(W, H, spp, spp_offset, seed, prr, mis, maxdepth, rrdepth, 0, 0, integrator, int(env_is), pk)
        c = make_camera(cam)
        grad = np.zeros(n_mats * 4, np.float64)
        dl = None if dl_dimg is None else _f32(dl_dimg).reshape(-1)
        if tex_id < 0:
            check(lib().nr_adjoint_ptnee(self.h, C.byref(c), C.byref(p), None if dl is None else dl.ctypes.data,
                                         grad))
            return grad.reshape(n_mats, 4)
        gt = np.zeros((tex_shape[0], tex_shape[1], 3), np.float64)
        check(lib().nr_adjoint_ptnee_tex(self.h, C.byref(c), C.byref(p), None if dl is None else dl.ctypes.data,
                                         grad, tex_id, gt.ctypes.data))
        return grad.reshape(n_mats, 4), gt

    def set_texture_rgb32f(self, tex_id, rgb):
        rgb = _f32(rgb)
        check(lib().nr_scene_set_texture_rgb32f(self.h, tex_id, rgb.shape[1], rgb.shape[0], rgb))

    def set_texture_rgba8(self, tex_id, rgba):
        t = np.ascontiguousarray(rgba, np.uint8)
        check(lib().nr_scene_set_texture_rgba8(self.h, tex_id, t.shape[1], t.shape[0], t))

    def set_envmap(self, rgb, gain=1.0):
        rgb = _f32(rgb)
        check(lib().nr_scene_set_envmap(self.h, rgb.shape[1], rgb.shape[0], rgb, float(gain)))

    def set_material_ext(self, mat_id, ext):
        check(lib().nr_scene_set_material_ext(self.h, int(mat_id), _f32(ext)))

    def render_nrc(self, cache, cam, W, H, spp=1, seed=0, prr=0.8, mis=1, maxdepth=0, rrdepth=0, spp_offset=0,
                   accumulate=False, train_paths=65536, train_maxlen=5, spread=0.01, use_cache=True, train=True,
                   train_steps=4, min_depth=1, fetch=False, env_is=False):
        """One radiance-cache frame (nr_render_nrc).  Returns the frame statistics (and the fetched
        image when fetch=True)."""
        p = nr_params(W, H, spp, spp_offset, seed, prr, mis, maxdepth, rrdepth, 0, int(accumulate), 0, int(env_is))
        c = make_camera(cam)
        npar = nr_nrc_params(train_paths, train_maxlen, spread, int(use_cache), int(train), train_steps, min_depth)
        st = np.zeros(4, np.float64)
        check(lib().nr_render_nrc(self.h, cache.h, C.byref(c), C.byref(p), C.byref(npar), st.ctypes.data))
        stats = dict(loss=st[0], queries=int(st[1]), records=int(st[2]), trained=int(st[3]))
        if fetch:
            return stats, self.fetch(W, H)
        return stats

    def nrc_frame_records(self, cap=1 << 16):
        """The last cache frame's training records (n x 16 floats: float 11 = slot bits, 12..14 = beta)
        and targets (n x 4; xyz used) (nr_nrc_frame_records)."""
        recs = np.zeros((cap, 16), np.float32)
        tgt = np.zeros((cap, 4), np.float32)
        n = C.c_int32(0)
        check(lib().nr_nrc_frame_records(self.h, recs.ctypes.data, tgt.ctypes.data, cap, C.byref(n)))
        k = min(n.value, cap)
        return recs[:k], tgt[:k]

    def set_profiling(self, mode):
        check(lib().nr_set_profiling(self.h, mode))

    def set_short_stack(self, entries):
        """LDS stack entries of the first trace launch (rays needing more are traced again)."""
        check(lib().nr_scene_set_short_stack(self.h, entries))

    def set_trace_order(self, mode):
        """Order the wavefront's trace launches take rays in: 0 queue order, 1 direction octant per
        1024 rays, 2 sixty-four direction bins per 4096 rays, 3 (default) shadow rays alone by the
        64 bins per 256 rays, 4 the same per 1024 rays (results are the same)."""
        check(lib().nr_scene_set_trace_order(self.h, mode))

    def stats(self):
        out = np.zeros(16, np.float64)
        check(lib().nr_render_stats(self.h, out))
        keys = ["ms", "samples", "iterations", "closest_ms", "closest_launches", "rays", "node_fetch", "tri_tests",
                "shadow_ms", "shadow_rays", "shadow_node_fetch", "shadow_tri_tests", "shade_ms", "shade_launches",
                "shadow_launches", "retried"]
        return dict(zip(keys, out.tolist()))


class Nrc:
    """The radiance-cache network (nr_nrc): 16-float query rows, RGB outputs."""

    def __init__(self, seed=0, lr=1e-3, max_batch=1 << 16):
        self.h = lib().nr_nrc_create(seed, lr, max_batch)
        if not self.h:
            raise NuRayError(lib().nr_last_error().decode())

    def close(self):
        if getattr(self, "h", None):
            lib().nr_nrc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def weights(self):
        w = np.zeros(NRC_NPARAM, np.float32)
        check(lib().nr_nrc_get_weights(self.h, w, NRC_NPARAM))
        return w

    def set_weights(self, w):
        check(lib().nr_nrc_set_weights(self.h, _f32(w).reshape(-1), NRC_NPARAM))

    def infer(self, q):
        q = _f32(q).reshape(-1, 16)
        out = np.zeros((len(q), 3), np.float32)
        check(lib().nr_nrc_infer(self.h, q, len(q), out))
        return out

    def train_step(self, q, t):
        q = _f32(q).reshape(-1, 16)
        loss = C.c_double()
        check(lib().nr_nrc_train_step(self.h, q, _f32(t).reshape(-1, 3), len(q), C.byref(loss)))
        return loss.value

    def bench(self, n, reps=20):
        """(ms per inference launch over n synthetic queries, ms per training step)."""
        a, b = C.c_double(), C.c_double()
        check(lib().nr_nrc_bench(self.h, n, reps, C.byref(a), C.byref(b)))
        return a.value, b.value

    def gradient(self, q, t):
        q = _f32(q).reshape(-1, 16)
        g = np.zeros(NRC_NPARAM, np.float32)
        loss = C.c_double()
        check(lib().nr_nrc_gradient(self.h, q, _f32(t).reshape(-1, 3), len(q), g, C.byref(loss)))
        return loss.value, g
