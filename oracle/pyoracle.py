"""ORACLE -- test infrastructure only.

ctypes front-end of oracle/_build/liboracle.so, the CPU restatement of the reference's ptnee
path.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module,
and only as the checker; the product (nuray_amd / libnuray.so) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        L.orc_scene_create.restype = vp
        L.orc_scene_create.argtypes = [f32p, vp, vp, i32p, C.c_int, f32p, C.c_int]
        L.orc_scene_destroy.argtypes = [vp]
        L.orc_scene_set_envmap.argtypes = [vp, C.c_int, C.c_int, f32p, C.c_float]
        L.orc_scene_set_texture_rgba8.argtypes = [vp, C.c_int, C.c_int, C.c_int, u8p]
        L.orc_bvh_build.argtypes = [vp]
        L.orc_bvh_depth.argtypes = [vp]
        L.orc_bvh_export.argtypes = [vp, i32p, i32p, f32p, i32p, i32p, i32p]
        L.orc_intersect.argtypes = [vp, C.c_int, f32p, f32p, i32p, f32p, f32p, f32p, C.c_int]
        L.orc_trav_stats.argtypes = [vp, C.c_int, f32p, f32p, u64p]
        L.orc_render.argtypes = [vp, f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                 C.c_int, C.c_uint32, C.c_float, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                 f64p, f64p, C.c_int, C.c_int]
        L.orc_env_sample.argtypes = [vp, C.c_int, f32p, f32p, f32p, C.POINTER(C.c_int)]
        L.orc_env_density.argtypes = [vp, C.c_int, f64p, f64p]
        L.orc_material_sample_density.argtypes = [vp, C.c_int, C.c_int, f64p, f64p, f64p, f64p]
        L.orc_philox.argtypes = [u32p, u32p, u32p]
        L.orc_srgb8_decode.argtypes = [C.c_int]
        L.orc_srgb8_decode.restype = C.c_float
        L.orc_camera_rays.argtypes = [f32p, C.c_int, f32p, f32p, C.c_float, C.c_float, f32p]
        L.orc_texture_lookup.argtypes = [vp, C.c_int, C.c_int, f32p, f32p, f32p]
        L.orc_material_eval.argtypes = [vp, C.c_int, C.c_int, f32p, f32p, f32p, f32p, C.c_int, f32p, f32p, f32p]
        L.orc_slab.argtypes = [f32p, f32p, f32p, f32p]
        L.orc_tri_test.argtypes = [f32p, f32p, f32p, f32p]
        L.orc_set_params.argtypes = [vp, f32p, C.c_int]
        L.orc_adjoint.argtypes = [vp, f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_float, C.c_int,
                                  C.c_int, C.c_int, f32p, f64p, C.c_int, C.c_int, C.c_int, vp, C.c_int]
        L.orc_scene_set_material_ext.argtypes = [vp, C.c_int, f32p]
        L.orc_scene_set_texture_rgb32f.argtypes = [vp, C.c_int, C.c_int, C.c_int, f32p]
        L.orc_nrc_encode.argtypes = [f32p, C.c_int, f32p]
        L.orc_nrc_forward.argtypes = [f32p, f32p, C.c_int, C.c_int, f64p]
        L.orc_nrc_grad.argtypes = [f32p, f32p, f32p, C.c_int, C.c_int, f64p]
        L.orc_nrc_grad.restype = C.c_double
        L.orc_env_radiance.argtypes = [vp, C.c_int, f32p, f32p]
        L.orc_nrc_frame.argtypes = [vp, f32p, C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_float, C.c_int,
                                    C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, f32p, C.c_int,
                                    f64p, f64p, f32p, f32p, f32p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.POINTER(C.c_int)]
        L.orc_light_pick.argtypes = [vp, C.c_float]
        L.orc_light_p.argtypes = [vp]
        L.orc_light_p.restype = C.c_float
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


class Scene:
    """Oracle scene built from a scene dict of nuray_amd.scenes."""

    def __init__(self, sc, build_bvh=True):
        L = lib()
        p = _f32(sc["p"])
        n = None if sc.get("n") is None else _f32(sc["n"])
        uv = None if sc.get("uv") is None else _f32(sc["uv"])
        self._keep = (p, n, uv)
        self.h = L.orc_scene_create(p, None if n is None else n.ctypes.data, None if uv is None else uv.ctypes.data,
                                    np.ascontiguousarray(sc["mat"], np.int32), len(p), _f32(sc["mats"]),
                                    len(sc["mats"]))
        self.n_tris = len(p)
        env = sc.get("envmap")
        if env is not None:
            rgb = _f32(env["rgb"])
            L.orc_scene_set_envmap(self.h, rgb.shape[1], rgb.shape[0], rgb, float(env.get("gain", 1.0)))
        for mid, ext in (sc.get("mats_ext") or {}).items():   # Blinn-Phong Ks rows
            L.orc_scene_set_material_ext(self.h, int(mid), _f32(ext))
        for tid, tex in enumerate(sc.get("textures") or []):
            if np.asarray(tex).dtype == np.float32:   # linear RGB float texture
                t = _f32(tex)
                L.orc_scene_set_texture_rgb32f(self.h, tid, t.shape[1], t.shape[0], t)
            else:
                t = np.ascontiguousarray(tex, np.uint8)
                L.orc_scene_set_texture_rgba8(self.h, tid, t.shape[1], t.shape[0], t)
        self.n_nodes = L.orc_bvh_build(self.h) if build_bvh else 0
        if self.n_nodes < 0:
            raise RuntimeError("oracle SAH build: no split has a finite cost")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_destroy(self.h)
            self.h = None

    def depth(self):
        return lib().orc_bvh_depth(self.h)

    def export(self):
        n = self.n_nodes
        left = np.zeros(n, np.int32); right = np.zeros(n, np.int32)
        box = np.zeros((n, 6), np.float32)
        lf = np.zeros(n, np.int32); lc = np.zeros(n, np.int32)
        order = np.zeros(self.n_tris, np.int32)
        lib().orc_bvh_export(self.h, left, right, box, lf, lc, order)
        return dict(left=left, right=right, box=box, leaf_first=lf, leaf_count=lc, prim_order=order)

    def intersect(self, o, d, brute=False):
        o = _f32(o); d = _f32(d)
        n = len(o)
        tri = np.zeros(n, np.int32); t = np.zeros(n, np.float32)
        b1 = np.zeros(n, np.float32); b2 = np.zeros(n, np.float32)
        lib().orc_intersect(self.h, n, o, d, tri, t, b1, b2, int(brute))
        return tri, t, b1, b2

    def trav_stats(self, o, d):
        out = np.zeros(2, np.uint64)
        lib().orc_trav_stats(self.h, len(o), _f32(o), _f32(d), out)
        return int(out[0]), int(out[1])

    def render(self, cam, W, H, spp, spp0=0, seed=0, prr=0.8, mis=1, maxdepth=0, rrdepth=0, fp64=False,
               nthreads=0, rect=None, sums=None, integrator=0, env_is=False):
        x0, y0, x1, y1 = rect if rect is not None else (0, 0, W, H)
        if sums is None:
            sums = (np.zeros(W * H * 3), np.zeros(W * H * 3))
        r = lib().orc_render(self.h, _f32(cam), W, H, x0, y0, x1, y1, spp0, spp, seed, prr, mis, maxdepth,
                             rrdepth, int(fp64), nthreads, sums[0], sums[1], integrator, int(env_is))
        if r != 0:
            raise RuntimeError("oracle render failed")
        return sums

    def nrc_frame(self, cam, W, H, spp, w, sample0=0, seed=0, prr=0.8, mis=1, maxdepth=0, rrdepth=0, train_paths=65536,
                  train_maxlen=5, c_spread=0.01, use_cache=True, min_depth=1, nthreads=0, rec_cap=None):
        """One radiance-cache frame with the network w answering every query (orc_nrc_frame.cpp).
        Returns a dict: per-pixel sums sum / sum2, the training records recs (n x 16 floats, slot
        and depth order), the radiance their paths had collected before them lb (n x 3), their
        targets tgt (n x 3), the number of queries, and qdepth: the query vertex of every camera
        path (-1: none)."""
        rec_cap = rec_cap if rec_cap is not None else train_paths * max(1, train_maxlen - 1)
        s, s2 = np.zeros(W * H * 3), np.zeros(W * H * 3)
        recs = np.zeros((max(rec_cap, 1), 16), np.float32)
        lb = np.zeros((max(rec_cap, 1), 3), np.float32)
        tgt = np.zeros((max(rec_cap, 1), 3), np.float32)
        nr, nq = C.c_int(0), C.c_int(0)
        qd = np.zeros(W * H * spp, np.int32)
        r = lib().orc_nrc_frame(self.h, _f32(cam), W, H, spp, sample0, seed, prr, mis, maxdepth, rrdepth, train_paths,
                                train_maxlen, c_spread, int(use_cache), min_depth, _f32(w), nthreads, s, s2, recs, lb,
                                tgt, rec_cap, C.byref(nr), C.byref(nq), qd.ctypes.data_as(C.POINTER(C.c_int)))
        if r != 0:
            raise RuntimeError("oracle cache frame failed")
        n = min(nr.value, rec_cap)
        return dict(sum=s, sum2=s2, recs=recs[:n], lb=lb[:n], tgt=tgt[:n], queries=nq.value, qdepth=qd)

    def texture_lookup(self, tid, u, v):
        u = _f32(u); v = _f32(v)
        out = np.zeros((len(u), 3), np.float32)
        lib().orc_texture_lookup(self.h, tid, len(u), u, v, out)
        return out

    def material_eval(self, m, wo, n, wi=None, u3=None):
        wo = _f32(wo); n = _f32(n); k = len(wo)
        wi_in = _f32(wi if wi is not None else np.zeros((k, 3)))
        u_in = _f32(u3 if u3 is not None else np.zeros((k, 3)))
        wi_out = np.zeros((k, 3), np.float32); f = np.zeros((k, 3), np.float32); pdf = np.zeros(k, np.float32)
        lib().orc_material_eval(self.h, m, k, wo, n, wi_in, u_in, 1 if wi is None else 0, wi_out, f, pdf)
        return wi_out, f, pdf

    def env_sample(self, u4):
        """Sky sampler of the optional environment sampling: directions, densities, accepted flags."""
        u4 = _f32(u4)
        n = len(u4)
        d = np.zeros((n, 3), np.float32); p = np.zeros(n, np.float32); ok = np.zeros(n, np.int32)
        lib().orc_env_sample(self.h, n, u4, d, p, ok.ctypes.data_as(C.POINTER(C.c_int)))
        return d, p, ok.astype(bool)

    def env_density(self, d):
        d = np.ascontiguousarray(d, np.float64).reshape(-1, 3)
        out = np.zeros(len(d))
        lib().orc_env_density(self.h, len(d), d, out)
        return out

    def sample_density(self, m, wo, n, wi):
        """True density of material m's BSDF sampler at wi (Lambert and GGX-reflect, FP64)."""
        a = [np.ascontiguousarray(x, np.float64).reshape(-1, 3) for x in (wo, n, wi)]
        out = np.zeros(len(a[2]))
        lib().orc_material_sample_density(self.h, m, len(a[2]), a[0], a[1], a[2], out)
        return out

    def set_texture_rgb32f(self, tid, rgb):
        rgb = _f32(rgb)
        lib().orc_scene_set_texture_rgb32f(self.h, tid, rgb.shape[1], rgb.shape[0], rgb)

    def set_params(self, theta):
        theta = _f32(theta)
        lib().orc_set_params(self.h, theta.reshape(-1), len(theta))

    def adjoint(self, cam, W, H, spp, dldi, n_mats, spp0=0, seed=0, prr=0.8, mis=1, maxdepth=0, rrdepth=0,
                nthreads=0, tex_id=-1, tex_shape=None, integrator=0):
        """FP64 gradient (n_mats, 4) of sum(dldi * mean image) for the current parameters; with
        tex_id >= 0 also the texel gradient (h, w, 3) of that texture."""
        g = np.zeros(n_mats * 4, np.float64)
        gt = np.zeros((tex_shape[0], tex_shape[1], 3), np.float64) if tex_id >= 0 else None
        r = lib().orc_adjoint(self.h, _f32(cam), W, H, spp, spp0, seed, prr, mis, maxdepth, rrdepth,
                              _f32(dldi).reshape(-1), g, n_mats, nthreads, tex_id,
                              None if gt is None else gt.ctypes.data, integrator)
        if r != 0:
            raise RuntimeError("oracle adjoint failed")
        return g.reshape(n_mats, 4) if gt is None else (g.reshape(n_mats, 4), gt)

    def env_radiance(self, d):
        """Radiance of rays that leave the scene along the unit directions d (n x 3)."""
        d = _f32(d).reshape(-1, 3)
        out = np.zeros((len(d), 3), np.float32)
        lib().orc_env_radiance(self.h, len(d), d, out)
        return out

    def light_pick(self, u):
        return lib().orc_light_pick(self.h, float(u))

    def light_p(self):
        return lib().orc_light_p(self.h)


def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().orc_philox(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def slab(mn, mx, o, d):
    return bool(lib().orc_slab(_f32(mn), _f32(mx), _f32(o), _f32(d)))


def tri_test(p9, o, d):
    out = np.zeros(3, np.float32)
    lib().orc_tri_test(_f32(p9), _f32(o), _f32(d), out)
    return out


def srgb8_decode(c):
    return lib().orc_srgb8_decode(int(c))


def camera_rays(cam, px, py, W, H):
    px = _f32(px); py = _f32(py)
    out = np.zeros((len(px), 3), np.float32)
    lib().orc_camera_rays(_f32(cam), len(px), px, py, float(W), float(H), out)
    return out


# ---- radiance-cache network (orc_nrc.cpp)
NRC_NPARAM = 5 * 4096 + 16 * 64


def nrc_encode(q):
    q = _f32(q).reshape(-1, 16)
    out = np.zeros((len(q), 64), np.float32)
    lib().orc_nrc_encode(q, len(q), out)
    return out


def nrc_forward(w, q, bf16=False):
    q = _f32(q).reshape(-1, 16)
    y = np.zeros((len(q), 3), np.float64)
    lib().orc_nrc_forward(_f32(w), q, len(q), int(bf16), y)
    return y


def nrc_grad(w, q, t, bf16=False):
    """(mean relative L2 loss, gradient of NRC_NPARAM doubles)."""
    q = _f32(q).reshape(-1, 16)
    g = np.zeros(NRC_NPARAM, np.float64)
    loss = lib().orc_nrc_grad(_f32(w), q, _f32(t).reshape(-1, 3), len(q), int(bf16), g)
    return loss, g
