"""The large textured scene (9,188,616 triangles: a 2048^2-cell heightfield in 16 regions with
4096^2 RGBA8 albedo textures, one region on a 2048^2 float texture, 10,000 icospheres of which
10 % are GGX dielectrics, four area lights, the sky) on the GPU against the oracle: bit-exact
traversal, same-seed render agreement, and the texel gradient of the 2048^2 texture."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def large(nr, po, scenes):
    t0 = time.time()
    sc = scenes.large_textured()
    print(f"\nscene: {len(sc['p'])} triangles, generated in {time.time() - t0:.1f} s", flush=True)
    t0 = time.time()
    dev = nr.Scene(sc)
    print(f"device scene + SAH build {time.time() - t0:.1f} s ({dev.info()})", flush=True)
    t0 = time.time()
    ref = po.Scene(sc)
    print(f"oracle scene + SAH build {time.time() - t0:.1f} s", flush=True)
    return sc, dev, ref


def test_large_traversal_bit_exact(large):
    sc, dev, ref = large
    rng = np.random.Generator(np.random.PCG64(8))
    n = 1 << 16
    o = np.concatenate([np.tile(np.array(sc["camera"]["pos"], np.float32), (n // 2, 1)),
                        rng.uniform([0, 0, 0], [2000, 300, 2000], (n // 2, 3))]).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d[: n // 2, 1] = -np.abs(d[: n // 2, 1]) - 0.3
    d[: n // 2, 2] = -np.abs(d[: n // 2, 2]) - 0.5
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    g = dev.intersect(o, d)
    r = ref.intersect(o, d)
    print(f"hits: {np.count_nonzero(r[0] >= 0)} of {n}")
    assert np.array_equal(g[0], r[0])
    for k in (1, 2, 3):
        assert np.array_equal(g[k].view(np.uint32), r[k].view(np.uint32))


def test_large_render_same_seed(large, scenes):
    sc, dev, ref = large
    W, H, spp = 160, 90, 2
    cam = dict(sc["camera"], aspect=W / H)
    kw = dict(maxdepth=12, rrdepth=3)
    g = dev.render(cam, W, H, spp, seed=5, **kw)["sum"] / spp
    s, _ = ref.render(scenes.camera_array(cam), W, H, spp, seed=5, **kw)
    m = s.reshape(H, W, 3) / spp
    rel = np.abs(g - m) / (np.abs(m) + 1e-3)
    frac = float((rel > 1e-3).any(-1).mean())
    print(f"same seed: {frac:.2%} of pixels beyond 1e-3, means {g.mean():.5f} / {m.mean():.5f}")
    assert frac < 0.03 and abs(g.mean() - m.mean()) < 3e-3 * m.mean()


def test_large_texel_gradient(large, scenes):
    """The adjoint of the 2048^2 float texture (12.6 M parameters), recompute-only.  (1) Against
    central differences of the GPU's own render with common random numbers through in-place texture
    updates (texels enter linearly and move no path: exact up to rounding), on the texels with the
    largest gradients.  (2) Against the FP64 oracle: on this scene's 1-unit triangles the FP32 and
    FP64 paths part on about 2 % of pixels (the same-seed render test), and each texel sums few
    samples, so the comparison is of shape (correlation) and size, not of every entry."""
    sc, dev, ref = large
    W, H, spp = 96, 54, 4
    cam = dict(sc["camera"], aspect=W / H)
    kw = dict(maxdepth=12, rrdepth=3)
    M = len(sc["mats"])
    dldi = np.random.Generator(np.random.PCG64(9)).normal(size=(H, W, 3)).astype(np.float32)
    t0 = time.time()
    g, gt = dev.adjoint(cam, W, H, spp, M, seed=7, dl_dimg=dldi, tex_id=16, tex_shape=(2048, 2048), **kw)
    t_gpu = time.time() - t0
    assert np.count_nonzero(gt) > 100
    tex = sc["textures"][16].copy()
    for f in np.argsort(-np.abs(gt).reshape(-1))[:5]:
        y, x, c = np.unravel_index(f, gt.shape)
        tp, tm = tex.copy(), tex.copy()
        tp[y, x, c] += 1e-3
        tm[y, x, c] -= 1e-3
        vals = []
        for t in (tp, tm):
            dev.set_texture_rgb32f(16, t)
            vals.append((dev.render(cam, W, H, spp, seed=7, **kw)["sum"] / spp * dldi).sum())
        fd = (vals[0] - vals[1]) / (float(tp[y, x, c]) - float(tm[y, x, c]))
        print(f"texel ({y},{x},{c}): adjoint {gt[y, x, c]:.6e} central difference {fd:.6e}")
        assert abs(fd - gt[y, x, c]) <= 5e-3 * abs(fd) + 1e-7
    dev.set_texture_rgb32f(16, tex)
    go, gto = ref.adjoint(scenes.camera_array(cam), W, H, spp, dldi, M, seed=7, tex_id=16, tex_shape=(2048, 2048),
                          **kw)
    et = np.linalg.norm(gt - gto) / np.linalg.norm(gto)
    corr = float(np.corrcoef(gt.reshape(-1), gto.reshape(-1))[0, 1])
    print(f"2048^2 texel gradient ({np.count_nonzero(gto)} nonzero): relative L2 vs FP64 oracle {et:.2e}, "
          f"correlation {corr:.4f}; GPU adjoint {t_gpu * 1e3:.0f} ms (recompute-only)")
    assert corr > 0.99 and et < 0.15


def test_large_render_fp64_mre(large, scenes):
    """A crop of the large scene (the centre 160 x 90 of a 640 x 360 frame, 8 spp, same seed)
    against the FP64 oracle of the same estimator on the same random numbers: mean |gpu - fp64| /
    (|fp64| + 1e-2).  On this scene's unit-sized triangles FP32 and FP64 hit points part on a few
    per cent of paths (see the same-seed test), which bounds how close the two can be."""
    sc, dev, ref = large
    W, H, spp = 640, 360, 8
    rect = (240, 135, 400, 225)
    cam = dict(sc["camera"], aspect=W / H)
    kw = dict(maxdepth=12, rrdepth=3)
    g = dev.render(cam, W, H, spp, seed=11, **kw)["sum"] / spp
    sums = (np.zeros(W * H * 3), np.zeros(W * H * 3))
    ref.render(scenes.camera_array(cam), W, H, spp, seed=11, rect=rect, sums=sums, fp64=True, **kw)
    m = sums[0].reshape(H, W, 3) / spp
    x0, y0, x1, y1 = rect
    gc, mc = g[y0:y1, x0:x1], m[y0:y1, x0:x1]
    mre = float((np.abs(gc - mc) / (np.abs(mc) + 1e-2)).mean())
    print(f"9.19 M-triangle scene crop: mean relative error against the FP64 oracle {mre:.2e} "
          f"(means {gc.mean():.5f} / {mc.mean():.5f})")
    assert mre < 3e-2


def test_large_full_size_frame(large, scenes):
    """The scene's own frame size, 3840 x 2160, at 4 spp (maxdepth 12, RR from vertex 3) as one
    render (33 M paths: one wavefront batch), against the oracle on two 64 x 36 crops rendered with
    the full frame's pixel indices: the same seed (few pixels off, as on the small frame), and the
    mean relative error against the FP64 oracle no larger than the FP32 oracle's own (at 4 spp on
    unit-sized triangles FP32 and FP64 hit points part on a few per cent of paths, so that error is
    set by the arithmetic, not by the kernels)."""
    sc, dev, ref = large
    W, H, spp = 3840, 2160, 4
    cam = dict(sc["camera"], aspect=W / H)
    kw = dict(maxdepth=12, rrdepth=3)
    t0 = time.time()
    g = dev.render(cam, W, H, spp, seed=13, **kw)["sum"].reshape(H, W, 3) / spp
    t = time.time() - t0
    print(f"3840x2160 x {spp} spp: {t * 1e3:.0f} ms including the FP64 readback "
          f"({W * H * spp / t / 1e6:.0f} Msamples/s)")
    for rect in ((1888, 1062, 1952, 1098), (200, 1900, 264, 1936)):
        x0, y0, x1, y1 = rect
        gc = g[y0:y1, x0:x1]
        s32 = ref.render(scenes.camera_array(cam), W, H, spp, seed=13, rect=rect, **kw)[0].reshape(H, W, 3)
        s64 = ref.render(scenes.camera_array(cam), W, H, spp, seed=13, rect=rect, fp64=True, **kw)[0].reshape(H, W, 3)
        m32, m64 = s32[y0:y1, x0:x1] / spp, s64[y0:y1, x0:x1] / spp
        frac = float((np.abs(gc - m32) / (np.abs(m32) + 1e-3) > 1e-3).any(-1).mean())
        mre = float((np.abs(gc - m64) / (np.abs(m64) + 1e-2)).mean())
        mre32 = float((np.abs(m32 - m64) / (np.abs(m64) + 1e-2)).mean())
        print(f"crop {rect}: same seed {frac:.2%} of pixels beyond 1e-3; MRE against FP64 {mre:.2e} (FP32 oracle "
              f"{mre32:.2e}); means {gc.mean():.5f} / {m64.mean():.5f}")
        assert m64.mean() > 1e-3
        assert frac < 0.05
        assert mre <= 1.25 * mre32 + 2e-3
        assert abs(gc.mean() - m64.mean()) < 2e-2 * m64.mean()


def test_large_full_size_texel_adjoint(large):
    """The 2048^2 texel adjoint at 1920 x 1080 x 4 spp (8.3 M samples, the benchmark's size): the
    gradient's directional derivative along a random +-1 pattern over all 12.6 M texels against a
    central difference of the GPU render with common random numbers (texel values enter the image
    linearly along each path and move no path)."""
    sc, dev, _ = large
    W, H, spp = 1920, 1080, 4
    cam = dict(sc["camera"], aspect=W / H)
    kw = dict(maxdepth=12, rrdepth=3)
    M = len(sc["mats"])
    rng = np.random.Generator(np.random.PCG64(31))
    dldi = rng.normal(size=(H, W, 3)).astype(np.float32)
    t0 = time.time()
    _, gt = dev.adjoint(cam, W, H, spp, M, seed=7, dl_dimg=dldi, tex_id=16, tex_shape=(2048, 2048), **kw)
    t_adj = time.time() - t0
    tex = sc["textures"][16].copy()
    v = rng.choice(np.array([-1.0, 1.0], np.float32), size=tex.shape)
    h = 1e-3
    vals = []
    for sgn in (1.0, -1.0):
        t = (tex + np.float32(sgn * h) * v).astype(np.float32)
        dev.set_texture_rgb32f(16, t)
        vals.append((dev.render(cam, W, H, spp, seed=7, **kw)["sum"].reshape(H, W, 3) / spp * dldi).sum())
    dev.set_texture_rgb32f(16, tex)
    # each texel's exact float32 step between the two renders
    step = (tex + np.float32(h) * v).astype(np.float32).astype(np.float64) - \
        (tex - np.float32(h) * v).astype(np.float32).astype(np.float64)
    fd = vals[0] - vals[1]
    gv = float((gt.astype(np.float64) * step).sum())
    print(f"1080p x {spp} spp texel adjoint: {np.count_nonzero(gt)} nonzero texel gradients, {t_adj * 1e3:.0f} ms "
          f"with readback; gradient . step {gv:.6e}, central difference {fd:.6e}")
    assert np.count_nonzero(gt) > 10000
    assert abs(fd - gv) <= 5e-3 * abs(fd)
