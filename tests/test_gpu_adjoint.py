"""Differentiable ptnee on the GPU: path-replay gradients against (1) the FP64 oracle's forward
recursion on the same random numbers, (2) central finite differences with common random numbers
(h = 1e-3), (3) an Adam descent that recovers the hidden parameters."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(scenes, W=64, H=64):
    return scenes.adjoint_problem(width=W, height=H)


def _render_loss(dev, sc, theta, spp, seed, target):
    dev.set_params(theta)
    dev.render(sc["camera"], sc["width"], sc["height"], spp, seed=seed, fetch=False)
    return dev.loss_l2(target)


def test_adjoint_matches_fp64_oracle(nr, po, scenes):
    sc, theta_star, theta0 = _problem(scenes)
    W, H, spp = sc["width"], sc["height"], 8
    M = len(theta0)
    rng = np.random.Generator(np.random.PCG64(3))
    dldi = rng.normal(size=(H, W, 3)).astype(np.float32) * 1e-3
    dev = nr.Scene(sc)
    dev.set_params(theta0)
    dev.render(sc["camera"], W, H, spp, seed=11, fetch=False)
    g_gpu = dev.adjoint(sc["camera"], W, H, spp, M, seed=11, dl_dimg=dldi)
    ref = po.Scene(sc)
    ref.set_params(theta0)
    g_orc = ref.adjoint(scenes.camera_array(sc["camera"]), W, H, spp, dldi, M, seed=11)
    err = np.linalg.norm(g_gpu - g_orc) / np.linalg.norm(g_orc)
    print(f"adjoint vs FP64 oracle: relative L2 error {err:.2e}, |g| = {np.linalg.norm(g_orc):.3e}")
    assert err < 2e-3
    big = np.abs(g_orc) > 1e-3 * np.abs(g_orc).max()
    rel = np.abs(g_gpu - g_orc)[big] / np.abs(g_orc)[big]
    assert np.median(rel) < 1e-3


def test_albedo_gradient_vs_central_fd(nr, scenes):
    """Albedo enters each sample multilinearly: with common random numbers the central difference
    of the loss matches the adjoint closely."""
    sc, theta_star, theta0 = _problem(scenes)
    W, H, spp = sc["width"], sc["height"], 32
    M = len(theta0)
    dev = nr.Scene(sc)
    tgt = dev.render(sc["camera"], W, H, 128, seed=99)["rgb"]
    dev.set_params(theta0)
    dev.render(sc["camera"], W, H, spp, seed=7, fetch=False)
    dev.loss_l2(tgt)
    g = dev.adjoint(sc["camera"], W, H, spp, M, seed=7)
    h = 1e-3
    checked = 0
    for (mi, ch) in [(0, 0), (5, 1), (17, 2), (40, 0), (57, 1), (62, 2)]:
        tp, tm = theta0.copy(), theta0.copy()
        tp[mi, ch] += h
        tm[mi, ch] -= h
        fd = (_render_loss(dev, sc, tp, spp, 7, tgt) - _render_loss(dev, sc, tm, spp, 7, tgt)) / (2 * h)
        print(f"albedo m{mi} c{ch}: adjoint {g[mi, ch]:.6e} fd {fd:.6e}")
        assert abs(g[mi, ch] - fd) <= 1e-2 * abs(fd) + 1e-9
        checked += 1
    assert checked == 6


def test_roughness_gradient_vs_central_fd(nr, scenes):
    """GGX roughness moves sampled directions, so per realisation the common-random-number
    difference (paths move) and the adjoint (directions detached, score term added) differ; both
    estimate the derivative of the expected image.  Their means over independent seeds are
    compared within the standard errors.  A moving direction crosses visibility edges, which
    gives the CRN difference a variance of order 1/h: h = 1e-2 here (the albedo check uses 1e-3),
    with a curvature bias below the tolerance.  The four roughness parameters with the largest
    derivative are checked, and the check must resolve each of them."""
    sc, theta_star, theta0 = _problem(scenes, 128, 128)
    W, H, spp = sc["width"], sc["height"], 256
    M = len(theta0)
    dev = nr.Scene(sc)
    rng = np.random.Generator(np.random.PCG64(5))
    w = rng.uniform(0.5, 1.5, size=(H, W, 3)).astype(np.float32) / (W * H)
    dev.set_params(theta0)
    dev.render(sc["camera"], W, H, spp, seed=77, fetch=False)
    g0 = dev.adjoint(sc["camera"], W, H, spp, M, seed=77, dl_dimg=w)[:64, 3]
    params = [int(i) for i in np.argsort(-np.abs(g0))[:4]]
    h = 1e-2
    n_seeds = 24
    adj, fd = {p: [] for p in params}, {p: [] for p in params}
    for s in range(n_seeds):
        dev.set_params(theta0)
        dev.render(sc["camera"], W, H, spp, seed=1000 + s, fetch=False)
        g = dev.adjoint(sc["camera"], W, H, spp, M, seed=1000 + s, dl_dimg=w)
        for p in params:
            adj[p].append(g[p, 3])
            vals, steps = [], []
            for sgn in (1, -1):
                t = theta0.copy()
                t[p, 3] += sgn * h
                steps.append(float(t[p, 3]))
                dev.set_params(t)
                vals.append((dev.render(sc["camera"], W, H, spp, seed=1000 + s)["sum"] / spp * w).sum())
            fd[p].append((vals[0] - vals[1]) / (steps[0] - steps[1]))
    for p in params:
        a, f = np.array(adj[p]), np.array(fd[p])
        sa, sf = a.std(ddof=1) / np.sqrt(n_seeds), f.std(ddof=1) / np.sqrt(n_seeds)
        z = (a.mean() - f.mean()) / np.hypot(sa, sf)
        print(f"roughness m{p}: adjoint {a.mean():.5e} +- {sa:.1e}, fd {f.mean():.5e} +- {sf:.1e}, "
              f"z = {z:.2f}, rel diff {abs(a.mean() - f.mean()) / abs(f.mean()):.3f}")
        assert abs(z) < 4
        assert abs(f.mean()) > 5 * sf and abs(a.mean()) > 5 * sa


def test_adjoint_roughness_columns_match_oracle(nr, po, scenes):
    sc, theta_star, theta0 = _problem(scenes, 48, 48)
    W, H, spp = sc["width"], sc["height"], 16
    M = len(theta0)
    w = np.full((H, W, 3), 1.0 / (W * H), np.float32)
    dev = nr.Scene(sc)
    dev.set_params(theta0)
    dev.render(sc["camera"], W, H, spp, seed=5, fetch=False)
    g = dev.adjoint(sc["camera"], W, H, spp, M, seed=5, dl_dimg=w)[:64, 3]
    ref = po.Scene(sc)
    ref.set_params(theta0)
    r = ref.adjoint(scenes.camera_array(sc["camera"]), W, H, spp, w, M, seed=5)[:64, 3]
    err = np.linalg.norm(g - r) / np.linalg.norm(r)
    print(f"roughness columns vs FP64 oracle: relative L2 error {err:.2e}")
    assert err < 5e-3


def _adam(theta, g, m, v, it, lr=1e-2, b1=0.9, b2=0.999):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    step = lr * (m / (1 - b1 ** it)) / (np.sqrt(v / (1 - b2 ** it)) + 1e-8)
    return m, v, theta - step


def test_adam_inverse_rendering_256(nr, scenes):
    """The inverse-rendering workload: 256x256 target at 1024 spp from hidden parameters, start
    perturbed by +-30 %, per iteration a 32-spp render, L2 loss, path-replay adjoint to all
    256 parameters and Adam (lr 1e-2), 200 iterations.  Progress is measured on the converged
    image: the excess MSE of a 1024-spp render over the noise floor must fall by 4x."""
    sc, theta_star, theta0 = _problem(scenes, 256, 256)
    W, H = 256, 256
    M = len(theta0)
    cam = sc["camera"]
    dev = nr.Scene(sc)
    dev.set_params(theta_star)
    tgt = dev.render(cam, W, H, 1024, seed=1234)["rgb"]
    floor = float(((dev.render(cam, W, H, 1024, seed=4321)["rgb"] - tgt) ** 2).mean())
    emit = sc["mats"][:, 0] == scenes.LIGHT
    theta = theta0.astype(np.float64)
    m = np.zeros_like(theta); v = np.zeros_like(theta)
    lo = np.array([1e-3, 1e-3, 1e-3, 0.02]); hi = np.array([1.0, 1.0, 1.0, 1.0])
    import time
    t0 = time.perf_counter()
    for it in range(1, 201):
        dev.set_params(theta.astype(np.float32))
        dev.render(cam, W, H, 32, seed=it, fetch=False)
        dev.loss_l2(tgt)
        g = dev.adjoint(cam, W, H, 32, M, seed=it)
        g[emit] = 0
        m, v, theta = _adam(theta, g, m, v, it)
        theta = np.where(emit[:, None], theta0, np.clip(theta, lo, hi))
    ms = (time.perf_counter() - t0) * 1e3 / 200

    def excess(th):
        dev.set_params(th.astype(np.float32))
        return float(((dev.render(cam, W, H, 1024, seed=999)["rgb"] - tgt) ** 2).mean()) - floor
    e0, e1 = excess(theta0), excess(theta)
    p0 = np.abs(theta0 - theta_star)[~emit].mean(0)
    p1 = np.abs(theta - theta_star)[~emit].mean(0)
    print(f"{ms:.2f} ms/iteration; excess image MSE {e0:.3e} -> {e1:.3e} (noise floor {floor:.3e}); "
          f"mean |theta - theta*| rgb,alpha {np.round(p0, 4)} -> {np.round(p1, 4)}")
    assert e1 < 0.25 * e0


def test_texel_gradients_match_oracle(nr, po, scenes):
    """Albedo textures (Lambert floor, GGX-reflect back wall): the texel gradients and the material
    rows against the FP64 oracle on the same random numbers."""
    sc = scenes.textured_room(width=48, height=36)
    W, H, spp = 48, 36, 8
    M = len(sc["mats"])
    rng = np.random.Generator(np.random.PCG64(21))
    dldi = rng.normal(size=(H, W, 3)).astype(np.float32) * 1e-2
    dev = nr.Scene(sc)
    ref = po.Scene(sc)
    cam = sc["camera"]
    for tid in (0, 1):
        shape = sc["textures"][tid].shape[:2]
        dev.render(cam, W, H, spp, seed=17, fetch=False)
        g, gt = dev.adjoint(cam, W, H, spp, M, seed=17, dl_dimg=dldi, tex_id=tid, tex_shape=shape)
        go, gto = ref.adjoint(scenes.camera_array(cam), W, H, spp, dldi, M, seed=17, tex_id=tid, tex_shape=shape)
        et = np.linalg.norm(gt - gto) / np.linalg.norm(gto)
        em = np.linalg.norm(g - go) / np.linalg.norm(go)
        print(f"texture {tid} {shape}: texel gradient relative L2 {et:.2e}, material rows {em:.2e}")
        assert et < 2e-3 and em < 2e-3
        assert np.all(g[tid, :3] == 0)


def test_recompute_only_replay_equals_stored(nr, scenes):
    """Without a matching single-batch render the adjoint replays each sample forward first; the
    gradient equals the one computed from the stored radiance."""
    sc = scenes.textured_room(width=40, height=30)
    W, H, spp = 40, 30, 8
    M = len(sc["mats"])
    dldi = np.random.Generator(np.random.PCG64(2)).normal(size=(H, W, 3)).astype(np.float32)
    dev = nr.Scene(sc)
    dev.render(sc["camera"], W, H, spp, seed=3, fetch=False)
    g1, t1 = dev.adjoint(sc["camera"], W, H, spp, M, seed=3, dl_dimg=dldi, tex_id=0, tex_shape=(32, 32))
    dev.render(sc["camera"], W, H, 2, seed=99, fetch=False)          # other samples: stored radiance unusable
    g2, t2 = dev.adjoint(sc["camera"], W, H, spp, M, seed=3, dl_dimg=dldi, tex_id=0, tex_shape=(32, 32))
    assert np.linalg.norm(g1 - g2) <= 1e-6 * np.linalg.norm(g1)
    assert np.linalg.norm(t1 - t2) <= 1e-6 * np.linalg.norm(t1)


def test_texel_gradient_vs_central_fd_and_in_place_update(nr, scenes):
    """Texels enter linearly: central differences with common random numbers (h = 1e-3) through
    in-place texture updates (nr_scene_set_texture_rgb32f on the built scene) match the gradient."""
    sc = scenes.textured_room(width=48, height=36)
    W, H, spp = 48, 36, 16
    M = len(sc["mats"])
    dldi = np.random.Generator(np.random.PCG64(8)).normal(size=(H, W, 3)).astype(np.float32)
    dev = nr.Scene(sc)
    tex = sc["textures"][0].copy()
    dev.render(sc["camera"], W, H, spp, seed=4, fetch=False)
    _, gt = dev.adjoint(sc["camera"], W, H, spp, M, seed=4, dl_dimg=dldi, tex_id=0, tex_shape=tex.shape[:2])
    for f in np.argsort(-np.abs(gt).reshape(-1))[:6]:
        y, x, c = np.unravel_index(f, gt.shape)
        tp, tm = tex.copy(), tex.copy()
        tp[y, x, c] += 1e-3
        tm[y, x, c] -= 1e-3
        vals = []
        for t in (tp, tm):
            dev.set_texture_rgb32f(0, t)
            vals.append((dev.render(sc["camera"], W, H, spp, seed=4)["sum"] / spp * dldi).sum())
        fd = (vals[0] - vals[1]) / (float(tp[y, x, c]) - float(tm[y, x, c]))
        print(f"texel ({y},{x},{c}): adjoint {gt[y, x, c]:.6e} fd {fd:.6e}")
        assert abs(fd - gt[y, x, c]) <= 5e-3 * abs(fd) + 1e-7
    dev.set_texture_rgb32f(0, tex)


def test_adjoint_through_dielectrics_and_rgba8_texture(nr, po, scenes):
    """Material zoo (GGX dielectric spheres on the paths, RGBA8 textures): material rows and the
    floor texture's texel gradient against the FP64 oracle."""
    sc = scenes.material_zoo(48, 36)
    W, H, spp = 48, 36, 8
    M = len(sc["mats"])
    dldi = np.random.Generator(np.random.PCG64(6)).normal(size=(H, W, 3)).astype(np.float32) * 1e-2
    dev = nr.Scene(sc)
    dev.render(sc["camera"], W, H, spp, seed=23, fetch=False)
    g, gt = dev.adjoint(sc["camera"], W, H, spp, M, seed=23, dl_dimg=dldi, tex_id=0, tex_shape=(64, 64))
    ref = po.Scene(sc)
    go, gto = ref.adjoint(scenes.camera_array(sc["camera"]), W, H, spp, dldi, M, seed=23, tex_id=0,
                          tex_shape=(64, 64))
    em = np.linalg.norm(g - go) / np.linalg.norm(go)
    et = np.linalg.norm(gt - gto) / np.linalg.norm(gto)
    print(f"zoo adjoint: material rows {em:.2e}, floor texels {et:.2e}")
    # refraction makes FP32 and FP64 paths part more often than reflection does, and each texel
    # sums few samples here: the texel comparison is looser than the material rows'
    assert em < 5e-3 and et < 2e-2
    assert not g[6:8].any()   # the dielectrics' rows


@pytest.mark.parametrize("pk", [False, True, "auto"])
def test_recompute_by_wavefront_equals_stored(nr, scenes, pk):
    """On a scene too large for the flat leaf list the adjoint recomputes its samples' radiance with
    the wavefront (path_kernel False) or one lane per path (True, and "auto" at this size), leaving
    the accumulated image alone; the gradient equals the one taken from the render's stored
    per-sample radiance up to the order of the float atomic sums."""
    sc = scenes.terrain_spheres(n_cells=48, n_spheres=60, width=96, height=54)
    W, H, spp, M = 96, 54, 8, len(sc["mats"])
    kw = dict(maxdepth=6, rrdepth=2)
    dldi = np.random.Generator(np.random.PCG64(9)).normal(size=(H, W, 3)).astype(np.float32) * 1e-2
    dev = nr.Scene(sc)
    dev.render(sc["camera"], W, H, spp, seed=5, fetch=False, **kw)
    g_stored = dev.adjoint(sc["camera"], W, H, spp, M, seed=5, dl_dimg=dldi, **kw)
    dev.render(sc["camera"], W, H, spp, seed=6, fetch=False, **kw)  # its radiance is not the adjoint's
    before = dev.fetch(W, H)["sum"].copy()
    g_re = dev.adjoint(sc["camera"], W, H, spp, M, seed=5, dl_dimg=dldi, path_kernel=pk, **kw)
    after = dev.fetch(W, H)["sum"]
    dev.close()
    err = np.linalg.norm(g_re - g_stored) / np.linalg.norm(g_stored)
    print(f"recompute (path_kernel={pk}) vs stored radiance: relative L2 {err:.2e}, |g| {np.linalg.norm(g_stored):.3e}")
    assert np.linalg.norm(g_stored) > 0
    assert err <= 1e-6
    assert np.array_equal(before, after)


def test_scene_change_after_render_is_not_replayed_stale(nr, scenes):
    """The adjoint reuses the last render's per-sample radiance and vertex records only while the
    scene is unchanged: render(theta0), set_params(theta1), adjoint must equal render(theta1),
    adjoint -- and likewise after an in-place texture update.  (Gradients sum per block in FP32
    before the FP64 accumulation, so two replays agree to ~1e-7, not bit for bit; a stale replay
    is off by percents.)"""
    sc, theta_star, theta0 = _problem(scenes, 32, 32)
    W, H, spp = 32, 32, 8
    M = len(theta0)
    dldi = np.random.Generator(np.random.PCG64(31)).normal(size=(H, W, 3)).astype(np.float32) * 1e-3
    dev = nr.Scene(sc)
    dev.set_params(theta0)
    dev.render(sc["camera"], W, H, spp, seed=5, fetch=False)
    dev.set_params(theta_star)
    g_stale = dev.adjoint(sc["camera"], W, H, spp, M, seed=5, dl_dimg=dldi)
    dev.render(sc["camera"], W, H, spp, seed=5, fetch=False)
    g_fresh = dev.adjoint(sc["camera"], W, H, spp, M, seed=5, dl_dimg=dldi)
    dev.set_params(theta0)
    dev.render(sc["camera"], W, H, spp, seed=5, fetch=False)
    g_old = dev.adjoint(sc["camera"], W, H, spp, M, seed=5, dl_dimg=dldi)
    dev.close()
    err = np.abs(g_stale - g_fresh).max() / np.abs(g_fresh).max()
    gap = np.abs(g_old - g_fresh).max() / np.abs(g_fresh).max()
    print(f"params changed after the render: max relative difference {err:.2e} (gradient at theta0: {gap:.2e})")
    assert err < 1e-5 and gap > 1e-2
    ts = scenes.textured_room(width=24, height=20)
    W, H, spp = 24, 20, 4
    M = len(ts["mats"])
    dldi = np.random.Generator(np.random.PCG64(32)).normal(size=(H, W, 3)).astype(np.float32) * 1e-3
    tex = ts["textures"][0]
    dev = nr.Scene(ts)
    dev.render(ts["camera"], W, H, spp, seed=6, fetch=False)
    dev.set_texture_rgb32f(0, tex[::-1].copy() * 0.5)
    g_stale, t_stale = dev.adjoint(ts["camera"], W, H, spp, M, seed=6, dl_dimg=dldi, tex_id=0, tex_shape=tex.shape[:2])
    dev.render(ts["camera"], W, H, spp, seed=6, fetch=False)
    g_fresh, t_fresh = dev.adjoint(ts["camera"], W, H, spp, M, seed=6, dl_dimg=dldi, tex_id=0, tex_shape=tex.shape[:2])
    dev.close()
    assert np.abs(g_stale - g_fresh).max() <= 1e-5 * np.abs(g_fresh).max()
    assert np.abs(t_stale - t_fresh).max() <= 1e-5 * np.abs(t_fresh).max()
