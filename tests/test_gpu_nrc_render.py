"""Radiance-cache rendering on the GPU: with the cache switched off the frame is exactly the
ptnee render (same Philox numbers, same sums), training paths leave the render slots untouched,
and a trained cache lowers the error of a few-sample image."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cam(sc):
    return sc["camera"]


@pytest.mark.parametrize("which", ["box", "terrain", "terrain_sky_sampling"])
def test_cache_off_equals_ptnee(nr, scenes, which):
    if which == "box":
        sc = scenes.box_room()
        kw = dict(maxdepth=0, rrdepth=0)
    else:
        sc = scenes.terrain_spheres(n_cells=64, n_spheres=100, width=160, height=90)
        kw = dict(maxdepth=10, rrdepth=3, env_is=which == "terrain_sky_sampling")
    W, H = sc["width"], sc["height"]
    dev = nr.Scene(sc)
    ref = dev.render(_cam(sc), W, H, 4, seed=21, **kw)
    cache = nr.Nrc(seed=2)
    st, got = dev.render_nrc(cache, _cam(sc), W, H, spp=4, seed=21, use_cache=False, train=True, fetch=True, **kw)
    assert np.array_equal(got["sum"], ref["sum"]) and np.array_equal(got["sum2"], ref["sum2"])
    # training vertices are recorded (tail queries need 5-vertex paths, rare under an open sky)
    assert st["records"] > 0 and st["trained"] == min(st["records"], 1 << 16)
    assert st["queries"] > 0 or which != "box"
    assert np.isfinite(st["loss"])
    st2, got2 = dev.render_nrc(cache, _cam(sc), W, H, spp=4, seed=21, use_cache=False, train=False,
                               train_paths=0, fetch=True, **kw)
    assert np.array_equal(got2["sum"], ref["sum"]) and st2["queries"] == 0 and st2["trained"] == 0


def _box_setup(nr, scenes, ref_spp=8192):
    sc = scenes.box_room()
    W, H = sc["width"], sc["height"]
    kw = dict(maxdepth=0, rrdepth=0)
    dev = nr.Scene(sc)
    r = dev.render(sc["camera"], W, H, ref_spp, seed=999, **kw)
    ref = r["sum"] / ref_spp
    var_ref = float(np.maximum(r["sum2"] / ref_spp - ref ** 2, 0).mean() / ref_spp)
    return sc, W, H, kw, dev, ref, var_ref


def test_trained_cache_gains_over_plain_pt(nr, scenes):
    """Box room, reference semantics (no depth limit, RR at every vertex), 1 spp frames.
    MSE against an 8192-spp reference (its own variance removed) is fitted as b2 + s2 / K.
    The trained cache must (1) have a bias far below the 1e-2 RMSE target, (2) cost well under
    half a plain-PT frame at a similar per-frame variance, so (3) its time to RMSE 1e-2
    (training included) beats plain PT; an untrained cache, by contrast, has an error far above
    the target."""
    import time
    sc, W, H, kw, dev, ref, var_ref = _box_setup(nr, scenes)
    cam = sc["camera"]

    def mse(K):
        return float(((dev.fetch(W, H)["sum"] / K - ref) ** 2).mean()) - var_ref

    def frames(K, fn):
        nr.synchronize()
        t0 = time.perf_counter()
        for k in range(K):
            fn(k)
        nr.synchronize()
        return (time.perf_counter() - t0) / K

    pt_t = frames(128, lambda k: dev.render(cam, W, H, 1, seed=5, spp_offset=k, accumulate=k > 0, fetch=False, **kw))
    pt_s2 = mse(128) * 128
    untrained = nr.Nrc(seed=1)
    frames(64, lambda k: dev.render_nrc(untrained, cam, W, H, spp=1, seed=9, spp_offset=k, accumulate=k > 0,
                                        train=False, train_paths=0, **kw))
    untrained_mse = mse(64)
    cache = nr.Nrc(seed=1)
    warm_t = frames(128, lambda k: dev.render_nrc(cache, cam, W, H, spp=1, seed=3, spp_offset=k, **kw)) * 128
    fit = {}
    for K in (16, 256):
        t = frames(K, lambda k: dev.render_nrc(cache, cam, W, H, spp=1, seed=7 + K, spp_offset=k, accumulate=k > 0,
                                               **kw))
        fit[K] = (t, mse(K))
    s2 = (fit[16][1] - fit[256][1]) / (1 / 16 - 1 / 256)
    b2 = fit[256][1] - s2 / 256
    eps2 = 1e-4
    pt_time = pt_t * pt_s2 / eps2
    nrc_time = warm_t + fit[256][0] * s2 / (eps2 - b2)
    print(f"PT {pt_t * 1e3:.2f} ms/frame, s2 {pt_s2:.4f} -> {pt_time:.2f} s; cache {fit[256][0] * 1e3:.2f} ms/frame, "
          f"s2 {s2:.4f}, bias2 {b2:.2e}, warm-up {warm_t:.2f} s -> {nrc_time:.2f} s; untrained 64-frame mse "
          f"{untrained_mse:.3e} vs plain PT {pt_s2 / 64:.3e}")
    assert b2 < 0.1 * eps2
    assert fit[256][0] < 0.5 * pt_t
    assert s2 < 1.5 * pt_s2
    assert nrc_time < pt_time
    assert untrained_mse - s2 / 64 > 10 * eps2   # without training the cache's bias dominates
