"""ptnee render parity on the GPU against the oracle's restatement of RendererPTLS
(rendererptls.cpp:13-236).

* same seed: GPU and oracle consume the same Philox numbers, so per-sample paths coincide except
  where a last-ulp difference of a transcendental flips a branch; per-pixel means must agree
  to 1e-3 relative on almost every pixel;
* independent seeds at high spp: per-pixel 3-sigma z-test;
* FP64 oracle of the same estimator: mean relative error stated per result."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mean(sums, spp):
    return sums / spp


def _oracle(po, sc, spp, seed, fp64=False, **kw):
    o = po.Scene(sc)
    s, s2 = o.render(_cam(sc), sc["width"], sc["height"], spp, seed=seed, fp64=fp64, **kw)
    H, W = sc["height"], sc["width"]
    return s.reshape(H, W, 3), s2.reshape(H, W, 3)


def _cam(sc):
    from nuray_amd.scenes import camera_array

    return camera_array(sc["camera"])


def _gpu(nr, sc, spp, seed, **kw):
    dev = nr.Scene(sc)
    out = dev.render(sc["camera"], sc["width"], sc["height"], spp, seed=seed, **kw)
    dev.close()
    return out


def _report(label, **vals):
    print(label + ": " + ", ".join(f"{k}={v:.4g}" for k, v in vals.items()))


@pytest.mark.parametrize("mode", ["reference", "config"])
def test_box_room_same_seed(nr, po, scenes, mode):
    sc = scenes.box_room()
    kw = dict(maxdepth=0, rrdepth=0) if mode == "reference" else dict(maxdepth=8, rrdepth=3)
    spp = 8
    g = _gpu(nr, sc, spp, seed=123, **kw)
    s, _ = _oracle(po, sc, spp, seed=123, **kw)
    m_g, m_o = g["sum"] / spp, s / spp
    rel = np.abs(m_g - m_o) / (np.abs(m_o) + 1e-3)
    frac = float((rel > 1e-3).any(-1).mean())
    _report(f"same-seed {mode}", frac_pixels_off=frac, max_rel=float(rel.max()), mean_gpu=float(m_g.mean()),
            mean_oracle=float(m_o.mean()))
    assert frac < 0.02
    assert abs(m_g.mean() - m_o.mean()) / m_o.mean() < 2e-3


def _ztest(m1, v1, n1, m2, v2, n2):
    se = np.sqrt(v1 / n1 + v2 / n2)
    ok = se > 0
    z = np.zeros_like(m1)
    z[ok] = (m1[ok] - m2[ok]) / se[ok]
    return z, ok


@pytest.mark.parametrize("mode", ["reference", "config"])
def test_box_room_ztest_independent_seeds(nr, po, scenes, mode):
    """Per-pixel 3-sigma z-test between the GPU render and the oracle at 256 spp, different seeds."""
    sc = scenes.box_room()
    kw = dict(maxdepth=0, rrdepth=0) if mode == "reference" else dict(maxdepth=8, rrdepth=3)
    spp = 256
    g = _gpu(nr, sc, spp, seed=1, **kw)
    s, s2 = _oracle(po, sc, spp, seed=2, **kw)
    m1, m2 = g["sum"] / spp, s / spp
    v1 = np.maximum(g["sum2"] / spp - m1 ** 2, 0) * spp / (spp - 1)
    v2 = np.maximum(s2 / spp - m2 ** 2, 0) * spp / (spp - 1)
    z, ok = _ztest(m1, v1, spp, m2, v2, spp)
    frac3 = float((np.abs(z[ok]) > 3).mean())
    _report(f"z-test {mode}", frac_abs_z_gt3=frac3, mean_z=float(z[ok].mean()), mean_z2=float((z[ok] ** 2).mean()))
    # heavy-tailed per-pixel estimator: allow twice the Gaussian 0.27 %
    assert frac3 < 0.0054
    assert abs(float(z[ok].mean())) < 0.05


def test_box_room_fp64_oracle_mre(nr, po, scenes):
    """Mean relative error of the GPU image against the FP64 oracle of the same estimator, same
    random numbers (so the difference is arithmetic precision, not Monte Carlo noise)."""
    sc = scenes.box_room()
    spp = 16
    g = _gpu(nr, sc, spp, seed=5)
    s, _ = _oracle(po, sc, spp, seed=5, fp64=True)
    m_g, m_o = g["sum"] / spp, s / spp
    mre = float((np.abs(m_g - m_o) / (np.abs(m_o) + 1e-2)).mean())
    _report("fp64 oracle", mre=mre)
    assert mre < 5e-3


def test_terrain_same_seed(nr, po, scenes):
    sc = scenes.terrain_spheres(n_cells=64, n_spheres=100, width=160, height=90)
    spp = 4
    kw = dict(maxdepth=10, rrdepth=3)
    g = _gpu(nr, sc, spp, seed=77, **kw)
    s, _ = _oracle(po, sc, spp, seed=77, **kw)
    m_g, m_o = g["sum"] / spp, s / spp
    rel = np.abs(m_g - m_o) / (np.abs(m_o) + 1e-3)
    frac = float((rel > 1e-3).any(-1).mean())
    _report("terrain same-seed", frac_pixels_off=frac, mean_gpu=float(m_g.mean()), mean_oracle=float(m_o.mean()))
    assert frac < 0.03


def test_split_renders_equal_one_render(nr, scenes):
    """Samples are keyed by (seed, pixel, sample): 2 x 4 spp with offsets == one 8 spp render."""
    sc = scenes.box_room()
    dev = nr.Scene(sc)
    a = dev.render(sc["camera"], 320, 240, 8, seed=9)
    dev.render(sc["camera"], 320, 240, 4, seed=9, fetch=False)
    b = dev.render(sc["camera"], 320, 240, 4, seed=9, spp_offset=4, accumulate=True)
    c = dev.render(sc["camera"], 320, 240, 8, seed=9, batch_spp=3)
    np.testing.assert_allclose(a["sum"], b["sum"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(a["sum"], c["sum"], rtol=1e-12, atol=1e-12)
    assert np.array_equal(a["rgb8"], b["rgb8"])


@pytest.mark.parametrize("env_is", [False, True])
def test_overlapped_launches_equal_serial_ones(nr, scenes, env_is):
    """Each bounce's shadow launch runs on a second stream beside the next bounce's closest-hit
    launch unless kernel timing is on (nr_set_profiling), which keeps every launch on one stream.
    The two schedules must give the same image bit for bit: the overlapped launches share no data
    (the next shade step waits for the shadow launch), so a difference would mean a race.  Several
    batches (batch_spp) and sky shadow rays (env_is) included."""
    sc = scenes.terrain_spheres(n_cells=48, n_spheres=60, width=128, height=72)
    dev = nr.Scene(sc)
    kw = dict(seed=6, maxdepth=10, rrdepth=3, batch_spp=5, env_is=env_is)
    dev.set_profiling(0)
    a = dev.render(sc["camera"], 128, 72, 16, **kw)
    dev.set_profiling(1)
    b = dev.render(sc["camera"], 128, 72, 16, **kw)
    st = dev.stats()
    dev.set_profiling(0)
    c = dev.render(sc["camera"], 128, 72, 16, **kw)
    dev.close()
    print(f"env_is {env_is}: {st['iterations']:.0f} bounces, closest {st['closest_ms']:.2f} ms, "
          f"shadow {st['shadow_ms']:.2f} ms (serial run), image mean {a['sum'].mean() / 16:.5f}")
    for r in (b, c):
        assert np.array_equal(r["sum"], a["sum"]) and np.array_equal(r["sum2"], a["sum2"])


def test_progress_and_cancel(nr, scenes):
    sc = scenes.box_room()
    dev = nr.Scene(sc)
    seen = []
    dev.render(sc["camera"], 64, 48, 8, batch_spp=2, progress=seen.append, fetch=False)
    assert seen and abs(seen[-1] - 1.0) < 1e-6 and seen == sorted(seen)
    flag = np.ones(1, np.int32)  # set before the call: the render stops at the first check
    rc = dev.render(sc["camera"], 64, 48, 8, cancel=flag, fetch=False)
    assert rc == 1


def test_cancel_mid_render_from_another_thread(nr, scenes):
    """con_flag = 0 from another thread (rendererptls.cpp:150,215,226-230): a second host thread
    writes the cancel word 0.15 s into a render of 2000 one-sample batches; the call returns
    NR_CANCELLED soon after, and the sums it leaves are exactly those of the whole batches it
    completed (equal, bit for bit, to a fresh render of that many samples per pixel)."""
    import threading
    import time
    sc = scenes.box_room()
    W, H = 320, 240
    dev = nr.Scene(sc)
    dev.render(sc["camera"], W, H, 4, batch_spp=1, fetch=False)   # warm-up
    flag = np.zeros(1, np.int32)
    t_set = []

    def raise_flag():
        time.sleep(0.15)
        t_set.append(time.perf_counter())
        flag[0] = 1
    th = threading.Thread(target=raise_flag)
    th.start()
    rc = dev.render(sc["camera"], W, H, 2000, seed=9, batch_spp=1, cancel=flag, fetch=False)
    t_ret = time.perf_counter()
    th.join()
    done = int(round(dev.stats()["samples"] / (W * H)))
    got = dev.fetch(W, H)["sum"].copy()
    print(f"cancelled after {done} of 2000 samples per pixel, returned {1e3 * (t_ret - t_set[0]):.1f} ms after "
          f"the flag was raised")
    assert rc == 1 and 0 < done < 2000
    assert t_ret - t_set[0] < 0.5
    ref = dev.render(sc["camera"], W, H, done, seed=9, batch_spp=1)["sum"]
    assert np.array_equal(got, ref)


def test_rgb8_display_transform(nr, scenes):
    """rgb8 = trunc(clamp(pow(mean, 1/2.2) * 255, 0, 255)) (rendererptls.cpp:175-181)."""
    sc = scenes.box_room()
    g = _gpu(nr, sc, 4, seed=3)
    m = (g["sum"] / 4).astype(np.float32)
    exp = np.clip(np.power(m, np.float32(1 / 2.2)) * np.float32(255), 0, 255).astype(np.int32)
    assert np.abs(exp - g["rgb8"].astype(np.int32)).max() <= 1
    np.testing.assert_allclose(g["rgb"], m, rtol=1e-6)


def test_material_zoo_same_seed(nr, po, scenes):
    """Textured Lambert and GGX-reflect, translucent GGX dielectric spheres with smooth normals,
    the sky through the open front: the full-feature shade kernel against the oracle."""
    sc = scenes.material_zoo()
    spp = 8
    g = _gpu(nr, sc, spp, seed=31)
    s, _ = _oracle(po, sc, spp, seed=31)
    m_g, m_o = g["sum"] / spp, s / spp
    rel = np.abs(m_g - m_o) / (np.abs(m_o) + 1e-3)
    frac = float((rel > 1e-3).any(-1).mean())
    _report("zoo same-seed", frac_pixels_off=frac, mean_gpu=float(m_g.mean()), mean_oracle=float(m_o.mean()))
    assert frac < 0.03
    assert abs(m_g.mean() - m_o.mean()) / m_o.mean() < 3e-3


def test_material_zoo_ztest_independent_seeds(nr, po, scenes):
    sc = scenes.material_zoo()
    spp = 256
    g = _gpu(nr, sc, spp, seed=1)
    s, s2 = _oracle(po, sc, spp, seed=2)
    m1, m2 = g["sum"] / spp, s / spp
    v1 = np.maximum(g["sum2"] / spp - m1 ** 2, 0) * spp / (spp - 1)
    v2 = np.maximum(s2 / spp - m2 ** 2, 0) * spp / (spp - 1)
    z, ok = _ztest(m1, v1, spp, m2, v2, spp)
    frac3 = float((np.abs(z[ok]) > 3).mean())
    _report("zoo z-test", frac_abs_z_gt3=frac3, mean_z=float(z[ok].mean()))
    assert frac3 < 0.0054
    assert abs(float(z[ok].mean())) < 0.05


def test_material_zoo_fp64_oracle_mre(nr, po, scenes):
    sc = scenes.material_zoo()
    spp = 16
    g = _gpu(nr, sc, spp, seed=5)
    s, _ = _oracle(po, sc, spp, seed=5, fp64=True)
    mre = float((np.abs(g["sum"] / spp - s / spp) / (np.abs(s / spp) + 1e-2)).mean())
    _report("zoo fp64 oracle", mre=mre)
    assert mre < 1e-2
