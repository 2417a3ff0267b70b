"""Sky sampling (nr_params.env_is) on the GPU: the environment map sampled as a light at Lambert
and GGX-reflect vertices, weighted against BSDF sampling so that the expected image stays the
reference's.  Against the oracle's restatement of the same estimator at the same seed; against the
GPU's own render without it (per-pixel z-test: same expectation) with the variance it saves; and
the cases that must be refused or be no-ops."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _var(s, s2, n):
    m = s / n
    return m, np.maximum(s2 / n - m ** 2, 0) * n / (n - 1)


def _oracle(po, scenes, sc, spp, seed, **kw):
    s, s2 = po.Scene(sc).render(scenes.camera_array(sc["camera"]), sc["width"], sc["height"], spp, seed=seed, **kw)
    return s.reshape(sc["height"], sc["width"], 3), s2.reshape(sc["height"], sc["width"], 3)


def _gpu(nr, sc, spp, seed, **kw):
    dev = nr.Scene(sc)
    out = dev.render(sc["camera"], sc["width"], sc["height"], spp, seed=seed, **kw)
    dev.close()
    return out


@pytest.mark.parametrize("which", ["terrain", "zoo"])
def test_env_sampling_same_seed_as_oracle(nr, po, scenes, which):
    sc = scenes.terrain_spheres(48, 60, 128, 72) if which == "terrain" else scenes.material_zoo(128, 96)
    kw = dict(maxdepth=8, rrdepth=3, env_is=True)
    spp = 8
    g = _gpu(nr, sc, spp, seed=17, **kw)
    s, _ = _oracle(po, scenes, sc, spp, seed=17, **kw)
    m_g, m_o = g["sum"] / spp, s / spp
    rel = np.abs(m_g - m_o) / (np.abs(m_o) + 1e-3)
    frac = float((rel > 1e-3).any(-1).mean())
    print(f"{which} sky sampling same seed: frac pixels off {frac:.4f}, mean gpu {m_g.mean():.5f} "
          f"oracle {m_o.mean():.5f}")
    assert frac < 0.03
    assert abs(m_g.mean() - m_o.mean()) / m_o.mean() < 3e-3


@pytest.mark.parametrize("which", ["terrain", "zoo"])
def test_env_sampling_keeps_expectation_and_lowers_variance(nr, scenes, which):
    sc = scenes.terrain_spheres(48, 60, 128, 72) if which == "terrain" else scenes.material_zoo(128, 96)
    spp = 1024
    kw = dict(maxdepth=8, rrdepth=3)
    a = _gpu(nr, sc, spp, seed=3, **kw)
    b = _gpu(nr, sc, spp, seed=4, env_is=True, **kw)
    m0, v0 = _var(a["sum"], a["sum2"], spp)
    m1, v1 = _var(b["sum"], b["sum2"], spp)
    se = np.sqrt(v0 / spp + v1 / spp)
    ok = se > 1e-6
    z = (m1 - m0)[ok] / se[ok]
    frac3 = float((np.abs(z) > 3).mean())
    ratio = float(v0.mean() / v1.mean())
    print(f"{which}: sky sampling vs BSDF only: frac |z|>3 {frac3:.4f}, mean z {z.mean():.4f}, "
          f"image means {m0.mean():.5f} / {m1.mean():.5f}, variance ratio {ratio:.2f}")
    assert frac3 < 0.0054 and abs(float(z.mean())) < 0.05
    assert abs(m1.mean() - m0.mean()) / m0.mean() < 2e-3
    # open sky over the terrain: a large saving; the zoo sees the sky only through its open front,
    # where most sky samples are blocked, so there it only has to stay unbiased
    if which == "terrain":
        assert ratio > 3.0


def test_env_sampling_without_a_sky_is_a_no_op(nr, scenes):
    sc = scenes.box_room()
    a = _gpu(nr, sc, 4, seed=9)
    b = _gpu(nr, sc, 4, seed=9, env_is=True)
    assert np.array_equal(a["sum"], b["sum"]) and np.array_equal(a["sum2"], b["sum2"])


def test_env_sampling_refusals(nr, scenes):
    sc = scenes.terrain_spheres(16, 10, 32, 18)
    dev = nr.Scene(sc)
    with pytest.raises(nr.NuRayError, match="ptnee"):
        dev.render(sc["camera"], 32, 18, 2, integrator=1, env_is=True)
    dev.render(sc["camera"], 32, 18, 2, env_is=True, fetch=False)
    with pytest.raises(nr.NuRayError, match="environment sampling"):
        dev.adjoint(sc["camera"], 32, 18, 2, len(sc["mats"]), dl_dimg=np.zeros((18, 32, 3), np.float32), env_is=True)
    dev.close()
