"""The benchmark's own workload -- the 684,296-triangle heightfield + 2000 spheres + 4 emissive
quads + 1024x512 sky scene at 1280x720, max depth 10, Russian roulette from vertex 3 -- rendered
in full on the GPU and compared with the oracle on crops of the same frame: the Philox streams
are keyed by the pixel index of the full frame, so a crop rendered by the oracle consumes the same
random numbers as the GPU's full frame.

* same seed, 64 spp (the benchmark's frame): per-pixel means agree to 1e-3 relative on almost
  every crop pixel;
* independent seeds, 256 spp: per-pixel 3-sigma z-test on the crops."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CROPS = [(100, 60, 148, 108), (600, 300, 648, 348), (1100, 500, 1148, 548), (640, 80, 688, 128)]


@pytest.fixture(scope="module")
def headline(nr, po, scenes):
    sc = scenes.terrain_spheres()
    return sc, nr.Scene(sc), po.Scene(sc)


def _oracle_crops(po_scene, scenes, sc, spp, seed, kw):
    W, H = sc["width"], sc["height"]
    sums = (np.zeros(W * H * 3), np.zeros(W * H * 3))
    for rect in CROPS:
        po_scene.render(scenes.camera_array(sc["camera"]), W, H, spp, seed=seed, rect=rect, sums=sums, **kw)
    return sums[0].reshape(H, W, 3), sums[1].reshape(H, W, 3)


def _crop(a):
    return np.concatenate([a[y0:y1, x0:x1].reshape(-1, 3) for (x0, y0, x1, y1) in CROPS])


def test_headline_frame_same_seed(headline, scenes):
    sc, dev, ref = headline
    kw = dict(maxdepth=10, rrdepth=3)
    spp = 64
    g = dev.render(sc["camera"], sc["width"], sc["height"], spp, seed=7, **kw)
    s, _ = _oracle_crops(ref, scenes, sc, spp, 7, kw)
    m_g, m_o = _crop(g["sum"]) / spp, _crop(s) / spp
    rel = np.abs(m_g - m_o) / (np.abs(m_o) + 1e-3)
    frac = float((rel > 1e-3).any(-1).mean())
    print(f"headline frame crops, same seed: {frac:.4f} of pixels off, means {m_g.mean():.5f} / {m_o.mean():.5f}")
    assert frac < 0.03
    assert abs(m_g.mean() - m_o.mean()) / m_o.mean() < 3e-3


def test_headline_frame_ztest(headline, scenes):
    sc, dev, ref = headline
    kw = dict(maxdepth=10, rrdepth=3)
    spp = 256
    g = dev.render(sc["camera"], sc["width"], sc["height"], spp, seed=1, **kw)
    s, s2 = _oracle_crops(ref, scenes, sc, spp, 2, kw)
    m1, m2 = _crop(g["sum"]) / spp, _crop(s) / spp
    v1 = np.maximum(_crop(g["sum2"]) / spp - m1 ** 2, 0) * spp / (spp - 1)
    v2 = np.maximum(_crop(s2) / spp - m2 ** 2, 0) * spp / (spp - 1)
    se = np.sqrt(v1 / spp + v2 / spp)
    ok = se > 1e-6
    z = (m1 - m2)[ok] / se[ok]
    frac3 = float((np.abs(z) > 3).mean())
    print(f"headline frame crops, z-test: {frac3:.4f} of channels beyond 3 sigma, mean z {z.mean():.4f}")
    assert frac3 < 0.0054
    assert abs(float(z.mean())) < 0.06


def test_headline_frame_fp64_mre(headline, scenes):
    """The benchmark's frame (64 spp, seed 7) against the FP64 oracle of the same estimator on the
    same random numbers, on the crops: mean |gpu - fp64| / (|fp64| + 1e-2) (the figure the
    benchmark line reports as mre_fp64)."""
    sc, dev, ref = headline
    kw = dict(maxdepth=10, rrdepth=3)
    spp = 64
    g = dev.render(sc["camera"], sc["width"], sc["height"], spp, seed=7, **kw)
    s, _ = _oracle_crops(ref, scenes, sc, spp, 7, dict(kw, fp64=True))
    m_g, m_o = _crop(g["sum"]) / spp, _crop(s) / spp
    mre = float((np.abs(m_g - m_o) / (np.abs(m_o) + 1e-2)).mean())
    print(f"headline frame crops: mean relative error against the FP64 oracle {mre:.2e}")
    assert mre < 5e-3
