"""Radiance-cache rendering on the GPU: with the cache switched off the frame is exactly the
ptnee render (same Philox numbers, same sums), training paths leave the render slots untouched,
and a trained cache lowers the error of a few-sample image."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cam(sc):
    return sc["camera"]


@pytest.mark.parametrize("which", ["box", "terrain"])
def test_cache_off_equals_ptnee(nr, scenes, which):
    if which == "box":
        sc = scenes.box_room()
        kw = dict(maxdepth=0, rrdepth=0)
    else:
        sc = scenes.terrain_spheres(n_cells=64, n_spheres=100, width=160, height=90)
        kw = dict(maxdepth=10, rrdepth=3)
    W, H = sc["width"], sc["height"]
    dev = nr.Scene(sc)
    ref = dev.render(_cam(sc), W, H, 4, seed=21, **kw)
    cache = nr.Nrc(seed=2)
    st, got = dev.render_nrc(cache, _cam(sc), W, H, spp=4, seed=21, use_cache=False, train=True, fetch=True, **kw)
    assert np.array_equal(got["sum"], ref["sum"]) and np.array_equal(got["sum2"], ref["sum2"])
    # training vertices are recorded (tail queries need 5-vertex paths, rare under an open sky)
    assert st["records"] > 0 and st["trained"] == min(st["records"], 1 << 16)
    assert st["queries"] > 0 or which == "terrain"
    assert np.isfinite(st["loss"])
    st2, got2 = dev.render_nrc(cache, _cam(sc), W, H, spp=4, seed=21, use_cache=False, train=False,
                               train_paths=0, fetch=True, **kw)
    assert np.array_equal(got2["sum"], ref["sum"]) and st2["queries"] == 0 and st2["trained"] == 0
