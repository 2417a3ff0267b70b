"""Scene changes after the BVH is built (Renderer::setEnvmap at any time, renderer.h:30; new or
resized textures; Blinn-Phong Ks rows): the library re-uploads the material table, the texture
pool and the sky sampler without rebuilding the BVH, and a render afterwards equals, bit for bit,
the render of a scene built with the final data from the start."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _render(dev, sc, **kw):
    return dev.render(sc["camera"], sc["width"], sc["height"], 4, seed=17, maxdepth=8, rrdepth=3, **kw)


@pytest.mark.parametrize("env_is", [False, True])
def test_envmap_set_after_build(nr, scenes, env_is):
    sc = scenes.terrain_spheres(n_cells=32, n_spheres=40, width=96, height=54)
    ref_dev = nr.Scene(sc)
    ref = _render(ref_dev, sc, env_is=env_is)
    bare = dict(sc)
    bare.pop("envmap")
    dev = nr.Scene(bare)
    first = _render(dev, bare, env_is=env_is)               # no sky yet
    dev.set_envmap(sc["envmap"]["rgb"], sc["envmap"].get("gain", 1.0))
    got = _render(dev, sc, env_is=env_is)
    assert not np.array_equal(first["sum"], got["sum"])
    assert np.array_equal(got["sum"], ref["sum"]) and np.array_equal(got["sum2"], ref["sum2"])
    # same size again with other texels and gain: replaced in place (sky sampler rebuilt too)
    sky2 = np.ascontiguousarray(sc["envmap"]["rgb"][:, ::-1])
    dev.set_envmap(sky2, 0.5)
    ref2 = dict(sc, envmap=dict(rgb=sky2, gain=0.5))
    got2 = _render(dev, ref2, env_is=env_is)
    want2 = _render(nr.Scene(ref2), ref2, env_is=env_is)
    assert np.array_equal(got2["sum"], want2["sum"])


def test_texture_and_material_ext_after_build(nr, scenes):
    sc = scenes.material_zoo_ext(64, 48)
    dev = nr.Scene(sc)
    _render(dev, sc)
    # a resized RGBA8 texture 0 and a new Ks row on the Blinn-Phong sphere
    rng = np.random.Generator(np.random.PCG64(3))
    tex0 = rng.integers(0, 256, (32, 48, 4), dtype=np.uint8)
    dev.set_texture_rgba8(0, tex0)
    dev.set_material_ext(10, [0.3, 0.2, 0.1, -1])
    got = _render(dev, sc)
    final = dict(sc, textures=[tex0] + list(sc["textures"][1:]), mats_ext={**sc["mats_ext"], 10: [0.3, 0.2, 0.1, -1]})
    want = _render(nr.Scene(final), final)
    assert np.array_equal(got["sum"], want["sum"]) and np.array_equal(got["sum2"], want["sum2"])
