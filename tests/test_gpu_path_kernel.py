"""nr_params.path_kernel: each sample's whole path traced by one lane in one launch per batch
(the forward replay of the adjoint) instead of the wavefront's launch per bounce.  Same Philox
numbers, same traversal, same BSDF code, same order of additions: the image must equal the
wavefront's bit for bit, for every material model, both integrators and split batches; the
adjoint reuses its per-sample radiance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _both(nr, sc, spp, **kw):
    dev = nr.Scene(sc)
    a = dev.render(sc["camera"], sc["width"], sc["height"], spp, seed=23, **kw)
    b = dev.render(sc["camera"], sc["width"], sc["height"], spp, seed=23, path_kernel=True, **kw)
    dev.close()
    return a, b


@pytest.mark.parametrize("case", ["box_reference", "zoo", "zoo_ext_pt", "terrain_batches"])
def test_path_kernel_equals_wavefront(nr, scenes, case):
    if case == "box_reference":
        sc, kw = scenes.box_room(), dict(maxdepth=0, rrdepth=0)
    elif case == "zoo":
        sc, kw = scenes.material_zoo(96, 72), dict(maxdepth=8, rrdepth=3)
    elif case == "zoo_ext_pt":
        sc, kw = scenes.material_zoo_ext(96, 72), dict(maxdepth=8, rrdepth=3, integrator=1)
    else:
        sc, kw = scenes.terrain_spheres(n_cells=48, n_spheres=60, width=128, height=72), dict(maxdepth=10, rrdepth=3,
                                                                                             batch_spp=3)
    a, b = _both(nr, sc, 8, **kw)
    diff = np.abs(a["sum"] - b["sum"]).max()
    print(f"{case}: max |sum difference| {diff:.3e}, image mean {a['sum'].mean() / 8:.5f}")
    assert np.array_equal(a["sum"], b["sum"]) and np.array_equal(a["sum2"], b["sum2"])
    assert np.array_equal(a["rgb8"], b["rgb8"])


@pytest.mark.parametrize("maxdepth,rrdepth,integrator", [(0, 0, 0), (4, 2, 0), (6, 3, 1)])
def test_adjoint_after_path_kernel_render(nr, scenes, maxdepth, rrdepth, integrator):
    """The replay reads the vertex records of the render it follows (the hit of each of the first
    16 vertices and its NEE shadow-ray outcome) -- written by the wavefront's shade and shadow
    launches, or by the path kernel itself -- and traces past the 16th vertex (without a depth
    limit some paths are longer).  Both renders give the same gradient, with and without a depth
    limit and for the pt integrator (no NEE)."""
    sc, theta_star, theta0 = scenes.adjoint_problem(width=48, height=48)
    W, H, M = 48, 48, len(theta0)
    kw = dict(maxdepth=maxdepth, rrdepth=rrdepth, integrator=integrator)
    dldi = np.random.Generator(np.random.PCG64(4)).normal(size=(H, W, 3)).astype(np.float32) * 1e-2
    dev = nr.Scene(sc)
    dev.set_params(theta0)
    dev.render(sc["camera"], W, H, 8, seed=5, fetch=False, **kw)
    g1 = dev.adjoint(sc["camera"], W, H, 8, M, seed=5, dl_dimg=dldi, **kw)
    dev.render(sc["camera"], W, H, 8, seed=5, fetch=False, path_kernel=True, **kw)
    g2 = dev.adjoint(sc["camera"], W, H, 8, M, seed=5, dl_dimg=dldi, **kw)
    assert np.linalg.norm(g1) > 0
    # the replay reads the same stored radiance; the gradient rows are summed with float atomics whose
    # order changes from run to run, so the two agree to rounding, not bit for bit
    err = np.linalg.norm(g1 - g2) / np.linalg.norm(g1)
    print(f"adjoint after wavefront vs after path kernel render: relative L2 difference {err:.2e}")
    assert err <= 1e-6


def test_path_kernel_refusals(nr, scenes):
    sc = scenes.terrain_spheres(n_cells=16, n_spheres=10, width=32, height=18)
    dev = nr.Scene(sc)
    with pytest.raises(nr.NuRayError, match="path kernel"):
        dev.render(sc["camera"], 32, 18, 2, env_is=True, path_kernel=True)
    dev.close()


def test_path_kernel_auto(nr, scenes):
    """path_kernel = 2 picks one lane per path (no wavefront shade launches) for a scene with a flat
    leaf list and for a small render of a larger scene, the wavefront for a render whose samples x
    triangles exceed 2e12 and with sky sampling; the image is the same either way."""
    small = scenes.box_room()
    dev = nr.Scene(small)
    dev.set_profiling(1)
    a = dev.render(small["camera"], 64, 48, 4, seed=3, path_kernel="auto")
    assert dev.stats()["shade_launches"] == 0
    b = dev.render(small["camera"], 64, 48, 4, seed=3)
    assert dev.stats()["shade_launches"] > 0
    assert np.array_equal(a["sum"], b["sum"]) and np.array_equal(a["sum2"], b["sum2"])
    dev.close()
    mid = scenes.terrain_spheres(n_cells=48, n_spheres=60, width=64, height=36)
    dev = nr.Scene(mid)
    dev.set_profiling(1)
    dev.render(mid["camera"], 64, 36, 2, seed=3, path_kernel="auto")
    assert dev.stats()["shade_launches"] == 0
    dev.render(mid["camera"], 64, 36, 2, seed=3, path_kernel="auto", env_is=True)
    assert dev.stats()["shade_launches"] > 0
    dev.close()
    big = scenes.terrain_spheres()  # 684,296 triangles: 1280 x 720 x 4 samples is past 2e12
    dev = nr.Scene(big)
    dev.set_profiling(1)
    dev.render(big["camera"], 1280, 720, 4, seed=3, maxdepth=4, rrdepth=2, path_kernel="auto", fetch=False)
    assert dev.stats()["shade_launches"] > 0
    dev.close()
