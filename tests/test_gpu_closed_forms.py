"""The product on the GPU against closed forms that do not pass through the oracle:

* the sky seen by a ray that leaves the scene: 1-pixel cameras with a 0.01 degree field aimed at
  chosen points of a 4 x 2 sky return the bilinear blend of the texels there, times the gain
  (rendererptls.cpp:16-28, texture.cpp:61-72, envmap.cpp:17-24), for both render modes;
* the default MIS weight (mis = 1, rendererptls.cpp:61-69,89): a Lambert plane under a 10 x 10
  emitter returns the quadrature of the reference's rule (weight of the emitter hit taken from the
  NEE direction), for both render modes, and not the textbook balance heuristic's value."""
import math

import numpy as np
import pytest

from test_oracle_pins_bsdf_env_mis import (_aim, _bilinear_black, _env_scene, _env_texels, _light_plane,
                                           _mis_expectations, _rel)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("path_kernel", [False, True])
def test_gpu_miss_radiance_closed_form(nr, path_kernel):
    t = _env_texels()
    H, W = t.shape[:2]
    gain = 1.25
    dev = nr.Scene(_env_scene(t, gain))
    pts = [(0.25, 0.4), (1.6, 0.7), (2.3, 1.2), (3.25, 1.3), (3.5, 0.25), (0.7, 1.6)]
    try:
        for x, y in pts:
            d = _aim(x, y, W, H)
            pitch = math.degrees(math.asin(d[1]))
            yaw = math.degrees(math.atan2(-d[0], -d[2]))
            cam = dict(pos=(0.0, 0.0, 0.0), euler=(yaw, pitch, 0.0), fov=0.01, aspect=1.0)
            out = dev.render(cam, 1, 1, 64, seed=3, path_kernel=path_kernel)
            got = out["sum"].reshape(3) / 64
            want = _bilinear_black(t, x, y) * gain
            print(f"path_kernel={path_kernel} point ({x}, {y}): {got} expected {want}")
            assert np.all(_rel(got, want) < 2e-3)
    finally:
        dev.close()


@pytest.mark.parametrize("path_kernel", [False, True])
def test_gpu_default_mis_weight_closed_form(nr, path_kernel):
    rho, Le, a, h = 0.5, 4.0, 10.0, 4.0
    dev = nr.Scene(_light_plane(rho, Le, a, h))
    spp = 1 << 21
    cam = dict(pos=(0.0, h / 2, 0.0), euler=(0.0, -90.0, 0.0), fov=0.01, aspect=1.0)
    out = dev.render(cam, 1, 1, spp, seed=13, mis=1, maxdepth=2, path_kernel=path_kernel)
    dev.close()
    m = float(out["sum"].reshape(-1)[0]) / spp
    se = math.sqrt(max(float(out["sum2"].reshape(-1)[0]) / spp - m * m, 0) / spp)
    e_nee, e_hit, e_text = _mis_expectations(rho, Le, a, h)
    expect, textbook = e_nee + e_hit, e_nee + e_text
    print(f"path_kernel={path_kernel}: {m:.6f} expected {expect:.6f} (se {se:.2e}); textbook rule {textbook:.6f}")
    assert abs(m - expect) < 5 * se + 2e-4 * expect
    assert abs(m - textbook) > 20 * se
