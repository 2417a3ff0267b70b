"""The `pt` integrator (rendererpt.cpp: no light sampling, emission counted on every hit) and the
Blinn-Phong and glass materials (matblinnphong.cpp, matglass.cpp) on the GPU, against the oracle.

* same seed: per-pixel means agree to 1e-3 relative on almost every pixel, for both integrators;
* independent seeds at 256 spp: per-pixel 3-sigma z-test against the oracle;
* closed forms on a Lambert plane under a square emitter: pt returns the direct light with the
  Lambert pdf's 1/0.99, ptnee (MIS weight fixed to 1/2) the reference's 1.5x-weighted direct light;
* the adjoint replays paths through Blinn-Phong and glass (those rows stay zero, as in the oracle)
  and with the pt integrator, against the FP64 oracle;
* the command line renders the Cornell box with a Blinn-Phong tall box (CornellBox-Mirror.mtl) with
  both integrators, byte-equal to the binding on an independently parsed scene."""
import os
import subprocess

import numpy as np
import pytest

from test_facade import ROOT, _parse_obj

pytestmark = pytest.mark.gpu
MIRROR = os.path.join(ROOT, "tests", "golden", "cornell", "CornellBox-Mirror.obj")
CLI = os.path.join(ROOT, "nuray_amd", "nuray_cli")


def _gpu(nr, sc, spp, seed, **kw):
    dev = nr.Scene(sc)
    out = dev.render(sc["camera"], sc["width"], sc["height"], spp, seed=seed, **kw)
    dev.close()
    return out


def _oracle(po, scenes, sc, spp, seed, **kw):
    s, s2 = po.Scene(sc).render(scenes.camera_array(sc["camera"]), sc["width"], sc["height"], spp, seed=seed, **kw)
    H, W = sc["height"], sc["width"]
    return s.reshape(H, W, 3), s2.reshape(H, W, 3)


def _z(m1, v1, n1, m2, v2, n2):
    se = np.sqrt(v1 / n1 + v2 / n2)
    ok = se > 1e-6
    return (m1 - m2)[ok] / se[ok]


def _var(s, s2, n):
    m = s / n
    return m, np.maximum(s2 / n - m ** 2, 0) * n / (n - 1)


@pytest.mark.parametrize("integrator", [0, 1])
def test_blinnphong_glass_same_seed(nr, po, scenes, integrator):
    sc = scenes.material_zoo_ext()
    spp = 8
    g = _gpu(nr, sc, spp, seed=41, integrator=integrator)
    s, _ = _oracle(po, scenes, sc, spp, seed=41, integrator=integrator)
    m_g, m_o = g["sum"] / spp, s / spp
    rel = np.abs(m_g - m_o) / (np.abs(m_o) + 1e-3)
    frac = float((rel > 1e-3).any(-1).mean())
    print(f"zoo-ext same-seed integrator {integrator}: frac pixels off {frac:.4f}, "
          f"mean gpu {m_g.mean():.5f} oracle {m_o.mean():.5f}")
    assert frac < 0.03
    assert abs(m_g.mean() - m_o.mean()) / m_o.mean() < 3e-3


@pytest.mark.parametrize("integrator", [0, 1])
def test_blinnphong_glass_ztest_independent_seeds(nr, po, scenes, integrator):
    sc = scenes.material_zoo_ext()
    spp = 256
    g = _gpu(nr, sc, spp, seed=1, integrator=integrator)
    s, s2 = _oracle(po, scenes, sc, spp, seed=2, integrator=integrator)
    m1, v1 = _var(g["sum"], g["sum2"], spp)
    m2, v2 = _var(s, s2, spp)
    z = _z(m1, v1, spp, m2, v2, spp)
    frac3 = float((np.abs(z) > 3).mean())
    print(f"zoo-ext z-test integrator {integrator}: frac |z|>3 {frac3:.4f}, mean z {z.mean():.4f}")
    assert frac3 < 0.0054
    assert abs(float(z.mean())) < 0.05


@pytest.mark.parametrize("integrator,factor", [(1, 1 / 0.99), (0, 0.5 * 2.0 + 0.5 / 0.99)])
def test_direct_light_closed_forms(nr, integrator, factor):
    """A Lambert plane under a square emitter: pt returns rho/(0.99 pi) E_direct (BSDF-sampled
    hits alone, Lambert pdf's 1/0.99), ptnee with the MIS weight fixed to 1/2 returns
    (2 x 1/2 + 1/(0.99) x 1/2) rho/pi E_direct (the light sampler's pdf is half the area pdf)."""
    from test_oracle_pins import light_over_plane

    rho = 0.5
    sc, cam, Ed = light_over_plane(rho, a=10.0, h=5.0)
    spp = 1 << 20
    dev = nr.Scene(sc)
    c = dict(pos=tuple(cam[:3]), euler=tuple(cam[3:6]), fov=float(cam[6]), aspect=float(cam[7]))
    out = dev.render(c, 1, 1, spp, seed=11, mis=0, maxdepth=2 if integrator == 0 else 0, integrator=integrator)
    dev.close()
    m = float(out["sum"].reshape(-1)[0]) / spp
    var = float(out["sum2"].reshape(-1)[0]) / spp - m * m
    se = np.sqrt(var / spp)
    expect = factor * rho / 3.14159 * Ed
    print(f"integrator {integrator}: {m:.6f} expected {expect:.6f} (se {se:.2e})")
    assert abs(m - expect) < 5 * se + 1e-3 * expect


@pytest.mark.parametrize("integrator", [0, 1])
def test_adjoint_through_blinnphong_and_glass(nr, po, scenes, integrator):
    sc = scenes.material_zoo_ext(48, 36)
    W, H, spp = 48, 36, 8
    M = len(sc["mats"])
    dldi = np.random.Generator(np.random.PCG64(9)).normal(size=(H, W, 3)).astype(np.float32) * 1e-2
    dev = nr.Scene(sc)
    dev.render(sc["camera"], W, H, spp, seed=29, fetch=False, integrator=integrator)
    g, gt = dev.adjoint(sc["camera"], W, H, spp, M, seed=29, dl_dimg=dldi, tex_id=0, tex_shape=(64, 64),
                        integrator=integrator)
    go, gto = po.Scene(sc).adjoint(scenes.camera_array(sc["camera"]), W, H, spp, dldi, M, seed=29, tex_id=0,
                                   tex_shape=(64, 64), integrator=integrator)
    em = np.linalg.norm(g - go) / np.linalg.norm(go)
    et = np.linalg.norm(gt - gto) / np.linalg.norm(gto)
    print(f"zoo-ext adjoint integrator {integrator}: material rows {em:.2e}, floor texels {et:.2e}")
    assert em < 5e-3 and et < 2e-2
    assert not g[[1, 9, 10]].any() and not go[[1, 9, 10]].any()   # Blinn-Phong and glass rows
    assert np.abs(g[[0, 2, 3, 4, 8]]).sum() > 0


def _mirror_scene():
    tris, mats = _parse_obj(MIRROR, 100.0, np.zeros(3, np.float32))
    names = list(mats.keys())
    rows, ext = [], {}
    for j, nme in enumerate(names):
        m = mats[nme]
        if sum(m["Ke"]) > 0:
            rows.append([1, *m["Ke"], 0.5, 1, -1, 0])
        elif sum(m["Ks"]) > 0:
            rows.append([4, *m["Kd"], m["Ns"], 1, -1, 0])
            ext[j] = [*m["Ks"], -1]
        else:
            rows.append([0, *m["Kd"], 0.5, 1, -1, 0])
    return dict(p=np.stack([t[0] for t in tris]).astype(np.float32), n=None, uv=None,
                mat=np.array([names.index(t[1]) for t in tris], np.int32), mats=np.array(rows, np.float32),
                mats_ext=ext)


@pytest.mark.parametrize("renderer,integrator", [("ptnee", 0), ("pt", 1)])
def test_cli_cornell_mirror_equals_binding(nr, tmp_path, renderer, integrator):
    from test_gpu_cli import _read_bmp

    out = str(tmp_path / f"mirror_{renderer}.bmp")
    W, H, spp = 96, 72, 8
    cmd = [CLI, MIRROR, "-s", "100", "--output", out, "--params", f"renderer={renderer}", f"imgw={W}",
           f"imgh={H}", f"spp={spp}", "campos=0,100,350", "cameuler=0,0,0", "camfov=40", "camasp=1.3333", "seed=5"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(r.stdout[-400:], r.stderr[-400:])
    assert r.returncode == 0
    got = _read_bmp(out)
    sc = _mirror_scene()
    assert (sc["mats"][:, 0] == 4).sum() == 1   # the tall box
    dev = nr.Scene(sc)
    ref = dev.render(dict(pos=(0, 100, 350), euler=(0, 0, 0), fov=40.0, aspect=1.3333), W, H, spp, seed=5,
                     integrator=integrator)["rgb8"]
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)
    assert ref.mean() > 5
