"""Scenes read by the C++ facade's Loader (include/nuray.hpp: scene description, OBJ / MTL, PNG
textures) and rendered by its Renderer on the GPU, against the oracle rendering the very triangles,
material rows, textures and sky the loader produced (tests/cpp/loader_render.cpp writes them):

* the reference's Cornell box (tests/golden/cornell/CornellBox-Original: 36 triangles, Lambert and
  an area light) with ptnee, and CornellBox-Mirror (a Blinn-Phong tall box) with pt;
* the reference's Mitsuba material-preview object (tests/golden/mitsuba/mitsuba-back.obj: 4,448
  faces, a Lambert inside and a Blinn-Phong backdrop whose map_Kd is the PNG default.png) under a
  sky with a bright patch.

For each: the same seed gives the same per-pixel means on almost every pixel (Philox streams keyed
by pixel and sample), and independent seeds pass a per-pixel 3-sigma z-test at 256 spp.  The
reference's defaults are used (prr 0.8, mis 1, no depth limit, Russian roulette at every vertex)."""
import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(ROOT, "tests", "golden")
CASES = {
    "cornell_ptnee": (os.path.join(GOLDEN, "cornell", "CornellBox-Original.obj") + " -s 100", "ptnee",
                      (0, 100, 350, 0, 0, 0, 40, 96 / 72), 96, 72, False),
    "cornell_mirror_pt": (os.path.join(GOLDEN, "cornell", "CornellBox-Mirror.obj") + " -s 100", "pt",
                          (0, 100, 350, 0, 0, 0, 40, 96 / 72), 96, 72, False),
    "mitsuba_ptnee": (os.path.join(GOLDEN, "mitsuba", "mitsuba-back.obj"), "ptnee",
                      (0.4, 4.0, 13.0, 0, -12, 0, 40, 1.0), 80, 80, True),
}


@pytest.fixture(scope="module")
def loader_render(tmp_path_factory):
    import nuray_amd

    exe = str(tmp_path_factory.mktemp("loader") / "loader_render")
    lib_dir = os.path.dirname(nuray_amd.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-o", exe, os.path.join(ROOT, "tests", "cpp", "loader_render.cpp"),
                    f"-L{lib_dir}", "-lnuray", "-lz", f"-Wl,-rpath,{lib_dir}"], check=True, timeout=300)
    return exe


def _run(exe, tmp_path, case, spp, seed):
    desc, renderer, cam, W, H, sky = CASES[case]
    out = str(tmp_path / f"{case}_{seed}.bin")
    cfg = f"renderer={renderer} imgw={W} imgh={H} spp={spp} seed={seed}"
    args = [exe, desc, out, cfg, ",".join(str(float(c)) for c in cam)] + (["sky"] if sky else [])
    print(f"{case}: loader_render {cfg}", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    print(r.stdout.strip(), r.stderr[-500:])
    assert r.returncode == 0
    raw = open(out, "rb").read()
    nt, nm, ntex, W, H, has_sky = struct.unpack("<6i", raw[:24])
    o = 24
    rec = np.frombuffer(raw[o:o + nt * 100], np.uint8).reshape(nt, 100)
    o += nt * 100
    fl = rec[:, :96].copy().view(np.float32).reshape(nt, 24)
    sc = dict(p=fl[:, :9].copy(), n=fl[:, 9:18].copy(), uv=fl[:, 18:24].copy(),
              mat=rec[:, 96:100].copy().view(np.int32).reshape(nt))
    rows = np.frombuffer(raw[o:o + nm * 48], np.float32).reshape(nm, 12)
    o += nm * 48
    sc["mats"] = rows[:, :8].copy()
    sc["mats_ext"] = {j: rows[j, 8:12].copy() for j in range(nm) if rows[j, 0] == 4}
    texs = []
    for _ in range(ntex):
        w, h = struct.unpack("<2i", raw[o:o + 8])
        o += 8
        texs.append(np.frombuffer(raw[o:o + w * h * 4], np.uint8).reshape(h, w, 4).copy())
        o += w * h * 4
    sc["textures"] = texs
    if has_sky:
        sw, sh = struct.unpack("<2i", raw[o:o + 8])
        gain = struct.unpack("<f", raw[o + 8:o + 12])[0]
        o += 12
        sc["envmap"] = dict(rgb=np.frombuffer(raw[o:o + sw * sh * 12], np.float32).reshape(sh, sw, 3).copy(), gain=gain)
        o += sw * sh * 12
    n3 = W * H * 3
    s = np.frombuffer(raw[o:o + n3 * 8], np.float64).reshape(H, W, 3)
    s2 = np.frombuffer(raw[o + n3 * 8:o + 2 * n3 * 8], np.float64).reshape(H, W, 3)
    return sc, np.asarray(cam, np.float32), W, H, s, s2, renderer


def _oracle(po, sc, cam, W, H, spp, seed, renderer):
    s, s2 = po.Scene(sc).render(cam, W, H, spp, seed=seed, integrator=1 if renderer == "pt" else 0)
    return s.reshape(H, W, 3), s2.reshape(H, W, 3)


@pytest.mark.parametrize("case", list(CASES))
def test_loader_scene_same_seed(loader_render, po, tmp_path, case):
    spp = 16
    sc, cam, W, H, g, _, renderer = _run(loader_render, tmp_path, case, spp, 5)
    if case.startswith("mitsuba"):
        assert len(sc["p"]) == 4448 and len(sc["textures"]) == 1 and sc["textures"][0].shape == (128, 128, 4)
        assert (sc["mats"][:, 6] >= 0).sum() == 1
    s, _ = _oracle(po, sc, cam, W, H, spp, 5, renderer)
    m_g, m_o = g / spp, s / spp
    rel = np.abs(m_g - m_o) / (np.abs(m_o) + 1e-3)
    frac = float((rel > 1e-3).any(-1).mean())
    print(f"{case} same seed: {frac:.4f} of pixels off, means {m_g.mean():.5f} / {m_o.mean():.5f}")
    assert m_o.mean() > 1e-3
    assert frac < 0.03
    assert abs(m_g.mean() - m_o.mean()) <= 3e-3 * m_o.mean()


@pytest.mark.parametrize("case", list(CASES))
def test_loader_scene_ztest(loader_render, po, tmp_path, case):
    spp = 256
    sc, cam, W, H, g, g2, renderer = _run(loader_render, tmp_path, case, spp, 1)
    s, s2 = _oracle(po, sc, cam, W, H, spp, 2, renderer)
    m1, m2 = g / spp, s / spp
    v1 = np.maximum(g2 / spp - m1 ** 2, 0) * spp / (spp - 1)
    v2 = np.maximum(s2 / spp - m2 ** 2, 0) * spp / (spp - 1)
    se = np.sqrt(v1 / spp + v2 / spp)
    ok = se > 1e-6
    z = (m1 - m2)[ok] / se[ok]
    frac3 = float((np.abs(z) > 3).mean())
    print(f"{case} z-test: {frac3:.4f} of channels beyond 3 sigma, mean z {z.mean():.4f} over {ok.sum()} channels")
    assert ok.sum() > 0.5 * ok.size
    assert frac3 < 0.0054
    assert abs(float(z.mean())) < 0.06
