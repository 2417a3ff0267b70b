"""The command line (nuray_cli, cmd.cpp:9-113) end to end on the GPU: the Cornell box OBJ/MTL
fixture rendered through the C++ facade and saved as BMP equals, byte for byte, the same scene
built from an independent parse and rendered through the Python binding."""
import os
import subprocess

import numpy as np
import pytest

from test_facade import OBJ, ROOT, _parse_obj

pytestmark = pytest.mark.gpu
CLI = os.path.join(ROOT, "nuray_amd", "nuray_cli")


def _read_bmp(path):
    d = open(path, "rb").read()
    off = int.from_bytes(d[10:14], "little")
    w = int.from_bytes(d[18:22], "little")
    h = int.from_bytes(d[22:26], "little", signed=True)
    row = (w * 3 + 3) & ~3
    img = np.zeros((abs(h), w, 3), np.uint8)
    for y in range(abs(h)):
        r = np.frombuffer(d[off + y * row: off + y * row + w * 3], np.uint8).reshape(w, 3)
        img[abs(h) - 1 - y if h > 0 else y] = r[:, ::-1]
    return img


def _scene_from_parse():
    tris, mats = _parse_obj(OBJ, 100.0, np.zeros(3, np.float32))
    names = list(mats.keys())
    rows = []
    for nme in names:
        m = mats[nme]
        if sum(m["Ke"]) > 0:
            rows.append([1, *m["Ke"], 0.5, 1, -1, 0])
        else:
            rows.append([0, *m["Kd"], 0.5, 1, -1, 0])
    return dict(p=np.stack([t[0] for t in tris]).astype(np.float32), n=None, uv=None,
                mat=np.array([names.index(t[1]) for t in tris], np.int32), mats=np.array(rows, np.float32))


def test_cli_cornell_equals_binding(nr, tmp_path):
    out = str(tmp_path / "cornell.bmp")
    W, H, spp = 96, 72, 8
    cmd = [CLI, OBJ, "-s", "100", "--output", out, "--params", "renderer=ptnee", f"imgw={W}", f"imgh={H}",
           f"spp={spp}", "campos=0,100,350", "cameuler=0,0,0", "camfov=40", "camasp=1.3333", "seed=3"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(r.stdout[-400:], r.stderr[-400:])
    assert r.returncode == 0
    got = _read_bmp(out)
    sc = _scene_from_parse()
    dev = nr.Scene(sc)
    ref = dev.render(dict(pos=(0, 100, 350), euler=(0, 0, 0), fov=40.0, aspect=1.3333), W, H, spp, seed=3)["rgb8"]
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)
    assert ref.mean() > 5


def test_cli_nrc_renderer(tmp_path):
    out = str(tmp_path / "cornell_nrc.bmp")
    cmd = [CLI, OBJ, "-s", "100", "--output", out, "--params", "renderer=nrc", "imgw=96", "imgh=72", "frames=16",
           "campos=0,100,350", "camfov=40", "camasp=1.3333", "nrc_train_paths=4096"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(r.stdout[-300:], r.stderr[-300:])
    assert r.returncode == 0
    img = _read_bmp(out)
    assert img.shape == (72, 96, 3) and img.mean() > 5


def test_cli_errors_are_reported(tmp_path):
    r = subprocess.run([CLI, str(tmp_path / "missing.obj"), "--output", str(tmp_path / "x.bmp")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "cannot open" in r.stderr


def test_render_function_backward_vs_central_difference(tmp_path):
    exe = str(tmp_path / "test_autograd")
    src = os.path.join(ROOT, "tests", "cpp", "test_autograd.cpp")
    lib_dir = os.path.join(ROOT, "nuray_amd")
    subprocess.run(["g++", "-std=c++17", "-O1", "-o", exe, src, f"-L{lib_dir}", "-lnuray", "-lz", f"-Wl,-rpath,{lib_dir}"],
                   check=True)
    r = subprocess.run([exe, OBJ], capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-500:])
    assert r.returncode == 0 and "OK" in r.stdout
