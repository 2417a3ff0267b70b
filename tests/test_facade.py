"""The C++ facade (include/nuray.hpp) on the host: Config, Loader and BMP checks compiled with g++,
and the loader's triangles against an independent parse of the Cornell box OBJ/MTL fixture
(tests/golden/cornell, the reference's CornellBox-Original, 36 triangles from 18 quads)."""
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
OBJ = os.path.join(HERE, "golden", "cornell", "CornellBox-Original.obj")


def _parse_obj(path, scale, pos):
    """Independent OBJ/MTL reading: quads (0,1,2), (2,3,0) (loader.cpp:336-354), negative indices
    relative, v * scale + position in float32 (loader.cpp:274)."""
    mats, cur, name = {}, None, None
    d = os.path.dirname(path)
    v, tris, order = [None], [], []
    for line in open(path):
        b = line.split()
        if not b:
            continue
        if b[0] == "mtllib":
            for ml in open(os.path.join(d, b[1])):
                c = ml.split()
                if not c or c[0].startswith("#"):
                    continue
                if c[0] == "newmtl":
                    name = c[1]
                    mats[name] = {"Kd": [0, 0, 0], "Ke": [0, 0, 0], "Ks": [0, 0, 0], "Ns": 10.0}
                elif c[0] in ("Kd", "Ke", "Ks"):
                    mats[name][c[0]] = [float(x) for x in c[1:4]]
                elif c[0] == "Ns":
                    mats[name]["Ns"] = float(c[1])
        elif b[0] == "usemtl":
            cur = b[1]
            if cur not in order:
                order.append(cur)
        elif b[0] == "v":
            v.append(np.float32(np.array([float(x) for x in b[1:4]], np.float32) * np.float32(scale)) + pos)
        elif b[0] == "f":
            idx = [int(t.split("/")[0]) for t in b[1:]]
            idx = [i + len(v) if i < 0 else i for i in idx]
            quads = [(0, 1, 2)] + ([(2, 3, 0)] if len(idx) == 4 else [])
            for q in quads:
                tris.append((np.concatenate([v[idx[k]] for k in q]), cur))
    return tris, mats


def test_facade_host_checks_and_loader(tmp_path):
    exe = str(tmp_path / "test_facade")
    src = os.path.join(HERE, "cpp", "test_facade.cpp")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-o", exe, src, "-lz"], check=True)
    dump = str(tmp_path / "dump.bin")
    mirror = os.path.join(os.path.dirname(OBJ), "CornellBox-Mirror.obj")
    r = subprocess.run([exe, OBJ, dump, str(tmp_path / "t.bmp"), mirror], capture_output=True, text=True)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and r.stdout.startswith("OK")
    raw = open(dump, "rb").read()
    nt, nm = np.frombuffer(raw[:8], np.int32)
    rec = np.frombuffer(raw[8: 8 + nt * 76], np.uint8).reshape(nt, 76)
    p = rec[:, :36].copy().view(np.float32).reshape(nt, 9)
    mat = rec[:, 72:76].copy().view(np.int32).reshape(nt)
    rows = np.frombuffer(raw[8 + nt * 76:], np.float32).reshape(nm, 8)
    tris, mats = _parse_obj(OBJ, 100.0, np.array([1, 2, 3], np.float32))
    assert nt == len(tris) == 36
    assert np.array_equal(p, np.stack([t[0] for t in tris]).astype(np.float32))
    # material rows in MTL order; the light (Ke > 0) becomes an emitter row
    names = list(mats.keys())
    for i, (t, m) in enumerate(tris):
        row = rows[mat[i]]
        k = names.index(m)
        assert mat[i] == k
        if sum(mats[m]["Ke"]) > 0:
            assert row[0] == 1 and np.allclose(row[1:4], mats[m]["Ke"])
        else:
            assert row[0] == 0 and np.allclose(row[1:4], mats[m]["Kd"])
