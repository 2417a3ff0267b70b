"""PNG textures through the C++ facade's loader (include/nuray.hpp, nuray::Texture::load), the
format Texture::load reads through QImage (texture.cpp:5-31) for the reference's Mitsuba scene
(scenes/mitsuba/mitsuba.mtl: map_Kd default.png).  Checked against Pillow's decoder (an
independent implementation) on the scene's own default.png and on images Pillow writes in every
colour type it offers, and against exact pixel values on images written here with each of the five
row filters forced, at 1-, 2-, 4-, 8- and 16-bit depths.  JPEG and interlaced PNG are refused with
a clear error.  Host only."""
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

from conftest import ROOT

GOLDEN_PNG = os.path.join(ROOT, "tests", "golden", "mitsuba", "default.png")


@pytest.fixture(scope="module")
def decoder(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("png") / "texture_decode")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-o", exe,
                    os.path.join(ROOT, "tests", "cpp", "texture_decode.cpp"), "-lz"], check=True)

    def run(path):
        out = path + ".rgba"
        r = subprocess.run([exe, path, out], capture_output=True, text=True)
        if r.returncode != 0:
            return None, r.stdout.strip()
        raw = open(out, "rb").read()
        w, h = struct.unpack("<ii", raw[:8])
        return np.frombuffer(raw[8:], np.uint8).reshape(h, w, 4), None
    return run


def _pil_rgba(path):
    from PIL import Image
    return np.asarray(Image.open(path).convert("RGBA"))


def test_reference_scene_texture(decoder):
    got, err = decoder(GOLDEN_PNG)
    assert err is None, err
    want = _pil_rgba(GOLDEN_PNG)
    assert got.shape == want.shape == (128, 128, 4)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("mode", ["RGB", "RGBA", "L", "LA", "P", "1"])
def test_pillow_written_images(decoder, tmp_path, mode):
    from PIL import Image
    rng = np.random.Generator(np.random.PCG64(3))
    base = np.zeros((37, 53, 3), np.uint8)
    yy, xx = np.mgrid[0:37, 0:53]
    base[..., 0] = (xx * 5) % 256
    base[..., 1] = (yy * 7 + xx) % 256
    base[..., 2] = rng.integers(0, 256, (37, 53))
    img = Image.fromarray(base, "RGB")
    if mode == "RGBA":
        a = Image.fromarray(rng.integers(0, 256, (37, 53)).astype(np.uint8), "L")
        img = img.copy()
        img.putalpha(a)
    elif mode == "P":
        img = img.quantize(colors=40)
    elif mode != "RGB":
        img = img.convert(mode)
    path = str(tmp_path / f"t_{mode}.png")
    img.save(path)
    got, err = decoder(path)
    assert err is None, err
    assert np.array_equal(got, _pil_rgba(path))


def _png(path, w, h, ctype, depth, rows, filters, plte=None, interlace=0):
    """Minimal encoder: rows = list of unfiltered scanline byte strings; each row filtered with the
    given filter type (0 None, 1 Sub, 2 Up, 3 Average, 4 Paeth)."""
    nch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    bpp = max(1, nch * depth // 8)
    out, prev = b"", bytes(len(rows[0]))
    for row, f in zip(rows, filters):
        enc = bytearray()
        for i, x in enumerate(row):
            a = row[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            if f == 0: p = 0
            elif f == 1: p = a
            elif f == 2: p = b
            elif f == 3: p = (a + b) // 2
            else:
                pp = a + b - c
                pa, pb, pc = abs(pp - a), abs(pp - b), abs(pp - c)
                p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
            enc.append((x - p) & 255)
        out += bytes([f]) + bytes(enc)
        prev = row

    def chunk(t, data):
        return struct.pack(">I", len(data)) + t + data + struct.pack(">I", zlib.crc32(t + data) & 0xFFFFFFFF)
    blob = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, interlace))
    if plte is not None:
        blob += chunk(b"PLTE", plte)
    blob += chunk(b"IDAT", zlib.compress(out, 6)) + chunk(b"IEND", b"")
    open(path, "wb").write(blob)


def test_every_filter_and_depth(decoder, tmp_path):
    rng = np.random.Generator(np.random.PCG64(7))
    w, h = 13, 10
    filters = [0, 1, 2, 3, 4, 4, 3, 2, 1, 0]
    # 8-bit RGB
    px = rng.integers(0, 256, (h, w, 3)).astype(np.uint8)
    p = str(tmp_path / "rgb8.png")
    _png(p, w, h, 2, 8, [px[y].tobytes() for y in range(h)], filters)
    got, err = decoder(p)
    assert err is None, err
    assert np.array_equal(got[..., :3], px) and np.all(got[..., 3] == 255)
    # 16-bit RGBA: the high byte of each sample
    px16 = rng.integers(0, 65536, (h, w, 4)).astype(">u2")
    p = str(tmp_path / "rgba16.png")
    _png(p, w, h, 6, 16, [px16[y].tobytes() for y in range(h)], filters)
    got, err = decoder(p)
    assert err is None, err
    assert np.array_equal(got, (px16 >> 8).astype(np.uint8))
    # 1-, 2-, 4-bit greyscale (scaled to 0..255) and a 4-bit palette
    for depth in (1, 2, 4):
        g = rng.integers(0, 1 << depth, (h, w))
        rows = []
        for y in range(h):
            bits = "".join(format(v, f"0{depth}b") for v in g[y])
            bits += "0" * (-len(bits) % 8)
            rows.append(int(bits, 2).to_bytes(len(bits) // 8, "big"))
        p = str(tmp_path / f"grey{depth}.png")
        _png(p, w, h, 0, depth, rows, filters)
        got, err = decoder(p)
        assert err is None, err
        assert np.array_equal(got[..., 0], (g * 255 // ((1 << depth) - 1)).astype(np.uint8)), depth
        assert np.array_equal(got[..., 0], got[..., 2])
        if depth == 4:
            pal = rng.integers(0, 256, (16, 3)).astype(np.uint8)
            p = str(tmp_path / "pal4.png")
            _png(p, w, h, 3, 4, rows, filters, plte=pal.tobytes())
            got, err = decoder(p)
            assert err is None, err
            assert np.array_equal(got[..., :3], pal[g])


def test_refused_formats(decoder, tmp_path):
    p = str(tmp_path / "x.jpg")
    open(p, "wb").write(b"\xff\xd8\xff\xe0" + bytes(64))
    got, err = decoder(p)
    assert got is None and "JPEG" in err
    p = str(tmp_path / "interlaced.png")
    _png(p, 4, 4, 2, 8, [bytes(12)] * 4, [0] * 4, interlace=1)
    got, err = decoder(p)
    assert got is None and "interlaced" in err
    p = str(tmp_path / "truncated.png")
    _png(p, 8, 8, 2, 8, [bytes(24)] * 8, [0] * 8)
    open(p, "r+b").truncate(45)   # inside the IDAT chunk
    got, err = decoder(p)
    assert got is None
