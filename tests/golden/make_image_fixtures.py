"""Writes the texture-decoder fixtures: small synthetic images in the formats
the reference's loader reads (stb_image, libopenntc.cpp:1302), encoded by
Pillow, and stb_image's own 8-bit RGBA decode of each (oracle/_ref/libstbi.so,
compiled from the reference's stb_image.h by `make -C oracle ref`) as the
expected output.

Run with an interpreter that has Pillow, from the repository root:
    /opt/conda/bin/python tests/golden/make_image_fixtures.py
Output: tests/golden/images/*.{jpg,png,bmp,tga,ppm,pgm} and stb_rgba.npz.
"""
import ctypes as C
import os

import numpy as np
from PIL import Image

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.join(ROOT, "tests", "golden", "images")


def source(w, h, seed):
    """Smooth gradients, a sharp diagonal edge, a disc and mild noise."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    r = 255 * x / max(w - 1, 1)
    g = 255 * y / max(h - 1, 1)
    b = np.where(x + 0.7 * y < 0.6 * w, 40.0, 220.0)
    disc = (x - 0.6 * w) ** 2 + (y - 0.4 * h) ** 2 < (0.2 * min(w, h)) ** 2
    g = np.where(disc, 255 - g, g)
    a = 255 * (0.5 + 0.5 * np.sin(x / 5.0) * np.cos(y / 7.0))
    img = np.stack([r, g, b, a], -1) + rng.normal(0, 6, (h, w, 4))
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def main():
    os.makedirs(OUT, exist_ok=True)
    rgba = source(48, 40, 1)
    odd = source(37, 23, 2)
    big = source(72, 56, 3)
    rgb = Image.fromarray(rgba[..., :3])
    files = {}

    def save(name, im, **kw):
        path = os.path.join(OUT, name)
        im.save(path, **kw)
        files[name] = path

    # JPEG: chroma subsampling 4:4:4 / 4:2:2 / 4:2:0, grey, progressive,
    # optimised Huffman tables, restart markers, sizes off the MCU grid
    save("q90_444.jpg", rgb, quality=90, subsampling=0)
    save("q90_422.jpg", rgb, quality=90, subsampling=1)
    save("q90_420.jpg", rgb, quality=90, subsampling=2)
    save("q60_420_opt.jpg", rgb, quality=60, subsampling=2, optimize=True)
    save("grey.jpg", rgb.convert("L"), quality=85)
    save("prog_420.jpg", rgb, quality=85, subsampling=2, progressive=True)
    save("prog_444.jpg", rgb, quality=95, subsampling=0, progressive=True)
    save("prog_grey.jpg", rgb.convert("L"), quality=80, progressive=True)
    save("odd_420.jpg", Image.fromarray(odd[..., :3]), quality=90, subsampling=2)
    save("odd_prog_422.jpg", Image.fromarray(odd[..., :3]), quality=90, subsampling=1, progressive=True)
    save("restart_420.jpg", rgb, quality=90, subsampling=2, restart_marker_blocks=3)
    save("big_prog_420.jpg", Image.fromarray(big[..., :3]), quality=75, subsampling=2, progressive=True)
    save("big_q100_444.jpg", Image.fromarray(big[..., :3]), quality=100, subsampling=0)
    # BMP: 24-bit, 32-bit with alpha, 8-bit palette, 8-bit grey, 1-bit
    save("rgb24.bmp", rgb)
    save("rgba32.bmp", Image.fromarray(rgba))
    save("pal8.bmp", rgb.quantize(64))
    save("grey8.bmp", rgb.convert("L"))
    save("mono1.bmp", rgb.convert("1"))
    save("odd_rgb24.bmp", Image.fromarray(odd[..., :3]))
    # PNG, TGA, PNM as Pillow writes them
    save("rgb.png", rgb)
    save("rgba.png", Image.fromarray(rgba))
    save("grey.png", rgb.convert("L"))
    save("grey_alpha.png", Image.fromarray(rgba).convert("LA"))
    save("pal.png", rgb.quantize(32))
    save("grey16.png", Image.fromarray((rgba[..., 0].astype(np.uint16) * 257 + 3)))
    save("rgb.tga", rgb)
    save("rgba_rle.tga", Image.fromarray(rgba), compression="tga_rle")
    save("grey.tga", rgb.convert("L"))
    save("rgb.ppm", rgb)
    save("grey.pgm", rgb.convert("L"))

    stb = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libstbi.so"))
    stb.stbi_load.restype = C.POINTER(C.c_ubyte)
    stb.stbi_load.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int]
    stb.stbi_image_free.argtypes = [C.c_void_p]
    expected = {}
    for name, path in sorted(files.items()):
        w, h, n = C.c_int(), C.c_int(), C.c_int()
        p = stb.stbi_load(path.encode(), C.byref(w), C.byref(h), C.byref(n), 4)
        assert p, name
        expected[name] = np.ctypeslib.as_array(p, shape=(h.value, w.value, 4)).copy()
        stb.stbi_image_free(p)
    np.savez_compressed(os.path.join(OUT, "stb_rgba.npz"), **expected)
    print(f"{len(files)} images, {sum(os.path.getsize(p) for p in files.values())} bytes")


if __name__ == "__main__":
    main()
