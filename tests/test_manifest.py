"""Texture manifest front end (GPU-free parts of Context::LoadManifest,
libopenntc.cpp:1103-1207 and 1296-1313): image decoding to RGBA8 against
images written here from the formats' specifications (PNG, ISO/IEC 15948:
every colour type and bit depth, all five row filters, tRNS, Adam7; TGA;
binary PNM) and against the reference's own loader, stb_image, on images
Pillow wrote (JPEG sequential and progressive, BMP, PNG, TGA, PNM:
tests/golden/images, made by tests/golden/make_image_fixtures.py), and the
manifest rules with the reference's error codes."""
import ctypes as C
import json
import os
import struct
import zlib

import numpy as np
import pytest

import openntc_amd as ntc
from openntc_amd import _lib as L

SUCCESS, INVALID_DIMENSION, INVALID_MANIFEST, FILE_NOT_FOUND = 0, 1, 3, 4


# ---- writers ------------------------------------------------------------------------
def _chunk(kind, data):
    return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)


def _pack_row(samples, depth):
    """One scanline of samples at a bit depth (MSB first below 8 bits)."""
    if depth == 8:
        return bytes(int(v) for v in samples)
    if depth == 16:
        return b"".join(struct.pack(">H", int(v)) for v in samples)
    out, acc, nbits = bytearray(), 0, 0
    for v in samples:
        acc = (acc << depth) | int(v)
        nbits += depth
        if nbits == 8:
            out.append(acc)
            acc, nbits = 0, 0
    if nbits:
        out.append(acc << (8 - nbits))
    return bytes(out)


def _filter_rows(rows, bpp, rng):
    """Each row with a random filter type 0-4 (the encoder side of the PNG
    filters: the decoder must undo every one)."""
    out, prev = bytearray(), bytes(len(rows[0]))
    for r in rows:
        ft = int(rng.integers(0, 5))
        f = bytearray(len(r))
        for i in range(len(r)):
            a = r[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            if ft == 0:
                pred = 0
            elif ft == 1:
                pred = a
            elif ft == 2:
                pred = b
            elif ft == 3:
                pred = (a + b) >> 1
            else:
                p = a + b - c
                pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
                pred = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
            f[i] = (r[i] - pred) & 0xFF
        out += bytes([ft]) + f
        prev = r
    return bytes(out)


ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]


def write_png(path, samples, ctype, depth, plte=None, trns=None, interlace=0, seed=0):
    """samples: (h, w, channels) integer array at `depth`."""
    h, w = samples.shape[:2]
    nc = samples.shape[2]
    bpp = max(1, nc * depth // 8)
    rng = np.random.default_rng(seed)
    if interlace == 0:
        rows = [_pack_row(samples[y].reshape(-1), depth) for y in range(h)]
        raw = _filter_rows(rows, bpp, rng)
    else:
        raw = b""
        for x0, y0, dx, dy in ADAM7:
            sub = samples[y0::dy, x0::dx]
            if sub.shape[0] == 0 or sub.shape[1] == 0:
                continue
            rows = [_pack_row(sub[y].reshape(-1), depth) for y in range(sub.shape[0])]
            raw += _filter_rows(rows, bpp, rng)
    data = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, interlace))
    if plte is not None:
        data += _chunk(b"PLTE", bytes(np.asarray(plte, np.uint8).reshape(-1)))
    if trns is not None:
        data += _chunk(b"tRNS", bytes(trns))
    z = zlib.compress(raw, 6)
    data += _chunk(b"IDAT", z[: len(z) // 2]) + _chunk(b"IDAT", z[len(z) // 2:]) + _chunk(b"IEND", b"")
    path.write_bytes(data)


def write_tga(path, rgba, grey=False, rle=False, top_left=True):
    h, w = rgba.shape[:2]
    img = rgba if top_left else rgba[::-1]
    if grey:
        px = img[..., 0].reshape(-1, 1)
    else:
        px = img[..., [2, 1, 0, 3]].reshape(-1, 4)
    bypp = px.shape[1]
    body = bytearray()
    if not rle:
        body = px.tobytes()
    else:
        i, n = 0, len(px)
        while i < n:
            j = i
            while j + 1 < n and j - i < 127 and (px[j + 1] == px[i]).all():
                j += 1
            if j > i:  # run
                body += bytes([0x80 | (j - i)]) + px[i].tobytes()
                i = j + 1
            else:
                k = i
                while k + 1 < n and k - i < 127 and not (px[k + 1] == px[k]).all():
                    k += 1
                body += bytes([k - i]) + px[i:k + 1].tobytes()
                i = k + 1
    typ = (11 if rle else 3) if grey else (10 if rle else 2)
    hdr = struct.pack("<BBBHHBHHHHBB", 0, 0, typ, 0, 0, 0, 0, 0, w, h, 8 * bypp, 0x20 if top_left else 0)
    path.write_bytes(hdr + bytes(body))


def decode(path):
    lib = ntc.load()
    w, h = C.c_int32(), C.c_int32()
    st = lib.ntc_image_decode_rgba8(str(path).encode(), C.byref(w), C.byref(h), None, 0)
    if st:
        return st, None
    out = np.zeros((h.value, w.value, 4), np.uint8)
    assert lib.ntc_image_decode_rgba8(str(path).encode(), C.byref(w), C.byref(h), out.ctypes.data, out.nbytes) == 0
    return 0, out


def _to8(v, depth):
    return (v >> 8) if depth == 16 else (v if depth == 8 else v * 255 // ((1 << depth) - 1))


# ---- PNG ----------------------------------------------------------------------------
@pytest.mark.parametrize("depth", [1, 2, 4, 8, 16])
@pytest.mark.parametrize("interlace", [0, 1])
def test_png_grey(tmp_path, depth, interlace):
    rng = np.random.default_rng(depth)
    s = rng.integers(0, 1 << depth, (13, 11, 1))
    p = tmp_path / "g.png"
    write_png(p, s, 0, depth, interlace=interlace, seed=depth)
    st, img = decode(p)
    g = _to8(s[..., 0], depth)
    assert st == 0 and np.array_equal(img, np.stack([g, g, g, np.full_like(g, 255)], -1))


@pytest.mark.parametrize("depth", [8, 16])
@pytest.mark.parametrize("ctype", [2, 4, 6])
@pytest.mark.parametrize("interlace", [0, 1])
def test_png_colour_and_alpha(tmp_path, depth, ctype, interlace):
    nc = {2: 3, 4: 2, 6: 4}[ctype]
    rng = np.random.default_rng(depth * 10 + ctype)
    s = rng.integers(0, 1 << depth, (9, 17, nc))
    p = tmp_path / "c.png"
    write_png(p, s, ctype, depth, interlace=interlace, seed=ctype)
    st, img = decode(p)
    v = _to8(s, depth)
    if ctype == 2:
        want = np.concatenate([v, np.full(v.shape[:2] + (1,), 255)], -1)
    elif ctype == 4:
        want = np.stack([v[..., 0], v[..., 0], v[..., 0], v[..., 1]], -1)
    else:
        want = v
    assert st == 0 and np.array_equal(img, want)


@pytest.mark.parametrize("depth", [1, 2, 4, 8])
def test_png_palette_with_transparency(tmp_path, depth):
    rng = np.random.default_rng(depth)
    n = 1 << depth
    plte = rng.integers(0, 256, (n, 3))
    trns = list(rng.integers(0, 256, max(1, n // 2)))
    idx = rng.integers(0, n, (10, 7, 1))
    p = tmp_path / "p.png"
    write_png(p, idx, 3, depth, plte=plte, trns=trns)
    st, img = decode(p)
    alpha = np.array([trns[i] if i < len(trns) else 255 for i in idx.reshape(-1)]).reshape(10, 7)
    want = np.concatenate([plte[idx[..., 0]], alpha[..., None]], -1)
    assert st == 0 and np.array_equal(img, want)


def test_png_colour_key_transparency(tmp_path):
    s = np.array([[[10, 20, 30], [1, 2, 3]], [[10, 20, 30], [4, 5, 6]]])
    p = tmp_path / "k.png"
    write_png(p, s, 2, 8, trns=struct.pack(">HHH", 10, 20, 30))
    st, img = decode(p)
    assert st == 0 and list(img[..., 3].reshape(-1)) == [0, 255, 0, 255]


def test_png_rejects_corrupt_data(tmp_path):
    s = np.random.default_rng(0).integers(0, 256, (8, 8, 3))
    p = tmp_path / "x.png"
    write_png(p, s, 2, 8)
    data = bytearray(p.read_bytes())
    data[-30] ^= 0xFF  # inside the compressed stream
    (tmp_path / "bad.png").write_bytes(bytes(data[:60]))
    assert decode(tmp_path / "bad.png")[0] == FILE_NOT_FOUND
    assert decode(tmp_path / "missing.png")[0] == FILE_NOT_FOUND


# ---- TGA / PNM ----------------------------------------------------------------------
@pytest.mark.parametrize("grey", [False, True])
@pytest.mark.parametrize("rle", [False, True])
@pytest.mark.parametrize("top_left", [False, True])
def test_tga(tmp_path, grey, rle, top_left):
    rng = np.random.default_rng(int(grey) * 4 + int(rle) * 2 + int(top_left))
    rgba = rng.integers(0, 4, (12, 9, 4)).astype(np.uint8) * 60  # runs for RLE
    if grey:
        rgba[..., 1] = rgba[..., 2] = rgba[..., 0]
        rgba[..., 3] = 255
    p = tmp_path / "t.tga"
    write_tga(p, rgba, grey, rle, top_left)
    st, img = decode(p)
    assert st == 0 and np.array_equal(img, rgba)


@pytest.mark.parametrize("kind,maxv", [("P5", 255), ("P6", 255), ("P6", 65535), ("P5", 1000)])
def test_pnm(tmp_path, kind, maxv):
    nc = 1 if kind == "P5" else 3
    rng = np.random.default_rng(maxv)
    s = rng.integers(0, maxv + 1, (5, 6, nc))
    body = s.astype(">u2").tobytes() if maxv > 255 else s.astype(np.uint8).tobytes()
    p = tmp_path / "i.pnm"
    p.write_bytes(f"{kind}\n# comment\n6 5\n{maxv}\n".encode() + body)
    st, img = decode(p)
    v = s >> 8 if maxv == 65535 else (s if maxv == 255 else (s * 255 + maxv // 2) // maxv)
    want = np.concatenate([np.repeat(v, 3 // nc, -1) if nc == 1 else v, np.full((5, 6, 1), 255)], -1)
    assert st == 0 and np.array_equal(img, want)


# ---- manifest ---------------------------------------------------------------------
def _parse(path, min_dim=0):
    lib = ntc.load()
    dim, count = C.c_int32(), C.c_int32()
    src = (L.ManifestSource * 16)()
    st = lib.ntc_manifest_parse(str(path).encode(), min_dim, C.byref(dim), C.byref(count), src)
    return st, dim.value, [src[i] for i in range(count.value)]


def _set(tmp_path, dim=64):
    """Three images of a PBR set written as PNG / TGA / PNM."""
    rng = np.random.default_rng(dim)
    imgs = {n: rng.integers(0, 256, (dim, dim, 4)).astype(np.uint8) for n in ("a", "n", "orm")}
    imgs["a"][..., 3] = 255
    write_png(tmp_path / "albedo.png", imgs["a"][..., :3], 2, 8)
    write_tga(tmp_path / "normal.tga", np.concatenate([imgs["n"][..., :3], np.full((dim, dim, 1), 255, np.uint8)], -1))
    (tmp_path / "orm.ppm").write_bytes(f"P6\n{dim} {dim}\n255\n".encode() + imgs["orm"][..., :3].tobytes())
    return imgs


def _manifest(tmp_path, textures, **top):
    m = dict(top)
    m["textures"] = textures
    p = tmp_path / "manifest.json"
    p.write_text(json.dumps(m))
    return p


PBR = [{"fileName": "albedo.png", "name": "base", "isSRGB": True, "semantics": {"Albedo": "RGB"}},
       {"fileName": "normal.tga", "semantics": {"Normal": "RGB"}},
       {"fileName": "orm.ppm", "verticalFlip": True, "semantics": {"Roughness": "G", "AO": "R", "Metallic": "B"}}]


def test_manifest_sources_in_semantic_order(tmp_path):
    _set(tmp_path)
    st, dim, src = _parse(_manifest(tmp_path, PBR, width=64, height=64), 64)
    assert st == SUCCESS and dim == 64
    # Albedo(1) < Metallic(6) < Normal(7) < AO(8) < Roughness(9)
    got = [(ntc.Semantic(s.semantic).name, s.channels.decode(), s.path.decode().rsplit("/", 1)[-1]) for s in src]
    assert got == [("Albedo", "RGB", "albedo.png"), ("Metallic", "B", "orm.ppm"), ("Normal", "RGB", "normal.tga"),
                   ("AO", "R", "orm.ppm"), ("Roughness", "G", "orm.ppm")]
    assert src[0].name == b"base" and src[0].is_srgb == 1 and src[1].vertical_flip == 1


def test_manifest_errors(tmp_path):
    _set(tmp_path)
    cases = [
        (dict(width=64), None, INVALID_MANIFEST),                                  # no "textures"
        (dict(), PBR, INVALID_MANIFEST),                                           # no width / height
        (dict(width=64, height=32), PBR, INVALID_MANIFEST),                        # not square
        (dict(width=16), PBR, INVALID_MANIFEST),                                   # below the minimum side
        (dict(width=96), PBR, INVALID_DIMENSION),                                  # not a power of two
        (dict(height=64), [{"fileName": "albedo.png"}], INVALID_MANIFEST),        # no semantics
        (dict(width=64), [{"semantics": {"Albedo": "RGB"}}], INVALID_MANIFEST),   # no file name
        (dict(width=64), [{"fileName": "albedo.png", "semantics": {"Albedo": "RG"}}], INVALID_MANIFEST),
        (dict(width=64), [{"fileName": "albedo.png", "semantics": {"None": "GR"}}], INVALID_MANIFEST),  # order
        (dict(width=64), [{"fileName": "albedo.png", "semantics": {"Albedo": 3}}], INVALID_MANIFEST),
        (dict(width=64), [{"fileName": "nothere.png", "semantics": {"Albedo": "RGB"}}], FILE_NOT_FOUND),
        (dict(width=128), PBR, INVALID_MANIFEST),                                  # images are 64^2
        (dict(width=64), [{"fileName": "albedo.png", "semantics": {"Albedo": "RGB"}}] * 6, INVALID_MANIFEST),  # 18 ch
    ]
    for top, tex, want in cases:
        if tex is None:
            p = tmp_path / "m.json"
            p.write_text(json.dumps(top))
        else:
            p = _manifest(tmp_path, tex, **top)
        assert _parse(p, 64)[0] == want, (top, tex)
    (tmp_path / "broken.json").write_text('{"width": 64, "textures": [')
    assert _parse(tmp_path / "broken.json", 64)[0] == INVALID_MANIFEST
    assert _parse(tmp_path / "absent.json", 64)[0] == FILE_NOT_FOUND
    # a "None" semantic takes its channel count from the string
    st, _, src = _parse(_manifest(tmp_path, [{"fileName": "albedo.png", "semantics": {"None": "RA"}}], width=64), 64)
    assert st == SUCCESS and src[0].channels == b"RA"


# ---- against the reference's loader (stb_image) --------------------------------------
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "images")
_STB = np.load(os.path.join(GOLDEN, "stb_rgba.npz"))


@pytest.mark.parametrize("name", sorted(_STB.files))
def test_decoder_matches_stb_image(name):
    """Every fixture decodes to stb_image's RGBA: bit for bit for the lossless
    formats (PNG, BMP, TGA, PNM); for JPEG within 3 levels, 99 % of values
    exact on average. stb_image's integer inverse DCT and fixed-point YCbCr
    conversion are restated here in float from T.81 and JFIF rather than
    reproduced (its chroma upsampling is matched), hence the rounding-level
    differences; the measured worst is 3 levels, mean 0.045."""
    st, got = decode(os.path.join(GOLDEN, name))
    assert st == 0, (name, st)
    want = _STB[name]
    assert got.shape == want.shape, (name, got.shape, want.shape)
    if name.endswith(".jpg"):
        d = np.abs(got.astype(np.int32) - want.astype(np.int32))
        assert d.max() <= 3 and d.mean() <= 0.06, (name, int(d.max()), float(d.mean()))
        assert (got[..., 3] == 255).all()
    else:
        assert np.array_equal(got, want), (name, int(np.count_nonzero(got != want)))


def test_stb_fixtures_are_stb_image_output():
    """The expected arrays are what the reference's stb_image (compiled from
    the reference tree into oracle/_ref by `make -C oracle ref`) decodes
    from the fixture files; skipped where the reference tree is absent."""
    so = os.path.join(os.path.dirname(GOLDEN), "..", "..", "oracle", "_ref", "libstbi.so")
    if not os.path.exists(so):
        pytest.skip("oracle/_ref/libstbi.so not built (needs the reference tree)")
    stb = C.CDLL(so)
    stb.stbi_load.restype = C.POINTER(C.c_ubyte)
    stb.stbi_load.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int]
    stb.stbi_image_free.argtypes = [C.c_void_p]
    for name in _STB.files:
        w, h, n = C.c_int(), C.c_int(), C.c_int()
        p = stb.stbi_load(os.path.join(GOLDEN, name).encode(), C.byref(w), C.byref(h), C.byref(n), 4)
        assert p, name
        live = np.ctypeslib.as_array(p, shape=(h.value, w.value, 4)).copy()
        stb.stbi_image_free(p)
        assert np.array_equal(live, _STB[name]), name


def test_truncated_and_unsupported_images_fail(tmp_path):
    """A JPEG cut short inside its first scan header and a BMP declaring
    run-length compression are not read: the loader reports a failed load
    (FileNotFound, as LoadTextures does for any image stb_image cannot
    decode, libopenntc.cpp:1302-1307)."""
    data = open(os.path.join(GOLDEN, "q90_420.jpg"), "rb").read()
    sos = data.index(b"\xff\xda")
    cut = tmp_path / "cut.jpg"
    cut.write_bytes(data[:sos + 6])
    assert decode(cut)[0] == FILE_NOT_FOUND
    # a BMP header declaring RLE8 compression
    bmp = bytearray(open(os.path.join(GOLDEN, "pal8.bmp"), "rb").read())
    struct.pack_into("<I", bmp, 30, 1)
    rle = tmp_path / "rle.bmp"
    rle.write_bytes(bytes(bmp))
    assert decode(rle)[0] == FILE_NOT_FOUND


def _peak_rss_kb():
    import resource
    return resource.getrusage(resource.RUSAGE_SELF).ru_maxrss


@pytest.mark.parametrize("kind", ["tga", "tga_rle", "png", "jpeg", "jpeg_progressive", "bmp", "pnm"])
def test_headers_claiming_huge_images_fail_without_allocating(tmp_path, kind):
    """Short files whose headers claim 65535 x 65535 (or up to 2^31) pixels:
    the decode fails (FileNotFound) before sizing anything from the header,
    and the process does not grow by gigabytes (a 65535^2 RGBA image is 16 GiB)."""
    p = tmp_path / f"huge.{kind}"
    W = H = 65535
    if kind.startswith("tga"):
        hdr = bytearray(18)
        hdr[2] = 10 if kind == "tga_rle" else 2
        struct.pack_into("<HH", hdr, 12, W, H)
        hdr[16], hdr[17] = 32, 0x20
        p.write_bytes(bytes(hdr) + bytes(64))
    elif kind == "png":
        ihdr = struct.pack(">IIBBBBB", W, H, 8, 6, 0, 0, 0)
        p.write_bytes(b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr) + _chunk(b"IDAT", zlib.compress(bytes(1000)))
                      + _chunk(b"IEND", b""))
    elif kind.startswith("jpeg"):
        sof = 0xC2 if kind == "jpeg_progressive" else 0xC0
        body = struct.pack(">BHHB", 8, H, W, 3) + bytes([1, 0x22, 0, 2, 0x11, 1, 3, 0x11, 1])
        p.write_bytes(b"\xff\xd8" + bytes([0xFF, sof]) + struct.pack(">H", len(body) + 2) + body + b"\xff\xd9")
    elif kind == "bmp":
        hdr = bytearray(54)
        hdr[0:2] = b"BM"
        struct.pack_into("<IIiiHHI", hdr, 10, 54, 40, 2 ** 31 - 1, -2 ** 31, 1, 24, 0)
        p.write_bytes(bytes(hdr) + bytes(64))
    else:
        p.write_bytes(b"P6\n65535 65535\n255\n" + bytes(64))
    before = _peak_rss_kb()
    assert decode(p)[0] == FILE_NOT_FOUND
    assert _peak_rss_kb() - before < 256 * 1024  # kB
