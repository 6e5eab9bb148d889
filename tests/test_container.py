"""The ONTC container through the C ABI, on the CPU: Dump / Load / FileData /
FillNTCConstants (openntc_runtime.cpp:156-468). The layout is the reference's:
a 16-byte header {'ONTC', 4, json bytes, blob bytes}, compact JSON with every
object's keys in sorted order (nlohmann::json's std::map objects, :188-225),
then the blobs g0_<level>..., g1_<level>..., decoder, each at a multiple of
256 bytes after the JSON with zero bytes between them (:227-246)."""
import ctypes as C
import json
import os
import struct

import numpy as np
import pytest

import openntc_amd as ntc
from openntc_amd import _lib as L

SUCCESS, ALLOC, FILE_NOT_FOUND, INVALID_FILE, WRITE_FAIL, INVALID_STATE = 0, 2, 4, 5, 6, 7
SEM_NAMES = ["None", "Albedo", "Alpha", "Displacement", "Emissive", "Gloss", "Metallic", "Normal",
             "AO", "Roughness", "Specular", "Transmission"]
PROFILES = ["bpp_0_2", "bpp_0_5", "bpp_1_0", "bpp_2_25"]


def _desc(rng, levels=3):
    """A CompressedData of `levels` levels with blob sizes that are not
    multiples of 256; returns (struct, kept numpy buffers)."""
    d = L.Compressed()
    keep = []
    g0 = [rng.integers(0, 256, 4 * int(rng.integers(1, 700)), dtype=np.uint8) for _ in range(levels)]
    g1 = [rng.integers(0, 256, 4 * int(rng.integers(1, 300)), dtype=np.uint8) for _ in range(levels)]
    dec = rng.integers(0, 256, 11840, dtype=np.uint8)
    off0 = off1 = 0
    for i in range(levels):
        d.g0[i], d.g1[i] = g0[i].ctypes.data, g1[i].ctypes.data
        d.g0_size[i], d.g1_size[i] = g0[i].nbytes, g1[i].nbytes
        d.g0_offset[i], d.g1_offset[i] = off0, off1
        off0 += g0[i].nbytes
        off1 += g1[i].nbytes
        d.g0_grid_dim[i] = 2048 >> (2 * i)
        d.g1_grid_dim[i] = 1024 >> (2 * i)
    d.g0_total_size, d.g1_total_size = off0, off1
    d.decoder, d.decoder_size = dec.ctypes.data, dec.nbytes
    d.g0_bits, d.g1_bits, d.g0_channels, d.g1_channels = 4, 2, 16, 12
    d.dim, d.mip_count, d.level_count = 4096, 11, levels
    sems = [1, 1, 1, 6, 7, 7, 7, 8, 9]
    d.channel_count = len(sems)
    for i, s in enumerate(sems):
        d.channel_semantics[i] = s
    d.profile = 3
    d.max_abs_a1, d.max_abs_a2, d.s_a1, d.s_a2 = 3.25, 1.1, np.float32(1.1 * 3.25 / 127), 0.0078125
    keep += g0 + g1 + [dec]
    return d, keep


def _dump(lib, path, d):
    return lib.ntc_container_dump(str(path).encode(), C.byref(d))


def _load(lib, path):
    f = C.c_void_p()
    r = lib.ntc_container_load(str(path).encode(), C.byref(f))
    return r, f


def _sorted_keys(o):
    if isinstance(o, dict):
        return list(o) == sorted(o) and all(_sorted_keys(v) for v in o.values())
    if isinstance(o, list):
        return all(_sorted_keys(v) for v in o)
    return True


def test_dump_layout_matches_reference_format(tmp_path):
    lib = ntc.load()
    d, keep = _desc(np.random.default_rng(1))
    p = tmp_path / "a.ontc"
    assert _dump(lib, p, d) == SUCCESS
    raw = p.read_bytes()
    magic, version, jsz, bsz = struct.unpack("<4I", raw[:16])
    assert magic == 0x43544E4F and raw[:4] == b"ONTC" and version == 4
    assert len(raw) == 16 + jsz + bsz
    text = raw[16:16 + jsz].decode()
    j = json.loads(text)
    assert _sorted_keys(j)
    # the whole JSON text is what a compact, key-sorted dump gives
    want = {
        "blobs": [], "calibration": {"max_abs_a1": 3.25, "max_abs_a2": float(np.float32(1.1)),
                                     "s_a1": float(np.float32(d.s_a1)), "s_a2": 0.0078125},
        "channel_semantics": [SEM_NAMES[d.channel_semantics[i]] for i in range(d.channel_count)],
        "dim": 4096, "g0": {"bits": 4, "channels": 16, "grid_dims": list(d.g0_grid_dim)},
        "g1": {"bits": 2, "channels": 12, "grid_dims": list(d.g1_grid_dim)},
        "level_count": d.level_count, "mip_count": 11, "profile": "bpp_2_25",
        "source": {"generator": "openntc", "version": 5}}
    names = [f"g0_{i}" for i in range(d.level_count)] + [f"g1_{i}" for i in range(d.level_count)] + ["decoder"]
    sizes = [d.g0_size[i] for i in range(d.level_count)] + [d.g1_size[i] for i in range(d.level_count)] + \
        [d.decoder_size]
    off = 0
    for n, s in zip(names, sizes):
        off = (off + 255) // 256 * 256
        want["blobs"].append({"name": n, "offset": off, "size": s})
        off += s
    assert bsz == off
    assert text == json.dumps(want, sort_keys=True, separators=(",", ":"))
    # blob bytes in place, zeros between them
    body = np.frombuffer(raw[16 + jsz:], np.uint8)
    covered = np.zeros(bsz, bool)
    bufs = keep[:d.level_count] + keep[d.level_count:2 * d.level_count] + [keep[-1]]
    for b, buf in zip(want["blobs"], bufs):
        assert np.array_equal(body[b["offset"]:b["offset"] + b["size"]], buf)
        covered[b["offset"]:b["offset"] + b["size"]] = True
    assert not np.any(body[~covered])


def test_dump_load_round_trip_is_byte_exact(tmp_path):
    lib = ntc.load()
    d, keep = _desc(np.random.default_rng(2), levels=4)
    a, b = tmp_path / "a.ontc", tmp_path / "b.ontc"
    assert _dump(lib, a, d) == SUCCESS
    r, f = _load(lib, a)
    assert r == SUCCESS and f.value
    try:
        got = L.Compressed()
        assert lib.ntc_file_data(f, C.byref(got)) == SUCCESS
        for name in ("g0_bits", "g1_bits", "g0_channels", "g1_channels", "dim", "mip_count", "level_count",
                     "channel_count", "profile", "max_abs_a1", "max_abs_a2", "s_a1", "s_a2", "decoder_size",
                     "g0_total_size", "g1_total_size"):
            assert getattr(got, name) == getattr(d, name), name
        for arr in ("g0_grid_dim", "g1_grid_dim", "channel_semantics", "g0_size", "g1_size", "g0_offset",
                    "g1_offset"):
            assert list(getattr(got, arr)) == list(getattr(d, arr)), arr
        # the levels of each pyramid are contiguous in FileData's packed copy
        for g in ("g0", "g1"):
            for i in range(d.level_count):
                ptr_i = getattr(got, g)[i]
                assert ptr_i == getattr(got, g)[0] + getattr(got, g + "_offset")[i]
                n = getattr(got, g + "_size")[i]
                assert C.string_at(ptr_i, n) == C.string_at(getattr(d, g)[i], n)
        assert C.string_at(got.decoder, got.decoder_size) == C.string_at(d.decoder, d.decoder_size)
        assert _dump(lib, b, got) == SUCCESS
    finally:
        lib.ntc_file_free(f)
    assert a.read_bytes() == b.read_bytes()


def test_fill_constants_layout():
    lib = ntc.load()
    d, keep = _desc(np.random.default_rng(3))
    buf = (C.c_uint8 * 240)()
    assert lib.ntc_fill_constants(C.byref(d), buf) == SUCCESS
    v = bytes(buf)
    g0d, g1d = struct.unpack("<8i", v[0:32]), struct.unpack("<8i", v[32:64])
    g0o, g1o = struct.unpack("<8I", v[64:96]), struct.unpack("<8I", v[96:128])
    assert list(g0d) == list(d.g0_grid_dim[:3]) + [0] * 5 and list(g1d) == list(d.g1_grid_dim[:3]) + [0] * 5
    assert list(g0o) == list(d.g0_offset[:3]) + [0] * 5 and list(g1o) == list(d.g1_offset[:3]) + [0] * 5
    b0, b1, c0, c1, dim, mips = struct.unpack("<6i", v[128:152])
    assert (b0, b1, c0, c1, dim, mips) == (4, 2, 16, 12, 4096, 11)
    r1, r2 = struct.unpack("<2f", v[152:160])
    assert r1 == np.float32(1) / np.float32(d.s_a1) and r2 == np.float32(128.0)
    (cc,) = struct.unpack("<i", v[160:164])
    assert cc == 9 and v[164:176] == bytes(12)
    assert list(struct.unpack("<16I", v[176:240])) == list(d.channel_semantics)


def _write(path, header, text, blob=b""):
    path.write_bytes(struct.pack("<4I", *header) + text + blob)


def _valid_json(levels=1, blobs=None, **over):
    j = {"blobs": blobs if blobs is not None else
         [{"name": f"g0_{i}", "offset": 0, "size": 4} for i in range(levels)] +
         [{"name": f"g1_{i}", "offset": 0, "size": 4} for i in range(levels)] +
         [{"name": "decoder", "offset": 0, "size": 8}],
         "calibration": {"max_abs_a1": 1.0, "max_abs_a2": 2.0, "s_a1": 0.5, "s_a2": 0.25},
         "channel_semantics": ["Albedo", "Normal"], "dim": 1024,
         "g0": {"bits": 4, "channels": 8, "grid_dims": [512, 128, 0, 0, 0, 0]},
         "g1": {"bits": 4, "channels": 8, "grid_dims": [256, 64, 0, 0, 0, 0]},
         "level_count": levels, "mip_count": 9, "profile": "bpp_1_0"}
    j.update(over)
    return j


def _file(path, j, blob=bytes(16), version=4, magic=0x43544E4F, blob_size=None):
    # keys in an order of their own and with spaces: Load takes any JSON
    text = json.dumps(j, indent=1).encode()
    _write(path, (magic, version, len(text), len(blob) if blob_size is None else blob_size), text, blob)


def test_load_accepts_any_key_order_and_whitespace(tmp_path):
    lib = ntc.load()
    p = tmp_path / "x.ontc"
    j = _valid_json(levels=2)
    j = dict(reversed(list(j.items())))
    _file(p, j)
    r, f = _load(lib, p)
    assert r == SUCCESS
    got = L.Compressed()
    lib.ntc_file_data(f, C.byref(got))
    assert got.level_count == 2 and got.dim == 1024 and got.profile == 2 and got.channel_count == 2
    assert list(got.channel_semantics[:2]) == [1, 7] and got.s_a2 == 0.25
    assert got.g0_offset[1] == 4 and got.g0_total_size == 8 and got.decoder_size == 8
    lib.ntc_file_free(f)


@pytest.mark.parametrize("case", [
    "missing", "short", "magic", "version", "json_past_end", "blob_past_end", "not_json", "no_g0",
    "g0_not_object", "dims_5", "dims_float", "bits_float", "no_dim", "cal_string", "profile_bad",
    "profile_missing", "sem_unknown", "sem_17", "sem_not_string", "levels_0", "levels_7", "mips_0", "mips_13",
    "blobs_missing", "blob_absent", "blob_negative", "blob_offset_huge", "blob_size_huge", "blob_wrap",
    "blob_not_object",
])
def test_load_error_codes(tmp_path, case):
    """Load's failures (openntc_runtime.cpp:307-456); a truncated file and a
    blob reaching past the file are InvalidFile, as are offsets and sizes
    whose sum would wrap 64 bits."""
    lib = ntc.load()
    p = tmp_path / "bad.ontc"
    want = INVALID_FILE
    j = _valid_json()
    big = 2 ** 63 - 1
    if case == "missing":
        p = tmp_path / "nope.ontc"
        want = FILE_NOT_FOUND
    elif case == "short":
        p.write_bytes(b"ONTC\x04\x00\x00")
    elif case == "magic":
        _file(p, j, magic=0x43544E4E)
    elif case == "version":
        _file(p, j, version=5)
    elif case == "json_past_end":
        text = json.dumps(j).encode()
        _write(p, (0x43544E4F, 4, len(text) + 100, 16), text, bytes(16))
    elif case == "blob_past_end":
        _file(p, j, blob=bytes(8), blob_size=16)
    elif case == "not_json":
        _write(p, (0x43544E4F, 4, 7, 0), b"{nope:1")
    else:
        mods = {
            "no_g0": lambda: j.pop("g0"),
            "g0_not_object": lambda: j.update(g0=[1, 2]),
            "dims_5": lambda: j["g0"].update(grid_dims=[1, 2, 3, 4, 5]),
            "dims_float": lambda: j["g1"].update(grid_dims=[1.5, 2, 3, 4, 5, 6]),
            "bits_float": lambda: j["g0"].update(bits=4.0),
            "no_dim": lambda: j.pop("dim"),
            "cal_string": lambda: j["calibration"].update(s_a1="0.5"),
            "profile_bad": lambda: j.update(profile="bpp_9"),
            "profile_missing": lambda: j.pop("profile"),
            "sem_unknown": lambda: j.update(channel_semantics=["Albedo", "Shininess"]),
            "sem_17": lambda: j.update(channel_semantics=["AO"] * 17),
            "sem_not_string": lambda: j.update(channel_semantics=[1]),
            "levels_0": lambda: j.update(level_count=0),
            "levels_7": lambda: j.update(level_count=7),
            "mips_0": lambda: j.update(mip_count=0),
            "mips_13": lambda: j.update(mip_count=13),
            "blobs_missing": lambda: j.pop("blobs"),
            "blob_absent": lambda: j.update(blobs=[b for b in j["blobs"] if b["name"] != "decoder"]),
            "blob_negative": lambda: j["blobs"][0].update(offset=-4),
            "blob_offset_huge": lambda: j["blobs"][0].update(offset=big),
            "blob_size_huge": lambda: j["blobs"][0].update(size=big),
            "blob_wrap": lambda: j["blobs"][0].update(offset=big, size=big),
            "blob_not_object": lambda: j["blobs"].append(3),
        }
        mods[case]()
        _file(p, j)
    r, f = _load(lib, p)
    assert r == want and not f.value


def test_dump_to_unwritable_path_fails(tmp_path):
    lib = ntc.load()
    d, keep = _desc(np.random.default_rng(4))
    assert _dump(lib, tmp_path / "no" / "such" / "dir.ontc", d) == WRITE_FAIL
    assert lib.ntc_container_dump(None, C.byref(d)) == INVALID_STATE
    f = C.c_void_p()
    assert lib.ntc_container_load(None, C.byref(f)) == FILE_NOT_FOUND
