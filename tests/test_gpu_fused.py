"""Fused BF16 training step (ntc_fused_train_step + ntc_decoder_update) against
the FP64 CPU oracle of the same step: gather with noise, forward, loss,
backward, latent scatter.

Two references: (1) the same step with the kernel's BF16 rounding points
(features, weights, activations, output and hidden-unit gradients rounded to
BF16 where the kernel rounds them, products accumulated in FP64), which the
kernel must match to 2e-3 relative in norm (FP32 accumulation order);
(2) the plain FP64 oracle, which bounds the precision loss of BF16: loss within
1e-2, decoder gradients within 5e-2; latent gradients within 1e-1, because
activation masks flip for pre-activations within BF16 rounding of zero."""
import ctypes as C
import zlib

import numpy as np
import pytest

import openntc_amd as ntc
from openntc_amd import _lib as L

pytestmark = pytest.mark.gpu

from tests import gpu_util as G  # noqa: E402

REL_TOL = 5e-2
REL_TOL_LATENT = 1e-1
REL_TOL_BF16 = 2e-3


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def bf16(x):
    """Round to BF16 (nearest even) and return as float64."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def act_f(act, z):
    if act == 1:
        return np.maximum(z, 0.0)
    return np.where(z < -1.5, 0.0, np.where(z < 1.5, z / 3 * (z + 1.5), z))


def act_b(act, z, g):
    if act == 1:
        return np.where(z > 0, g, 0.0)
    return np.where(z < -1.5, 0.0, np.where(z < 1.5, (2 * z + 1.5) / 3 * g, g))


def bf16_step(x32, yt, w0, w1, wout, O, act, scale):
    """The fused step with the kernel's BF16 rounding points."""
    xb, w0b, w1b, wob = bf16(x32), bf16(w0), bf16(w1), bf16(wout[:O])
    z0 = w0b @ xb
    a0 = bf16(act_f(act, z0.astype(np.float32)))
    z1 = w1b @ a0
    a1 = bf16(act_f(act, z1.astype(np.float32)))
    yh = (wob @ a1).astype(np.float32)
    r = yt.astype(np.float32) - yh
    sse = float(np.sum(r.astype(np.float64) ** 2))
    dy = bf16(np.float32(scale) * r)
    dz1 = bf16(act_b(act, z1, wob.T @ dy).astype(np.float32))
    dz0 = bf16(act_b(act, z0, w1b.T @ dz1).astype(np.float32))
    dx = w0b.T @ dz0
    return sse, dz0 @ xb.T, dz1 @ a0.T, dy @ a1.T, dx


def _case(oracle, kind, dim=256, mip_min=4):
    if kind == "c1_bilinear_relu":
        return oracle.layout_custom(4, 4, 8, 16, 2, 64, 1, 1, dim, 9, mip_min), 1
    if kind == "c3_bilinear_h128":
        return oracle.layout_custom(4, 4, 16, 16, 2, 128, 1, 1, dim, 9, mip_min), 1
    if kind == "c2_bilinear":
        return oracle.layout_custom(4, 2, 12, 16, 2, 64, 1, 1, 256), 1
    if kind == "ref_bpp_2_25":
        return oracle.layout_profile(3, 256), 0
    if kind == "ref_bpp_1_0":
        return oracle.layout_profile(2, 256), 0
    if kind == "ref_bpp_0_5":
        return oracle.layout_profile(1, 256), 0
    raise KeyError(kind)


def _step_case(oracle, kind, lod, n_tiles, dim=256, mip_min=4):
    """Inputs of one fused step, its FP64-oracle and BF16-rounding references,
    and the device-side arguments (without outputs)."""
    O = oracle
    Lay, act = _case(O, kind, dim, mip_min)
    Lay.activation = act
    lib = G.lib()
    rng = np.random.default_rng(zlib.crc32(f"{kind}-{lod}-{dim}".encode()) if dim != 256 else
                                zlib.crc32(f"{kind}-{lod}".encode()))
    level = O.level_for_lod(Lay, lod)
    fd = O.feature_desc(Lay, level)
    M = Lay.mip_dim[lod]
    T = min(M, 256)
    H, F, Od = Lay.hidden_dim, Lay.feature_dim, 9
    origins = rng.integers(0, M, 2 * n_tiles).astype(np.int32)
    g0 = rng.uniform(-0.5, 0.5, (fd["g0_dim"], fd["g0_dim"], fd["c0"])).astype(np.float32)
    g1 = rng.uniform(-0.5, 0.5, (fd["g1_dim"], fd["g1_dim"], fd["c1"])).astype(np.float32)
    w0 = (rng.uniform(-1, 1, (H, F)) / np.sqrt(F)).astype(np.float32)
    w1 = (rng.uniform(-1, 1, (H, H)) / np.sqrt(H)).astype(np.float32)
    wout = (rng.uniform(-1, 1, (16, H)) / np.sqrt(H)).astype(np.float32)
    pkg = rng.integers(0, 256, (M, M, Od)).astype(np.uint8)
    norm_lod = np.float32(lod / (Lay.mip_count - 1))
    step, seed = 5, 123

    # FP64 oracle of the step
    x = O.draw_features(fd, origins, T, M, norm_lod, g0.astype(np.float64), g1.astype(np.float64), seed=seed,
                        step=step, level=level, noise_on=True, d0=Lay.g0_delta, d1=Lay.g1_delta, dtype=np.float64)
    yt = O.draw_targets(origins, T, M, Od, pkg, np.float64)
    sse, dw0, dw1, dwout, dx = O.mlp_fwd_bwd(F, H, Od, act, x, yt, w0.astype(np.float64),
                                             w1.astype(np.float64), wout.astype(np.float64))
    rdg0, rdg1 = O.scatter(fd, origins, T, M, dx, g0.shape, g1.shape)
    # same step with the kernel's BF16 rounding points; FP32 features as the kernel gathers them
    x32 = O.draw_features(fd, origins, T, M, norm_lod, g0, g1, seed=seed, step=step, level=level, noise_on=True,
                          d0=Lay.g0_delta, d1=Lay.g1_delta)
    B = n_tiles * T * T
    scale = np.float32(-2.0) / (np.float32(Od) * np.float32(B))
    bsse, bw0, bw1, bwo, bdx = bf16_step(x32, yt, w0, w1, wout, Od, act, scale)
    bdg0, bdg1 = O.scatter(fd, origins, T, M, np.ascontiguousarray(bdx), g0.shape, g1.shape)
    # cells inside some texel's footprint (the scatter of ones reaches them)
    s0, s1 = O.scatter(fd, origins, T, M, np.ones_like(bdx), g0.shape, g1.shape)
    support = (np.abs(s0) > 0, np.abs(s1) > 0)

    fdc = G.feature_desc(fd)
    assert lib.ntc_fused_supported(G.byref(fdc), H, act) == 1
    dec = np.concatenate([w0.ravel(), w1.ravel(), wout.ravel()])
    assert dec.dtype == np.float32
    d_dec = G.dev(dec)
    tab = G.empty((lib.ntc_fused_table_elems(G.byref(fdc), H),), np.uint16)
    G.ok(lib.ntc_pack_decoder_bf16(G.byref(fdc), H, Od, G.p(d_dec), G.p(tab), None))
    og = G.dev(origins)
    d_g0, d_g1, d_pkg = G.dev(g0), G.dev(g1), G.dev(pkg)
    a = L.FusedArgs()
    a.feat = fdc
    a.tiles = G.tiles(og, n_tiles, T, M, float(norm_lod))
    a.noise = L.Noise(seed, step, level, 1, Lay.g0_delta, Lay.g1_delta)
    a.hidden_dim, a.out_dim, a.activation = H, Od, act
    a.g0, a.g1, a.weights, a.targets = G.p(d_g0), G.p(d_g1), G.p(tab), G.p(d_pkg)
    a.loss_scale = float(scale)
    a.scatter = 1
    return dict(a=a, fdc=fdc, H=H, F=F, Od=Od, act=act, B=B, dec=dec, g0=g0, g1=g1, pkg=pkg, support=support,
                fd=fd, origins=origins, T=T, M=M, bdx=np.ascontiguousarray(bdx), x32=x32,
                ref=(sse, dw0, dw1, dwout, rdg0, rdg1), bref=(bsse, bw0, bw1, bwo, bdg0, bdg1))


def _outputs(cs, blocks):
    """Fresh output buffers of one fused step with `blocks` partial slots."""
    n_dec = cs["dec"].size
    return dict(part=G.empty((blocks * n_dec,), np.float32), ssep=G.empty((blocks,), np.float64),
                dg0=G.empty(cs["g0"].shape, np.float32), dg1=G.empty(cs["g1"].shape, np.float32))


def _check_step(cs, out, blocks):
    """Reduces the partials (lr = 0) and checks them and the latent gradients
    against both references."""
    lib = G.lib()
    H, F, Od, dec = cs["H"], cs["F"], cs["Od"], cs["dec"]
    n_dec = dec.size
    grad = G.empty((n_dec,), np.float32)
    m = G.empty((n_dec,), np.float32)
    v = G.empty((n_dec,), np.float32)
    p = G.dev(dec)
    sse_out = G.empty((1,), np.float64)
    G.ok(lib.ntc_decoder_update(H, F, Od, blocks, G.p(out["part"]), 0.0, 0.9, 0.999, 1.0, 1.0, G.p(grad), G.p(m),
                                G.p(v), G.p(p), G.p(out["ssep"]), G.p(sse_out), None))
    g = G.host(grad)
    gw0, gw1, gwo = g[: H * F].reshape(H, F), g[H * F: H * F + H * H].reshape(H, H), g[H * F + H * H:].reshape(16, H)
    gdg0, gdg1 = G.host(out["dg0"]), G.host(out["dg1"])
    sse, dw0, dw1, dwout, rdg0, rdg1 = cs["ref"]
    bsse, bw0, bw1, bwo, bdg0, bdg1 = cs["bref"]
    # (1) against the BF16-rounding reference
    assert abs(G.host(sse_out)[0] - bsse) <= 1e-4 * bsse
    for got, want, name in ((gw0, bw0, "dW0"), (gw1, bw1, "dW1"), (gwo[:Od], bwo, "dWout"), (gdg0, bdg0, "dG0"),
                            (gdg1, bdg1, "dG1")):
        assert rel(got, want) < REL_TOL_BF16, (name, rel(got, want))
    # latent gradients only inside the texels' footprints, exactly zero
    # elsewhere (the tile-aware grid update reads and clears them only there)
    for got, sup, name in ((gdg0, cs["support"][0], "dG0"), (gdg1, cs["support"][1], "dG1")):
        outside = got[~sup]
        assert not outside.any(), (name, int(np.count_nonzero(outside)), float(np.max(np.abs(outside))))
    # (2) against the FP64 oracle; a batch of a few hundred texels has few
    # terms per decoder-gradient sum, where activation masks that flip
    # within BF16 rounding of zero weigh more
    tol, tol_lat = (REL_TOL, REL_TOL_LATENT) if cs["B"] >= 1024 else (0.25, 0.25)
    assert abs(G.host(sse_out)[0] - sse) <= 1e-2 * sse
    assert rel(gw0, dw0) < tol, rel(gw0, dw0)
    assert rel(gw1, dw1) < tol, rel(gw1, dw1)
    assert rel(gwo[:Od], dwout) < tol, rel(gwo[:Od], dwout)
    assert not gwo[Od:].any()
    assert rel(gdg0, rdg0) < tol_lat, rel(gdg0, rdg0)
    assert rel(gdg1, rdg1) < tol_lat, rel(gdg1, rdg1)
    # lr = 0 leaves the weights untouched
    assert np.array_equal(G.host(p), dec)


@pytest.mark.parametrize("kind", ["c1_bilinear_relu", "c3_bilinear_h128", "c2_bilinear", "ref_bpp_2_25",
                                  "ref_bpp_1_0", "ref_bpp_0_5"])
@pytest.mark.parametrize("lod,n_tiles", [(0, 2), (1, 1), (2, 3), (3, 4), (4, 5), (5, 2), (6, 7)])
def test_fused_step_vs_fp64_oracle(oracle, kind, lod, n_tiles):
    lib = G.lib()
    cs = _step_case(oracle, kind, lod, n_tiles)
    maxb = C.c_int32()
    G.ok(lib.ntc_fused_max_blocks(C.byref(maxb)))
    blocks = lib.ntc_fused_blocks_for(G.byref(cs["fdc"]), cs["H"], cs["act"], cs["B"], maxb.value)
    out = _outputs(cs, blocks)
    a = cs["a"]
    a.dg0, a.dg1, a.dw_partials, a.sse_partials, a.blocks = (G.p(out["dg0"]), G.p(out["dg1"]), G.p(out["part"]),
                                                             G.p(out["ssep"]), blocks)
    G.ok(lib.ntc_fused_train_step(G.byref(a), None), "ntc_fused_train_step")
    _check_step(cs, out, blocks)


@pytest.mark.parametrize("kind", ["c1_bilinear_relu", "c2_bilinear", "c3_bilinear_h128", "ref_bpp_2_25"])
@pytest.mark.parametrize("lod,n_tiles", [(0, 2), (1, 1), (2, 3), (3, 4), (4, 5), (5, 2), (6, 7)])
def test_fused_sets_step_vs_fp64_oracle(oracle, kind, lod, n_tiles):
    """The multi-set launch (ntc_fused_train_sets) at every compiled pass size:
    two sets, the same step in each (own outputs, own workgroup ranges of
    different sizes), each checked like the single-set kernel."""
    lib = G.lib()
    cs = _step_case(oracle, kind, lod, n_tiles)
    sizes = (C.c_int32 * 2)()
    n = lib.ntc_fused_pass_sizes(G.byref(cs["fdc"]), cs["H"], cs["act"], sizes)
    assert n >= 1
    rb = lib.ntc_fused_params_bytes()
    for pass_texels in list(sizes)[:n]:
        passes = -(-cs["B"] // pass_texels)
        counts = [min(passes, 200), min(passes, 37)]
        outs = [_outputs(cs, c) for c in counts]
        rec = np.zeros(2 * rb, np.uint8)
        for s in range(2):
            a = cs["a"]
            o = outs[s]
            a.dg0, a.dg1, a.dw_partials, a.sse_partials, a.blocks = (G.p(o["dg0"]), G.p(o["dg1"]), G.p(o["part"]),
                                                                     G.p(o["ssep"]), counts[s])
            G.ok(lib.ntc_fused_params_pack(G.byref(a), rec.ctypes.data + s * rb), "ntc_fused_params_pack")
        first = np.array([0, counts[0], counts[0] + counts[1]], np.int32)
        step = np.zeros(1, np.int32)
        grid = int(first[-1]) + 5  # idle workgroups past the last range
        G.ok(lib.ntc_fused_train_sets(G.byref(cs["fdc"]), cs["H"], cs["act"], int(pass_texels), G.p(G.dev(rec)),
                                      G.p(G.dev(first)), 2, G.p(G.dev(step)), grid, None), "ntc_fused_train_sets")
        for s in range(2):
            _check_step(cs, outs[s], counts[s])


def test_grid_update2_bitexact(oracle):
    rng = np.random.default_rng(3)
    n0, n1 = 4 * 1001, 4 * 333
    arrs = []
    for n in (n0, n1):
        arrs.append([(rng.standard_normal(n) * 1e-3).astype(np.float32),
                     (rng.standard_normal(n) * 1e-3).astype(np.float32),
                     rng.uniform(0, 1e-6, n).astype(np.float32),
                     rng.uniform(-1.1, 1.1, n).astype(np.float32)])
    d = [[G.dev(x) for x in a] for a in arrs]
    lr, c1, c2 = 0.007, 2.5, 40.0
    G.ok(G.lib().ntc_grid_update2_f32(n0, *[G.p(x) for x in d[0]], -1.0, 1.0 - 0.125,
                                      n1, *[G.p(x) for x in d[1]], -1.0, 1.0 - 0.5, lr, 0.9, 0.999, c1, c2, None))
    for (g, m, v, p), dd, delta in ((arrs[0], d[0], 0.125), (arrs[1], d[1], 0.5)):
        oracle.adam(lr, 0.9, 0.999, c1, c2, g, m, v, p)
        oracle.clamp_grid(delta, p)
        assert np.array_equal(G.host(dd[3]), p) and np.array_equal(G.host(dd[1]), m)
        assert np.array_equal(G.host(dd[2]), v) and not G.host(dd[0]).any()


def test_fused_context_training_tracks_plain(oracle):
    """The fused BF16 path and the plain FP32 path trained on the same seed
    reach the same quality (config-1 layout at 256^2)."""
    tex = ntc.synthetic_pbr(256, 20260)
    psnr = {}
    for path in (ntc.TrainPath.Plain, ntc.TrainPath.Fused):
        ctx = ntc.Context()
        lay = ntc.LayoutInfo(4, 4, 8, 16, 2, 64, ntc.FeatureLayout.Bilinear, ntc.Activation.ReLU)
        assert ctx.InitEx(layout=lay, min_dimension=64) == ntc.Result.Success
        assert ctx.LoadTextures(tex, 256) == ntc.Result.Success
        assert ctx.BeginTrainingEx(300, 1, 1e-2, 1e-3, -0.5, 0.5, path) == ntc.Result.Success
        prog = ctx.TrainUntilComplete()
        assert prog.phase == ntc.TrainPhase.TrainComplete, prog
        psnr[path] = ctx.Eval().psnr
    assert psnr[ntc.TrainPath.Fused] > 15.0
    assert abs(psnr[ntc.TrainPath.Fused] - psnr[ntc.TrainPath.Plain]) < 0.5, psnr



def test_grid_update_tiles_bitexact(oracle):
    """The footprint-aware dense update (gradient read and cleared only inside
    the tile footprints) leaves the level exactly as ntc_grid_update2_f32 does
    for the gradients of a fused scatter: 2048^2 mip, g0 1024^2 and g1 512^2 x
    16 channels, three tiles (one wrapping both edges, one overlapping it)."""
    O = oracle
    Lay = O.layout_custom(4, 4, 16, 16, 2, 128, 1, 1, 2048)
    act = 1
    Lay.activation = act
    lib = G.lib()
    rng = np.random.default_rng(11)
    fd = O.feature_desc(Lay, 0)
    M, T, H = Lay.mip_dim[0], 256, 128
    origins = np.array([M - 100, M - 37, M - 20, 150, 517, 1203], np.int32)
    n_tiles = len(origins) // 2
    fdc = G.feature_desc(fd)
    shapes = [(fd["g0_dim"], fd["g0_dim"], fd["c0"]), (fd["g1_dim"], fd["g1_dim"], fd["c1"])]
    init = [(rng.uniform(-1, 1, shp).astype(np.float32), (rng.standard_normal(shp) * 1e-3).astype(np.float32),
             rng.uniform(0, 1e-6, shp).astype(np.float32)) for shp in shapes]
    F = Lay.feature_dim
    dec = np.concatenate([(rng.uniform(-1, 1, H * F) / np.sqrt(F)).astype(np.float32),
                          (rng.uniform(-1, 1, H * H) / np.sqrt(H)).astype(np.float32),
                          (rng.uniform(-1, 1, 16 * H) / np.sqrt(H)).astype(np.float32)])
    d_dec = G.dev(dec)
    tab = G.empty((lib.ntc_fused_table_elems(G.byref(fdc), H),), np.uint16)
    G.ok(lib.ntc_pack_decoder_bf16(G.byref(fdc), H, 9, G.p(d_dec), G.p(tab), None))
    d_pkg = G.dev(rng.integers(0, 256, (M, M, 9)).astype(np.uint8))
    maxb = C.c_int32()
    G.ok(lib.ntc_fused_max_blocks(C.byref(maxb)))
    B = n_tiles * T * T
    blocks = lib.ntc_fused_blocks_for(G.byref(fdc), H, act, B, maxb.value)
    part = G.empty((blocks * dec.size,), np.float32)
    ssep = G.empty((blocks,), np.float64)
    sa = [[G.dev(x) for x in st] + [G.empty(st[0].shape, np.float32)] for st in init]  # p, m, v, g
    sb = [[G.dev(x) for x in st] + [G.empty(st[0].shape, np.float32)] for st in init]
    og = G.dev(origins)
    a = L.FusedArgs()
    a.feat = fdc
    a.tiles = G.tiles(og, n_tiles, T, M, 0.0)
    a.noise = L.Noise(123, 7, 0, 1, Lay.g0_delta, Lay.g1_delta)
    a.hidden_dim, a.out_dim, a.activation = H, 9, act
    a.g0, a.g1, a.weights, a.targets = G.p(sa[0][0]), G.p(sa[1][0]), G.p(tab), G.p(d_pkg)
    a.loss_scale = -2.0 / (9 * B)
    a.scatter = 1
    a.dg0, a.dg1, a.dw_partials, a.sse_partials, a.blocks = G.p(sa[0][3]), G.p(sa[1][3]), G.p(part), G.p(ssep), blocks
    G.ok(lib.ntc_fused_train_step(G.byref(a), None), "ntc_fused_train_step")
    grads = [G.host(sa[0][3]), G.host(sa[1][3])]
    assert np.count_nonzero(grads[0]) > 0 and np.count_nonzero(grads[1]) > 0
    for k in range(2):
        G.ok(lib.ntc_dev_upload(G.p(sb[k][3]), grads[k].ctypes.data, grads[k].nbytes), "upload")
    lr, b1c, b2c = 0.01, 1.0 / (1 - 0.9 ** 3), 1.0 / (1 - 0.999 ** 3)
    lo0, hi0, lo1, hi1 = -1.0, 1.0 - Lay.g0_delta, -1.0, 1.0 - Lay.g1_delta
    n = [int(np.prod(s)) for s in shapes]
    G.ok(lib.ntc_grid_update2_f32(n[0], G.p(sa[0][3]), G.p(sa[0][1]), G.p(sa[0][2]), G.p(sa[0][0]), lo0, hi0,
                                  n[1], G.p(sa[1][3]), G.p(sa[1][1]), G.p(sa[1][2]), G.p(sa[1][0]), lo1, hi1,
                                  lr, 0.9, 0.999, b1c, b2c, None))
    ad = L.GridAdam(G.p(sb[0][1]), G.p(sb[0][2]), G.p(sb[1][1]), G.p(sb[1][2]), lr, 0.9, 0.999, b1c, b2c,
                    lo0, hi0, lo1, hi1)
    G.ok(lib.ntc_grid_update_tiles_f32(G.byref(fdc), G.byref(a.tiles), G.p(sb[0][0]), G.p(sb[0][3]),
                                       G.p(sb[1][0]), G.p(sb[1][3]), G.byref(ad), None))
    for k in range(2):
        for j, name in enumerate("pmvg"):
            got, want = G.host(sb[k][j]), G.host(sa[k][j])
            assert np.array_equal(got, want), (k, name, int(np.count_nonzero(got != want)))
        assert not G.host(sb[k][3]).any()
    # every cell moved
    assert np.count_nonzero(G.host(sb[0][0]) == init[0][0]) < 0.01 * n[0]


# (g0 side, g1 side, g0 channels, g1 channels, mip side, tile side, origins)
_PART_CASES = {
    # three tiles, one wrapping both edges, one overlapping it
    "wrap_overlap": (1024, 512, 16, 16, 2048, 256, [2048 - 100, 2048 - 37, 2048 - 20, 150, 517, 1203]),
    # a coarse mip: 4 cells per texel, each rectangle half the grid wide; 3
    # and 2 channel quads
    "coarse_q3_q2": (256, 128, 12, 8, 64, 32, [10, 50, 33, 7]),
    # 16 tiles (the most) at random origins, one channel quad in g0
    "sixteen_q1": (512, 256, 4, 16, 2048, 256, list(np.random.default_rng(5).integers(0, 2048, 32))),
    # small level sides, one tile wrapping
    "tiny": (32, 16, 8, 4, 16, 8, [0, 0, 4, 2, 13, 9]),
}


@pytest.mark.gpu
@pytest.mark.parametrize("kind,Od,blocks", [("c1_bilinear_relu", 9, 253), ("c3_bilinear_h128", 9, 256),
                                            ("c2_bilinear", 4, 37), ("ref_bpp_2_25", 16, 1)])
def test_decoder_update_refreshes_tables(oracle, kind, Od, blocks):
    """ntc_decoder_update_bf16 (the update writes the BF16 tables it changes)
    against ntc_decoder_update followed by a full ntc_pack_decoder_bf16, bit
    for bit, over two updates in a row (Adam moments non-zero the second time)."""
    lib = G.lib()
    Lay, act = _case(oracle, kind)
    fd = oracle.feature_desc(Lay, 0)
    fdc = G.feature_desc(fd)
    H, F = Lay.hidden_dim, Lay.feature_dim
    rng = np.random.default_rng(zlib.crc32(f"tables-{kind}-{Od}".encode()))
    dec = np.concatenate([(rng.uniform(-1, 1, H * F) / np.sqrt(F)).astype(np.float32),
                          (rng.uniform(-1, 1, H * H) / np.sqrt(H)).astype(np.float32),
                          (rng.uniform(-1, 1, 16 * H) / np.sqrt(H)).astype(np.float32)])
    n_dec = dec.size
    n_tab = lib.ntc_fused_table_elems(G.byref(fdc), H)
    sides = []
    for _ in range(2):
        p = G.dev(dec)
        tab = G.empty((n_tab,), np.uint16)
        G.ok(lib.ntc_pack_decoder_bf16(G.byref(fdc), H, Od, G.p(p), G.p(tab), None))
        sides.append(dict(p=p, tab=tab, grad=G.empty((n_dec,), np.float32), m=G.dev(np.zeros(n_dec, np.float32)),
                          v=G.dev(np.zeros(n_dec, np.float32)), sse=G.empty((1,), np.float64)))
    for it, (lr, b1c, b2c) in enumerate(((1e-2, 10.0, 1000.0), (3e-3, 5.26, 500.25))):
        part = G.dev((rng.standard_normal(blocks * n_dec) * 1e-2).astype(np.float32))
        ssep = G.dev(rng.uniform(0, 1, blocks))
        a, b = sides
        G.ok(lib.ntc_decoder_update_bf16(G.byref(fdc), H, Od, blocks, G.p(part), lr, 0.9, 0.999, b1c, b2c,
                                         G.p(a["grad"]), G.p(a["m"]), G.p(a["v"]), G.p(a["p"]), G.p(a["tab"]),
                                         G.p(ssep), G.p(a["sse"]), None))
        G.ok(lib.ntc_decoder_update(H, F, Od, blocks, G.p(part), lr, 0.9, 0.999, b1c, b2c, G.p(b["grad"]),
                                    G.p(b["m"]), G.p(b["v"]), G.p(b["p"]), G.p(ssep), G.p(b["sse"]), None))
        G.ok(lib.ntc_pack_decoder_bf16(G.byref(fdc), H, Od, G.p(b["p"]), G.p(b["tab"]), None))
        for k in ("p", "tab", "grad", "m", "v", "sse"):
            ga, gb = G.host(a[k]), G.host(b[k])
            assert np.array_equal(ga.view(np.uint8), gb.view(np.uint8)), (it, k)
        # the update moved the parameters (the tables are not trivially equal)
        assert not np.array_equal(G.host(a["p"]), dec)


@pytest.mark.parametrize("case", list(_PART_CASES))
def test_grid_update_parts_match_tiles_update(case):
    """The cold part (cells outside every tile rectangle, gradient untouched)
    followed by the hot part (cells inside, gradient read and cleared) leave
    p, m, v and the gradient exactly as the one-pass footprint-aware update
    ntc_grid_update_tiles_f32 does, for a dense random gradient (both read it
    only inside the rectangles); and the cold part alone changes nothing
    inside the rectangles."""
    D0, D1, C0, C1, M, T, org = _PART_CASES[case]
    lib = G.lib()
    rng = np.random.default_rng(3)
    fdc = L.FeatureDesc(D0, D1, C0, C1, 1, 4, 2, 0, 128)
    shapes = [(D0, D0, C0), (D1, D1, C1)]
    init = [[rng.uniform(-1, 1, shp).astype(np.float32), (rng.standard_normal(shp) * 1e-3).astype(np.float32),
             rng.uniform(0, 1e-6, shp).astype(np.float32), (rng.standard_normal(shp) * 1e-3).astype(np.float32)]
            for shp in shapes]  # p, m, v, g
    origins = np.array(org, np.int32)
    og = G.dev(origins)
    tl = G.tiles(og, len(origins) // 2, T, M, 0.0)
    lr, b1c, b2c = 0.01, 1.0 / (1 - 0.9 ** 3), 1.0 / (1 - 0.999 ** 3)
    bounds = (-1.0, 1.0 - 2.0 / 16, -1.0, 1.0 - 2.0 / 16)
    runs = {}
    for label in ("one_pass", "parts", "cold_only"):
        st = [[G.dev(x) for x in grid] for grid in init]
        ad = L.GridAdam(G.p(st[0][1]), G.p(st[0][2]), G.p(st[1][1]), G.p(st[1][2]), lr, 0.9, 0.999, b1c, b2c,
                        *bounds)
        args = (G.byref(fdc), G.byref(tl), G.p(st[0][0]), G.p(st[0][3]), G.p(st[1][0]), G.p(st[1][3]), G.byref(ad))
        if label == "one_pass":
            G.ok(lib.ntc_grid_update_tiles_f32(*args, None), "tiles update")
        else:
            G.ok(lib.ntc_grid_update_tiles_part_f32(*args, 1, 0, None), "cold part")
            if label == "parts":
                G.ok(lib.ntc_grid_update_tiles_part_f32(*args, 2, 0, None), "hot part")
        runs[label] = [[G.host(x) for x in grid] for grid in st]
    for k in range(2):
        for j, name in enumerate("pmvg"):
            got, want = runs["parts"][k][j], runs["one_pass"][k][j]
            assert np.array_equal(got, want), (case, k, name, int(np.count_nonzero(got != want)))
        # inside the rectangles the cold part left everything as it was; the
        # one-pass update cleared the gradient there and only there
        hot = runs["one_pass"][k][3] != init[k][3]
        assert hot.any() and not hot.all()
        for j in range(4):
            assert np.array_equal(runs["cold_only"][k][j][hot], init[k][j][hot])
        assert np.array_equal(runs["cold_only"][k][3], init[k][3])


@pytest.mark.parametrize("pipelined", [False, True, "merged", "across"])
@pytest.mark.parametrize("case", list(_PART_CASES))
def test_grid_lazy_matches_eager(case, pipelined):
    """Twelve steps with fresh tile origins, gradients and Adam constants per
    step: the one-pass footprint-aware update of every step (eager) against
    the deferred form (catch-up of the step's footprint cells, the hot step
    there, the other cells' zero-gradient steps applied only at a flush).
    Pipelined: the catch-up of step k + 1 leaves out step k's rectangles and
    runs before step k's hot part (as it may beside it). Merged: step k's hot
    part and step k + 1's catch-up in one launch
    (ntc_grid_lazy_hot_catchup_f32). Across: pipelined, also when the flush
    at step 7 comes between the catch-up of step 7 and step 7 itself (the
    catch-up records the caught-up cells' counters, so the flush leaves them
    alone). A third of each grid starts with zero Adam state. After the flush
    at step 7 and at the end, p, m, v and the gradient agree bit for bit, and
    the cell counters hold the step count."""
    D0, D1, C0, C1, M, T, org = _PART_CASES[case]
    n_tiles, K, flush_at = len(org) // 2, 12, 7
    lib = G.lib()
    rng = np.random.default_rng(17)
    fdc = L.FeatureDesc(D0, D1, C0, C1, 1, 4, 2, 0, 128)
    shapes = [(D0, D0, C0), (D1, D1, C1)]
    init = [[rng.uniform(-1, 1, shp).astype(np.float32), (rng.standard_normal(shp) * 1e-3).astype(np.float32),
             rng.uniform(0, 1e-6, shp).astype(np.float32), np.zeros(shp, np.float32)] for shp in shapes]
    for p0, m0, v0, _ in init:
        # Adam state still zero (cells no gradient has reached: the catch-up
        # owes them only the clamp, p above the upper bound included), with
        # -0 in places, and rows where only m or only v is zero
        D = p0.shape[0]
        m0[: D // 3] = 0.0
        v0[: D // 3] = 0.0
        m0[: D // 6, ::3] = -0.0
        m0[D // 3: D // 2] = 0.0
        v0[D // 2: 2 * D // 3] = 0.0
    consts = np.stack([rng.uniform(0.002, 0.01, K), 1.0 / (1 - 0.9 ** np.arange(1, K + 1)),
                       1.0 / (1 - 0.999 ** np.arange(1, K + 1))], 1).astype(np.float32)
    bounds = (-1.0, 1.0 - 2.0 / 16, -1.0, 1.0 - 2.0 / 16)
    eager = [[G.dev(x) for x in grid] for grid in init]
    lazy = [[G.dev(x) for x in grid] for grid in init]
    last = [G.dev(np.zeros(D0 * D0, np.int32)), G.dev(np.zeros(D1 * D1, np.int32))]
    tab = G.dev(consts.ravel())

    def adam(st, k):
        lr, b1c, b2c = (float(x) for x in consts[k])
        return L.GridAdam(G.p(st[0][1]), G.p(st[0][2]), G.p(st[1][1]), G.p(st[1][2]), lr, 0.9, 0.999, b1c, b2c,
                          *bounds)

    def flush(upto):
        ad = adam(lazy, 0)
        G.ok(lib.ntc_grid_lazy_catchup_f32(G.byref(fdc), None, None, G.p(lazy[0][0]), G.p(lazy[1][0]), G.byref(ad),
                                           G.p(tab), upto, G.p(last[0]), G.p(last[1]), None), "flush")
        for k in range(2):
            G.ok(lib.ntc_fill_i32((D0, D1)[k] ** 2, upto, G.p(last[k]), None), "mark")

    def compare(step):
        for k in range(2):
            for j, name in enumerate("pmvg"):
                got, want = G.host(lazy[k][j]), G.host(eager[k][j])
                assert np.array_equal(got, want), (case, step, k, name, int(np.count_nonzero(got != want)))
            assert (G.host(last[k]) == step).all()

    ogs = [G.dev(rng.integers(0, M, 2 * n_tiles).astype(np.int32)) for _ in range(K)]
    tls = [G.tiles(og, n_tiles, T, M, 0.0) for og in ogs]

    def catch_up(k, skip):
        ad = adam(lazy, 0)
        G.ok(lib.ntc_grid_lazy_catchup_f32(G.byref(fdc), G.byref(tls[k]), G.byref(skip) if skip else None,
                                           G.p(lazy[0][0]), G.p(lazy[1][0]), G.byref(ad), G.p(tab), k,
                                           G.p(last[0]), G.p(last[1]), None), "catch-up")

    issued = -1
    for k in range(K):
        grads = [(rng.standard_normal(shp) * 1e-3).astype(np.float32) for shp in shapes]
        for st in (eager, lazy):
            for g in range(2):
                G.ok(lib.ntc_dev_upload(G.p(st[g][3]), grads[g].ctypes.data, grads[g].nbytes), "upload")
        ad = adam(eager, k)
        G.ok(lib.ntc_grid_update_tiles_f32(G.byref(fdc), G.byref(tls[k]), G.p(eager[0][0]), G.p(eager[0][3]),
                                           G.p(eager[1][0]), G.p(eager[1][3]), G.byref(ad), None), "eager")
        if issued != k:
            catch_up(k, None)
        ahead = pipelined and k + 1 < K and (k + 1 != flush_at or pipelined == "across")
        if ahead and pipelined != "merged":
            catch_up(k + 1, tls[k])
            issued = k + 1
        ad = adam(lazy, k)
        if ahead and pipelined == "merged":
            ad0 = adam(lazy, 0)
            G.ok(lib.ntc_grid_lazy_hot_catchup_f32(
                G.byref(fdc), G.byref(tls[k]), G.p(lazy[0][0]), G.p(lazy[0][3]), G.p(lazy[1][0]),
                G.p(lazy[1][3]), G.byref(ad), k, G.p(last[0]), G.p(last[1]), G.byref(fdc), G.byref(tls[k + 1]),
                G.byref(tls[k]), G.p(lazy[0][0]), G.p(lazy[1][0]), G.byref(ad0), G.p(tab), k + 1, G.p(last[0]),
                G.p(last[1]), None), "hot + catch-up")
            issued = k + 1
        else:
            G.ok(lib.ntc_grid_lazy_hot_f32(G.byref(fdc), G.byref(tls[k]), G.p(lazy[0][0]), G.p(lazy[0][3]),
                                           G.p(lazy[1][0]), G.p(lazy[1][3]), G.byref(ad), k, G.p(last[0]),
                                           G.p(last[1]), None), "hot")
        if k + 1 == flush_at:
            flush(k + 1)
            compare(k + 1)
    # before the final flush some cells are behind
    assert any((G.host(last[g]) < K).any() for g in range(2))
    flush(K)
    compare(K)


@pytest.mark.parametrize("seed", [0, 1])
def test_grid_lazy_flush_matches_eager_at_extreme_magnitudes(seed):
    """The catch-up's short forms (the square root without its scaled range,
    the division without operand scaling and fix-up, each taken only where
    bounds over the wave's steps allow) give the bits of the eager update:
    40 zero-gradient steps of every cell (the one-pass update with a zero
    gradient) against one flush from step 0, for Adam states spread over
    the whole FP32 range (m from denormal to 1e8 and zero of either sign,
    v from denormal to 1e36, where v bias2 overflows) and learning rates
    from 0 and 1e-38 to 1e10, with runs of ordinary values in between so
    that whole waves take the short forms."""
    D0, D1, C0, C1, M, T, org = _PART_CASES["coarse_q3_q2"]
    n_tiles, K = len(org) // 2, 40
    lib = G.lib()
    rng = np.random.default_rng(100 + seed)
    fdc = L.FeatureDesc(D0, D1, C0, C1, 1, 4, 2, 0, 128)
    shapes = [(D0, D0, C0), (D1, D1, C1)]

    def spread(shp, lo, hi, signed):
        x = (10.0 ** rng.uniform(lo, hi, shp)).astype(np.float32)
        if signed:
            x *= rng.choice(np.array([-1.0, 1.0], np.float32), shp)
        return x

    init = []
    for shp in shapes:
        p0 = rng.uniform(-1, 1, shp).astype(np.float32)
        m0 = spread(shp, -44, 8, True)
        v0 = spread(shp, -45, 36, False)
        D = shp[0]
        # ordinary rows (every wave of them takes the short forms), zero rows
        m0[: D // 2] = (rng.standard_normal((D // 2,) + shp[1:]) * 1e-3).astype(np.float32)
        v0[: D // 2] = rng.uniform(1e-9, 1e-6, (D // 2,) + shp[1:]).astype(np.float32)
        m0[D // 2: D // 2 + 4] = -0.0
        v0[D // 2 + 4: D // 2 + 8] = 0.0
        init.append([p0, m0, v0, np.zeros(shp, np.float32)])
    lrs = rng.uniform(0.002, 0.01, K)
    lrs[[3, 17]] = 0.0
    lrs[[5, 29]] = 1e-38
    lrs[[8, 33]] = 1e10
    lrs[21] = 3e-20
    consts = np.stack([lrs, 1.0 / (1 - 0.9 ** np.arange(1, K + 1)),
                       1.0 / (1 - 0.999 ** np.arange(1, K + 1))], 1).astype(np.float32)
    bounds = (-1.0, 1.0 - 2.0 / 16, -1.0, 1.0 - 2.0 / 16)
    eager = [[G.dev(x) for x in grid] for grid in init]
    lazy = [[G.dev(x) for x in grid] for grid in init]
    last = [G.dev(np.zeros(D0 * D0, np.int32)), G.dev(np.zeros(D1 * D1, np.int32))]
    tab = G.dev(consts.ravel())
    og = G.dev(np.array(org, np.int32))
    tl = G.tiles(og, n_tiles, T, M, 0.0)

    def adam(st, k):
        lr, b1c, b2c = (float(x) for x in consts[k])
        return L.GridAdam(G.p(st[0][1]), G.p(st[0][2]), G.p(st[1][1]), G.p(st[1][2]), lr, 0.9, 0.999, b1c, b2c,
                          *bounds)

    for k in range(K):
        ad = adam(eager, k)
        G.ok(lib.ntc_grid_update_tiles_f32(G.byref(fdc), G.byref(tl), G.p(eager[0][0]), G.p(eager[0][3]),
                                           G.p(eager[1][0]), G.p(eager[1][3]), G.byref(ad), None), "eager")
    ad = adam(lazy, 0)
    G.ok(lib.ntc_grid_lazy_catchup_f32(G.byref(fdc), None, None, G.p(lazy[0][0]), G.p(lazy[1][0]), G.byref(ad),
                                       G.p(tab), K, G.p(last[0]), G.p(last[1]), None), "flush")
    for g in range(2):
        for j, name in enumerate("pmv"):
            got, want = G.host(lazy[g][j]), G.host(eager[g][j])
            # NaN-free states: compare the bits
            assert not np.isnan(want).any()
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (
                seed, g, name, int(np.count_nonzero(got.view(np.uint32) != want.view(np.uint32))))
        # the extreme rows moved p too (the steps were not all no-ops)
        assert not np.array_equal(G.host(lazy[g][0]), init[g][0])


@pytest.mark.parametrize("kind,lod,n_tiles", [("c1_bilinear_relu", 0, 8), ("c1_bilinear_relu", 1, 4),
                                              ("c2_bilinear", 0, 8), ("c1_bilinear_relu", 2, 8)])
def test_fused_sets_partials_repeat_bit_for_bit(oracle, kind, lod, n_tiles):
    """Each workgroup's decoder-gradient and SSE partials depend only on its own
    passes, so repeated multi-set launches on the same inputs give the same
    bits, at every pass size and with more workgroups than CUs (two per CU
    where registers and LDS allow); only the latent-gradient atomics may sum
    in another order."""
    lib = G.lib()
    cs = _step_case(oracle, kind, lod, n_tiles)
    sizes = (C.c_int32 * 2)()
    n = lib.ntc_fused_pass_sizes(G.byref(cs["fdc"]), cs["H"], cs["act"], sizes)
    maxb = C.c_int32()
    G.ok(lib.ntc_fused_max_blocks(C.byref(maxb)))
    rb = lib.ntc_fused_params_bytes()
    for pass_texels in list(sizes)[:n]:
        passes = -(-cs["B"] // pass_texels)
        counts = [min(passes, maxb.value), min(passes, maxb.value)]
        outs = [_outputs(cs, c) for c in counts]
        rec = np.zeros(2 * rb, np.uint8)
        for s in range(2):
            a = cs["a"]
            o = outs[s]
            a.dg0, a.dg1, a.dw_partials, a.sse_partials, a.blocks = (G.p(o["dg0"]), G.p(o["dg1"]), G.p(o["part"]),
                                                                     G.p(o["ssep"]), counts[s])
            G.ok(lib.ntc_fused_params_pack(G.byref(a), rec.ctypes.data + s * rb))
        first = G.dev(np.array([0, counts[0], counts[0] + counts[1]], np.int32))
        d_rec, step = G.dev(rec), G.dev(np.zeros(1, np.int32))
        ref = None
        for rep in range(6):
            for o in outs:
                for k in ("dg0", "dg1"):
                    G.ok(lib.ntc_fill_zero(G.p(o[k]), o[k].nbytes, None))
            G.ok(lib.ntc_fused_train_sets(G.byref(cs["fdc"]), cs["H"], cs["act"], int(pass_texels), G.p(d_rec),
                                          G.p(first), 2, G.p(step), counts[0] + counts[1], None))
            got = [(G.host(o["part"]), G.host(o["ssep"]), G.host(o["dg0"]), G.host(o["dg1"])) for o in outs]
            if ref is None:
                ref = got
                # both sets saw the same inputs: the same partials, slot for slot
                assert np.array_equal(got[0][0], got[1][0]) and np.array_equal(got[0][1], got[1][1])
                continue
            for s in range(2):
                bad = int(np.count_nonzero(got[s][0] != ref[s][0]))
                assert bad == 0, (pass_texels, rep, s, bad)
                assert np.array_equal(got[s][1], ref[s][1]), (pass_texels, rep, s)
                # latent gradients: the same sums up to the atomics' order
                for k in (2, 3):
                    scale = float(np.max(np.abs(ref[s][k]))) + 1e-30
                    err = float(np.max(np.abs(got[s][k] - ref[s][k]))) / scale
                    assert err < 1e-5, (pass_texels, rep, s, k, err)


def _sets_step_outputs(lib, cs, pass_texels, workgroups=0):
    """One single-set multi-set launch at a pass size (on `workgroups`
    workgroups, default as many as passes up to one per CU); returns the
    reduced decoder gradient and the latent gradients."""
    rb = lib.ntc_fused_params_bytes()
    passes = -(-cs["B"] // pass_texels)
    maxb = C.c_int32()
    G.ok(lib.ntc_fused_max_blocks(C.byref(maxb)))
    count = min(passes, workgroups or maxb.value)
    out = _outputs(cs, count)
    a = cs["a"]
    a.dg0, a.dg1, a.dw_partials, a.sse_partials, a.blocks = (G.p(out["dg0"]), G.p(out["dg1"]), G.p(out["part"]),
                                                             G.p(out["ssep"]), count)
    rec = np.zeros(rb, np.uint8)
    G.ok(lib.ntc_fused_params_pack(G.byref(a), rec.ctypes.data))
    first = np.array([0, count], np.int32)
    G.ok(lib.ntc_fused_train_sets(G.byref(cs["fdc"]), cs["H"], cs["act"], int(pass_texels), G.p(G.dev(rec)),
                                  G.p(G.dev(first)), 1, G.p(G.dev(np.zeros(1, np.int32))), count, None))
    n_dec = cs["dec"].size
    grad = G.empty((n_dec,), np.float32)
    m = G.empty((n_dec,), np.float32)
    v = G.empty((n_dec,), np.float32)
    sse_out = G.empty((1,), np.float64)
    G.ok(lib.ntc_decoder_update(cs["H"], cs["F"], cs["Od"], count, G.p(out["part"]), 0.0, 0.9, 0.999, 1.0, 1.0,
                                G.p(grad), G.p(m), G.p(v), G.p(G.dev(cs["dec"])), G.p(out["ssep"]), G.p(sse_out),
                                None))
    return G.host(grad), G.host(out["dg0"]), G.host(out["dg1"]), float(G.host(sse_out)[0])


@pytest.mark.parametrize("kind", ["c1_bilinear_relu", "c2_bilinear"])
@pytest.mark.parametrize("lod,n_tiles", [(0, 2), (1, 1), (2, 3), (3, 4), (4, 5), (5, 2), (6, 7)])
def test_fused_pass_sizes_agree_elementwise(oracle, kind, lod, n_tiles):
    """The 64- and 128-texel pass kernels form the same BF16 products per
    texel and differ only in how the FP32 sums are grouped, so on the same
    inputs every decoder-gradient and latent-gradient element agrees to FP32
    rounding (relative to the largest element)."""
    lib = G.lib()
    cs = _step_case(oracle, kind, lod, n_tiles)
    sizes = (C.c_int32 * 2)()
    assert lib.ntc_fused_pass_sizes(G.byref(cs["fdc"]), cs["H"], cs["act"], sizes) == 2
    r128 = _sets_step_outputs(lib, cs, 128)
    r64 = _sets_step_outputs(lib, cs, 64)
    assert abs(r64[3] - r128[3]) <= 1e-9 * r128[3]
    for k, name in ((0, "dW"), (1, "dG0"), (2, "dG1")):
        scale = float(np.max(np.abs(r128[k]))) + 1e-30
        err = np.abs(r64[k] - r128[k]) / scale
        worst = int(np.argmax(err))
        assert float(err.max()) < 1e-5, (name, float(err.max()), worst, float(r64[k].ravel()[worst]),
                                         float(r128[k].ravel()[worst]))


@pytest.mark.parametrize("dim,lod,n_tiles", [(1024, 0, 1), (1024, 0, 3), (1024, 1, 2), (1024, 2, 1), (1024, 3, 2),
                                             (2048, 0, 1), (2048, 0, 2), (2048, 1, 1), (2048, 2, 2), (2048, 3, 1)])
def test_fused_pass_sizes_agree_on_large_mips(oracle, dim, lod, n_tiles):
    """The same element-wise agreement of the 64- and 128-texel kernels on
    the batched configuration's layout with 256^2 tiles inside larger mips
    (tiles that stop short of the mip edge as well as wrapping ones), and the
    64-texel step against both references."""
    lib = G.lib()
    cs = _step_case(oracle, "c1_bilinear_relu", lod, n_tiles, dim)
    r128 = _sets_step_outputs(lib, cs, 128)
    r64 = _sets_step_outputs(lib, cs, 64)
    assert abs(r64[3] - r128[3]) <= 1e-9 * r128[3]
    for k, name in ((0, "dW"), (1, "dG0"), (2, "dG1")):
        scale = float(np.max(np.abs(r128[k]))) + 1e-30
        err = np.abs(r64[k] - r128[k]) / scale
        worst = int(np.argmax(err))
        assert float(err.max()) < 1e-5, (name, float(err.max()), worst, float(r64[k].ravel()[worst]),
                                         float(r128[k].ravel()[worst]))


@pytest.mark.parametrize("dim,lod,n_tiles", [(256, 0, 1), (1024, 0, 1), (2048, 0, 1), (2048, 1, 2), (256, 3, 4)])
@pytest.mark.parametrize("workgroups", [16, 5, 1])
def test_fused_pass_sizes_agree_with_many_passes_per_workgroup(oracle, dim, lod, n_tiles, workgroups):
    """As above with few workgroups, so that each runs many passes through the
    cross-pass prefetch pipeline (a set of the batched configuration gets
    about 16 workgroups of a launch)."""
    lib = G.lib()
    cs = _step_case(oracle, "c1_bilinear_relu", lod, n_tiles, dim)
    r128 = _sets_step_outputs(lib, cs, 128, workgroups)
    r64 = _sets_step_outputs(lib, cs, 64, workgroups)
    full = _sets_step_outputs(lib, cs, 128)
    assert abs(r64[3] - r128[3]) <= 1e-9 * r128[3]
    for k, name in ((0, "dW"), (1, "dG0"), (2, "dG1")):
        scale = float(np.max(np.abs(full[k]))) + 1e-30
        for got, label in ((r64[k], "64"), (r128[k], "128")):
            err = np.abs(got - full[k]) / scale
            worst = int(np.argmax(err))
            assert float(err.max()) < 1e-5, (name, label, float(err.max()), worst, float(got.ravel()[worst]),
                                             float(full[k].ravel()[worst]))


@pytest.mark.parametrize("lod,n_tiles", [(7, 1), (8, 1), (7, 3), (8, 2)])
def test_fused_smallest_mips(oracle, lod, n_tiles):
    """Batches of 1-12 texels (the 2x2 and 1x1 mips of a full chain): both
    pass sizes against both references and against each other."""
    lib = G.lib()
    cs = _step_case(oracle, "c1_bilinear_relu", lod, n_tiles, 256, 1)
    assert cs["B"] == n_tiles * (256 >> lod) ** 2
    r128 = _sets_step_outputs(lib, cs, 128)
    r64 = _sets_step_outputs(lib, cs, 64)
    assert abs(r64[3] - r128[3]) <= 1e-9 * r128[3]
    for k, name in ((0, "dW"), (1, "dG0"), (2, "dG1")):
        scale = float(np.max(np.abs(r128[k]))) + 1e-30
        err = np.abs(r64[k] - r128[k]) / scale
        assert float(err.max()) < 1e-5, (name, float(err.max()))
    bsse, bw0, bw1, bwo, bdg0, bdg1 = cs["bref"]
    H, F, Od = cs["H"], cs["F"], cs["Od"]
    for r, label in ((r64, "64"), (r128, "128")):
        g = r[0]
        gw0, gw1, gwo = g[: H * F].reshape(H, F), g[H * F: H * F + H * H].reshape(H, H), g[H * F + H * H:].reshape(16, H)
        assert abs(r[3] - bsse) <= 1e-4 * bsse, label
        for got, want, name in ((gw0, bw0, "dW0"), (gw1, bw1, "dW1"), (gwo[:Od], bwo, "dWout"), (r[1], bdg0, "dG0"),
                                (r[2], bdg1, "dG1")):
            assert rel(got, want) < REL_TOL_BF16, (label, name, rel(got, want))
        for got, sup, name in ((r[1], cs["support"][0], "dG0"), (r[2], cs["support"][1], "dG1")):
            assert not got[~sup].any(), (label, name)
