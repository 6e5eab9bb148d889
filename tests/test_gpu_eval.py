"""Evaluation decode at reduced precision on the GPU (ntc_decode_sse_f32 and
Context::EvalEx) against the CPU oracle's restatement of the same
arithmetic: outputs bit-exact per precision (FP32, BF16, FP8 E4M3 with
per-layer scales, INT8 with the deployed decoder's semantics), per-channel
squared errors equal up to FP64 summation order, and every precision's PSNR
stated against the FP64 decode of the same state."""
import numpy as np
import pytest

import openntc_amd as ntc
from openntc_amd import _lib as L

pytestmark = pytest.mark.gpu

from tests import gpu_util as G  # noqa: E402
from tests.test_oracle import _int8_blob  # noqa: E402


def _decoder(rng, F, H, O):
    w0 = (rng.uniform(-1, 1, (H, F)) / np.sqrt(F)).astype(np.float32)
    w1 = (rng.uniform(-1, 1, (H, H)) / np.sqrt(H)).astype(np.float32)
    wo = np.zeros((16, H), np.float32)
    wo[:O] = rng.uniform(-1, 1, (O, H)) / np.sqrt(H)
    return w0, w1, wo, np.concatenate([w0.ravel(), w1.ravel(), wo.ravel()])


@pytest.mark.parametrize("H,F", [(64, 36), (128, 40)])
@pytest.mark.parametrize("act", [1, 0])
@pytest.mark.parametrize("precision", [0, 1, 2, 3])
def test_decode_kernel_matches_oracle(oracle, precision, act, H, F):
    O, B = 9, 3 * 256 + 77
    Fp = (F + 15) // 16 * 16
    rng = np.random.default_rng(100 * precision + 10 * act + H)
    w0, w1, wo, dec = _decoder(rng, F, H, O)
    x = rng.uniform(-1, 1, (F, B)).astype(np.float32)
    yt = rng.uniform(0, 1, (O, B)).astype(np.float32)
    lib = G.lib()
    table = np.zeros(lib.ntc_decode_table_bytes(precision, F, Fp, H), np.uint8)
    if precision == 3:
        s_a1, s_a2 = np.float32(1.1 * 1.2 / 127), np.float32(1.1 * 0.8 / 127)
        blob = _int8_blob(oracle, w0, w1, wo[:O], Fp, s_a1, s_a2)
        rcp = np.array([np.float32(1) / s_a1, np.float32(1) / s_a2], np.float32)
        G.ok(lib.ntc_build_decode_table_int8(F, Fp, H, O, blob.ctypes.data, float(rcp[0]), float(rcp[1]),
                                             table.ctypes.data))
        want = oracle.decode(3, act, F, Fp, H, O, x, blob=blob, rcp=rcp)
    else:
        ma0, ma1 = 1.7, 2.3
        G.ok(lib.ntc_build_decode_table_f(precision, F, H, O, dec.ctypes.data, ma0, ma1, table.ctypes.data))
        s8 = oracle.fp8_scales(F, H, O, dec, ma0, ma1)
        want = oracle.decode(precision, act, F, Fp, H, O, x, dec=dec, s8=s8)
    y = G.empty((O, B), np.float32)
    sse = G.empty((16,), np.float64)
    ws = G.empty((lib.ntc_decode_workspace_bytes(B),), np.uint8)
    G.ok(lib.ntc_decode_sse_f32(precision, act, F, Fp, H, O, B, G.p(G.dev(table)), G.p(G.dev(x)), G.p(G.dev(yt)),
                                G.p(sse), G.p(y), G.p(ws), None), "ntc_decode_sse_f32")
    got = G.host(y)
    assert np.array_equal(got, want), (precision, int(np.count_nonzero(got != want)))
    r = (yt - want).astype(np.float64)
    assert np.allclose(G.host(sse)[:O], np.sum(r * r, axis=1), rtol=1e-12, atol=0)
    # FP64 decode of the same weights: the precision's error is bounded
    d64 = oracle.decode_f64(act, F, H, O, x, dec.astype(np.float64))
    rel = np.max(np.abs(got - d64)) / np.max(np.abs(d64))
    assert rel < {0: 1e-5, 1: 3e-2, 2: 2e-1, 3: 2e-1}[precision], rel


@pytest.mark.parametrize("H,F", [(64, 36), (128, 40), (64, 92)])
@pytest.mark.parametrize("act", [1, 0])
@pytest.mark.parametrize("mfma,chains", [(4, 1), (5, 2)])
def test_mfma_decode_vs_chains(oracle, mfma, chains, act, H, F):
    """The decode on the matrix cores (4: BF16, 5: FP8 E4M3) has the rounding
    points (and FP8 scales) of the VALU decode of the same precision and
    differs from its fixed-order FMA chains only by FP32 accumulation order
    (a hidden activation on a rounding boundary may round the other way; an
    E4M3 step is 1/8 of the value): outputs close to the chains in norm, most
    within a few FP32 ulps; the squared errors are those of its own outputs;
    the error against the FP64 decode is the precision's."""
    O, B = 9, 5 * 256 + 77
    Fp = (F + 15) // 16 * 16
    rng = np.random.default_rng(500 + 10 * act + H + F + mfma)
    w0, w1, wo, dec = _decoder(rng, F, H, O)
    x = rng.uniform(-1, 1, (F, B)).astype(np.float32)
    yt = rng.uniform(0, 1, (O, B)).astype(np.float32)
    lib = G.lib()
    ma0, ma1 = 1.7, 2.3
    table = np.zeros(lib.ntc_decode_table_bytes(mfma, F, Fp, H), np.uint8)
    G.ok(lib.ntc_build_decode_table_f(mfma, F, H, O, dec.ctypes.data, ma0, ma1, table.ctypes.data))
    s8 = oracle.fp8_scales(F, H, O, dec, ma0, ma1)
    want = oracle.decode(chains, act, F, Fp, H, O, x, dec=dec, s8=s8)
    y = G.empty((O, B), np.float32)
    sse = G.empty((16,), np.float64)
    ws = G.empty((lib.ntc_decode_workspace_bytes(B),), np.uint8)
    G.ok(lib.ntc_decode_sse_f32(mfma, act, F, Fp, H, O, B, G.p(G.dev(table)), G.p(G.dev(x)), G.p(G.dev(yt)),
                                G.p(sse), G.p(y), G.p(ws), None), "ntc_decode_sse_f32")
    got = G.host(y)
    diff = np.abs(got - want)
    norm_tol = {4: 1e-3, 5: 1e-2}[mfma]
    assert np.linalg.norm(got - want) < norm_tol * np.linalg.norm(want), np.linalg.norm(got - want)
    if mfma == 4:
        ulp = np.spacing(np.maximum(np.abs(want), 1e-3).astype(np.float32))
        assert np.mean(diff <= 8 * ulp) > 0.97, np.mean(diff <= 8 * ulp)
    else:
        # the FP8 chains sum scaled E4M3 terms of both signs: the order moves
        # small outputs by many of their own ulps, so bound by the output range
        assert diff.max() < 2e-2 * np.abs(want).max(), diff.max() / np.abs(want).max()
    r = (yt - got).astype(np.float64)
    assert np.allclose(G.host(sse)[:O], np.sum(r * r, axis=1), rtol=1e-12, atol=0)
    assert not np.any(G.host(sse)[O:])
    d64 = oracle.decode_f64(act, F, H, O, x, dec.astype(np.float64))
    assert np.max(np.abs(got - d64)) / np.max(np.abs(d64)) < {4: 3e-2, 5: 2e-1}[mfma]


def _walk(oracle, Lay, pk, grids):
    """The evaluation walk's batches: (features, targets) per mip, as the
    oracle computes them from the given latents."""
    for mip in range(Lay.mip_count):
        M = Lay.mip_dim[mip]
        T = min(M, 256)
        per = (M + T - 1) // T
        lvl = oracle.level_for_lod(Lay, mip)
        fd = oracle.feature_desc(Lay, lvl)
        for start in range(0, per * per, 8):
            tiles = range(start, min(start + 8, per * per))
            origins = np.array([v for g in tiles for v in ((g % per) * T, (g // per) * T)], np.int32)
            x = oracle.draw_features(fd, origins, T, M, np.float32(mip / (Lay.mip_count - 1)), *grids[lvl])
            yt = oracle.draw_targets(origins, T, M, Lay.out_dim, pk(mip))
            yield x, yt


def _oracle_sse(oracle, Lay, pk, grids, dec, precisions, s8=None, blob=None, rcp=None):
    F, Fp, H, O = Lay.feature_dim, Lay.feature_dim_padded, Lay.hidden_dim, Lay.out_dim
    sse = {p: np.zeros(O) for p in precisions}
    count = 0
    for x, yt in _walk(oracle, Lay, pk, grids):
        for p in precisions:
            if p == "f64":
                y = oracle.decode_f64(Lay.activation, F, H, O, x.astype(np.float64), dec.astype(np.float64))
                r = yt.astype(np.float64) - y
            else:
                y = oracle.decode(p, Lay.activation, F, Fp, H, O, x, dec=dec, s8=s8, blob=blob, rcp=rcp)
                r = (yt - y).astype(np.float64)
            sse[p] += np.sum(r * r, axis=1)
        count += x.shape[1]
    return sse, count


def _fp8_scales(oracle, Lay, pk, grids, dec):
    P2 = oracle.Params(Lay)
    for l in range(Lay.level_count):
        P2.g0[l][...], P2.g1[l][...] = grids[l]
    P2.dec[...] = dec
    mx = oracle.Trainer(Lay, P2, pk, max_batch=65536).calibrate()
    return P2, oracle.fp8_scales(Lay.feature_dim, Lay.hidden_dim, Lay.out_dim, dec, mx[0], mx[1])


PRECISIONS = ((0, ntc.EvalPrecision.FP32), (1, ntc.EvalPrecision.BF16), (2, ntc.EvalPrecision.FP8),
              (3, ntc.EvalPrecision.INT8))


def test_context_eval_precisions_match_oracle(oracle):
    """A trained 256^2 context evaluated at every precision over all mips
    against the oracle's decode of the context's own trained state; then the
    decoder is changed and the cached reduced-precision tables must follow."""
    from tests.test_gpu_context import _make
    steps = 60
    ctx, Lay, P, sched, pk = _make(oracle, steps, ntc.TrainPath.Plain, grids=1)
    assert ctx.TrainUntilComplete().phase == ntc.TrainPhase.TrainComplete
    g0b, g1b, blob, cal = ctx.GetCompressedData()
    dec = ctx.read_decoder()
    O = Lay.out_dim
    grids = [(ctx.read_grid(0, l), ctx.read_grid(1, l)) for l in range(Lay.level_count)]
    P2, s8 = _fp8_scales(oracle, Lay, pk, grids, dec)
    rcp = np.array([np.float32(1) / np.float32(cal[2]), np.float32(1) / np.float32(cal[3])], np.float32)
    sse, count = _oracle_sse(oracle, Lay, lambda m: pk[m], grids, dec, (0, 1, 2, 3, "f64"), s8, blob, rcp)
    psnr64 = -10 * np.log10(sse["f64"].sum() / (count * O))
    # the oracle's own FP32 eval walk agrees with the per-batch restatement
    _, ocs = oracle.Trainer(Lay, P2, pk, max_batch=65536).eval(channel_sse=True)
    assert np.allclose(ocs, sse[0], rtol=1e-12, atol=0)
    got = {}
    for p, prec in PRECISIONS:
        ev = ctx.EvalEx(prec)
        got[p] = ev.channel_sse[:O].copy()
        assert np.allclose(ev.channel_sse[:O], sse[p], rtol=1e-10, atol=0), (prec, ev.channel_sse[:O], sse[p])
        want = -10 * np.log10(sse[p].sum() / (count * O))
        assert abs(ev.psnr - want) < 1e-8, (prec, ev.psnr, want)
        print(f"{prec.name}: psnr {ev.psnr:.4f} dB, vs FP64 decode {ev.psnr - psnr64:+.4f} dB")
        assert abs(ev.psnr - psnr64) < {0: 1e-3, 1: 0.2, 2: 1.5, 3: 1.5}[p]
    # the BF16 decode on the matrix cores: BF16's rounding points in the MFMA's
    # accumulation order
    ev = ctx.EvalEx(ntc.EvalPrecision.BF16_MFMA)
    assert np.allclose(ev.channel_sse[:O], sse[1], rtol=2e-3, atol=0), (ev.channel_sse[:O], sse[1])
    assert abs(ev.psnr - psnr64) < 0.2
    ev = ctx.EvalEx(ntc.EvalPrecision.FP8_MFMA)
    assert np.allclose(ev.channel_sse[:O], sse[2], rtol=2e-2, atol=0), (ev.channel_sse[:O], sse[2])
    assert abs(ev.psnr - psnr64) < 1.5
    # switching precisions rebuilds the tables: a second BF16 walk repeats the first
    assert np.array_equal(ctx.EvalEx(ntc.EvalPrecision.BF16).channel_sse[:O], got[1])
    # a new decoder invalidates the cached tables (BF16 and FP8, whose scales
    # come from a fresh calibration walk)
    dec2 = (dec * np.float32(0.75)).astype(np.float32)
    ctx.write_decoder(dec2)
    _, s8b = _fp8_scales(oracle, Lay, pk, grids, dec2)
    sse2, _ = _oracle_sse(oracle, Lay, lambda m: pk[m], grids, dec2, (1, 2), s8b)
    for p, prec in PRECISIONS[1:3]:
        ev = ctx.EvalEx(prec)
        assert np.allclose(ev.channel_sse[:O], sse2[p], rtol=1e-10, atol=0), (prec, ev.channel_sse[:O], sse2[p])


def test_eval_4096_precisions_vs_fp64_oracle(oracle):
    """The evaluation workload (the 2048^2 configuration's layout trained on a
    4096^2 set, every mip down to 1x1, 22.37 M texels): each precision's
    PSNR per channel group against the FP64 CPU decode of the same trained
    state, and FP32 per-channel errors equal to it up to FP32 rounding."""
    import bench
    w = bench.EVAL_WORKLOAD
    ctx, _ = bench.make_context(ntc.TrainPath.Fused, **w)
    assert ctx.TrainUntilComplete().phase == ntc.TrainPhase.TrainComplete
    ctx.GetCompressedData()
    lay = ctx.layout()
    lw = w["layout"]
    Lay = oracle.layout_custom(lw["g0_bits"], lw["g1_bits"], lw["g0_channels"], lw["g1_channels"], lw["g0_scale"],
                               lw["hidden"], 1, 1, w["dim"], 9, 1)
    assert Lay.mip_count == lay.mip_count and Lay.feature_dim == lay.feature_dim
    O = Lay.out_dim
    grids = [(ctx.read_grid(0, l), ctx.read_grid(1, l)) for l in range(Lay.level_count)]
    dec = ctx.read_decoder()
    sse, count = _oracle_sse(oracle, Lay, ctx.read_package, grids, dec, ("f64",))
    ref = bench.group_psnr(sse["f64"], count)
    ref_all = -10 * np.log10(sse["f64"].sum() / (count * O))
    tol = {0: 1e-3, 1: 0.2, 2: 3.0, 3: 3.0, 4: 0.2, 5: 3.0}
    rows = []
    for p, prec in PRECISIONS + ((4, ntc.EvalPrecision.BF16_MFMA), (5, ntc.EvalPrecision.FP8_MFMA)):
        ev = ctx.EvalEx(prec)
        rows.append((p, prec.name, ev.psnr, bench.group_psnr(ev.channel_sse, count), ev.channel_sse[:O].copy()))
    print(f"\nFP64 oracle: psnr {ref_all:.4f} dB " + " ".join(f"{g} {v:.4f}" for g, v in ref.items()))
    for p, name, psnr, grp, _ in rows:
        print(f"{name}: psnr {psnr:.4f} dB ({psnr - ref_all:+.4f}) " +
              " ".join(f"{g} {grp[g]:.4f} ({grp[g] - ref[g]:+.4f})" for g in grp))
    for p, name, psnr, grp, cs in rows:
        if p == 0:
            assert np.allclose(cs, sse["f64"], rtol=1e-4, atol=0)
        for g in grp:
            assert abs(grp[g] - ref[g]) < tol[p], (name, g, grp[g], ref[g])
        assert abs(psnr - ref_all) < tol[p]


@pytest.mark.parametrize("kind", ["c2_bilinear", "c3_h128", "ref_corners_hardgelu"])
@pytest.mark.parametrize("precision", [0, 1, 2, 3, 4, 5])
def test_decode_tiles_matches_gather_then_decode(oracle, kind, precision):
    """The decode kernel that forms each batch's features and targets itself
    (ntc_decode_tiles_sse_f32, as Context's evaluation walk runs it) against
    the separate gather, targets and decode kernels: per-channel squared
    errors equal bit for bit, for three tiles of 256^2 on a 512^2 mip (one
    wrapping both edges) and one 8^2 tile of an 8^2 mip."""
    O = 9
    args = {"c2_bilinear": (4, 2, 12, 16, 2, 64, 1, 1), "c3_h128": (4, 4, 16, 16, 2, 128, 1, 1),
            "ref_corners_hardgelu": (4, 4, 16, 12, 2, 64, 0, 0)}[kind]
    Lay = oracle.layout_custom(*args, 512)
    F, Fp, H, act = Lay.feature_dim, Lay.feature_dim_padded, Lay.hidden_dim, args[7]
    lib = G.lib()
    rng = np.random.default_rng(7 + precision)
    w0, w1, wo, dec = _decoder(rng, F, H, O)
    table = np.zeros(lib.ntc_decode_table_bytes(precision, F, Fp, H), np.uint8)
    if precision == 3:
        s_a1, s_a2 = np.float32(1.1 * 1.2 / 127), np.float32(1.1 * 0.8 / 127)
        blob = _int8_blob(oracle, w0, w1, wo[:O], Fp, s_a1, s_a2)
        G.ok(lib.ntc_build_decode_table_int8(F, Fp, H, O, blob.ctypes.data, float(np.float32(1) / s_a1),
                                             float(np.float32(1) / s_a2), table.ctypes.data))
    else:
        G.ok(lib.ntc_build_decode_table_f(precision, F, H, O, dec.ctypes.data, 1.7, 2.3, table.ctypes.data))
    d_table = G.dev(table)
    for M, T, origins in ((512, 256, [512 - 100, 512 - 37, 0, 0, 200, 311]), (8, 8, [3, 5])):
        lvl = oracle.level_for_lod(Lay, {512: 0, 8: 6}[M])
        fd = oracle.feature_desc(Lay, lvl)
        fdc = G.feature_desc(fd)
        g0 = G.dev(rng.uniform(-1, 1, fd["g0_dim"] ** 2 * fd["c0"]).astype(np.float32))
        g1 = G.dev(rng.uniform(-1, 1, fd["g1_dim"] ** 2 * fd["c1"]).astype(np.float32))
        pkg = G.dev(rng.integers(0, 256, M * M * O).astype(np.uint8))
        og = G.dev(np.array(origins, np.int32))
        n = len(origins) // 2
        tl = G.tiles(og, n, T, M, 0.25)
        B = n * T * T
        x = G.empty((F, B), np.float32)
        yt = G.empty((O, B), np.float32)
        nz = L.Noise(0, 0, 0, 0, 0.0, 0.0)
        G.ok(lib.ntc_draw_features_f32(G.byref(fdc), G.byref(tl), G.byref(nz), G.p(g0), G.p(g1), G.p(x), None))
        G.ok(lib.ntc_draw_targets_f32(G.byref(tl), O, G.p(pkg), G.p(yt), None))
        ws = G.empty((lib.ntc_decode_workspace_bytes(B),), np.uint8)
        sse_a = G.dev(np.zeros(16, np.float64))
        sse_b = G.dev(np.zeros(16, np.float64))
        G.ok(lib.ntc_decode_sse_f32(precision, act, F, Fp, H, O, B, G.p(d_table), G.p(x), G.p(yt), G.p(sse_a),
                                    None, G.p(ws), None), "decode")
        G.ok(lib.ntc_decode_tiles_sse_f32(precision, act, G.byref(fdc), G.byref(tl), G.p(g0), G.p(g1), G.p(pkg),
                                          Fp, H, O, G.p(d_table), G.p(sse_b), G.p(ws), None), "decode tiles")
        a, b = G.host(sse_a), G.host(sse_b)
        assert a[:O].min() > 0 and np.array_equal(a, b), (kind, precision, M, a[:O], b[:O])
