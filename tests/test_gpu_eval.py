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


def test_context_eval_precisions_match_oracle(oracle):
    """A trained 256^2 context evaluated at every precision over all mips
    against the oracle's decode of the context's own trained state."""
    from tests.test_gpu_context import _make
    steps = 60
    ctx, Lay, P, sched, pk = _make(oracle, steps, ntc.TrainPath.Plain, grids=1)
    assert ctx.TrainUntilComplete().phase == ntc.TrainPhase.TrainComplete
    g0b, g1b, blob, cal = ctx.GetCompressedData()
    dec = ctx.read_decoder()
    F, Fp, H, O = Lay.feature_dim, Lay.feature_dim_padded, Lay.hidden_dim, Lay.out_dim
    grids = [(ctx.read_grid(0, l), ctx.read_grid(1, l)) for l in range(Lay.level_count)]
    # the oracle's activation ranges of the same state (FP8 scales)
    P2 = oracle.Params(Lay)
    for l in range(Lay.level_count):
        P2.g0[l][...], P2.g1[l][...] = grids[l]
    P2.dec[...] = dec
    mx = oracle.Trainer(Lay, P2, pk, max_batch=65536).calibrate()
    s8 = oracle.fp8_scales(F, H, O, dec, mx[0], mx[1])
    rcp = np.array([np.float32(1) / np.float32(cal[2]), np.float32(1) / np.float32(cal[3])], np.float32)
    sse = {p: np.zeros(O) for p in (0, 1, 2, 3, "f64")}
    count = 0
    for mip in range(Lay.mip_count):
        M = Lay.mip_dim[mip]
        T = min(M, 256)
        per = (M + T - 1) // T
        origins = np.array([v for g in range(per * per) for v in ((g % per) * T, (g // per) * T)], np.int32)
        lvl = oracle.level_for_lod(Lay, mip)
        fd = oracle.feature_desc(Lay, lvl)
        x = oracle.draw_features(fd, origins, T, M, np.float32(mip / (Lay.mip_count - 1)), *grids[lvl])
        yt = oracle.draw_targets(origins, T, M, O, pk[mip])
        for p in (0, 1, 2, 3):
            y = oracle.decode(p, Lay.activation, F, Fp, H, O, x, dec=dec, s8=s8, blob=blob, rcp=rcp)
            sse[p] += np.sum(((yt - y).astype(np.float64)) ** 2, axis=1)
        y64 = oracle.decode_f64(Lay.activation, F, H, O, x.astype(np.float64), dec.astype(np.float64))
        sse["f64"] += np.sum((yt.astype(np.float64) - y64) ** 2, axis=1)
        count += x.shape[1]
    psnr64 = -10 * np.log10(sse["f64"].sum() / (count * O))
    # the oracle's own FP32 eval walk agrees with the per-mip restatement
    _, ocs = oracle.Trainer(Lay, P2, pk, max_batch=65536).eval(channel_sse=True)
    assert np.allclose(ocs, sse[0], rtol=1e-12, atol=0)
    for p, prec in ((0, ntc.EvalPrecision.FP32), (1, ntc.EvalPrecision.BF16), (2, ntc.EvalPrecision.FP8),
                    (3, ntc.EvalPrecision.INT8)):
        ev = ctx.EvalEx(prec)
        assert np.allclose(ev.channel_sse[:O], sse[p], rtol=1e-10, atol=0), (prec, ev.channel_sse[:O], sse[p])
        want = -10 * np.log10(sse[p].sum() / (count * O))
        assert abs(ev.psnr - want) < 1e-8, (prec, ev.psnr, want)
        print(f"{prec.name}: psnr {ev.psnr:.4f} dB, vs FP64 decode {ev.psnr - psnr64:+.4f} dB")
        assert abs(ev.psnr - psnr64) < {0: 1e-3, 1: 0.2, 2: 1.5, 3: 1.5}[p]
