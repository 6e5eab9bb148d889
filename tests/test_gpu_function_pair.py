"""The fused step as a forward / backward pair (ntc_step_forward /
ntc_step_backward), the reference's step split at the loss
(libopenntc.cpp:621-655 forward, :659-668 loss, :672-709 backward):

* the forward's outputs against the same forward with the kernel's BF16
  rounding points (2e-3) and the FP64 oracle's forward (2e-2);
* forward, then dL/dYhat formed in FP32 from its outputs exactly as the
  training step forms it, then backward: the training step's decoder-gradient
  partials bit for bit and its latent gradients to atomic-order rounding;
* the backward of a random upstream gradient against the FP64 chain rule
  (5e-2 decoder, 1e-1 latent; 0.25 below 1024 texels, as for the step);
* at the 4096^2 widths (40 -> 128 -> 128 -> 9) on a 64-texel batch, against
  central finite differences of <dYhat, Yhat(theta)> in FP64 through the
  oracle's gather (selected weights and latent values)."""
import ctypes as C

import numpy as np
import pytest

from openntc_amd import _lib as L

pytestmark = pytest.mark.gpu

from tests import gpu_util as G  # noqa: E402
from tests.test_gpu_fused import _outputs, _step_case, act_b, act_f, bf16, rel  # noqa: E402


def _blocks(cs):
    lib = G.lib()
    maxb = C.c_int32()
    G.ok(lib.ntc_fused_max_blocks(C.byref(maxb)))
    return lib.ntc_fused_blocks_for(G.byref(cs["fdc"]), cs["H"], cs["act"], cs["B"], maxb.value)


def _reduce(cs, out, blocks):
    lib = G.lib()
    grad = G.empty((cs["dec"].size,), np.float32)
    sse = G.empty((1,), np.float64)
    G.ok(lib.ntc_reduce_decoder_partials(cs["H"], cs["F"], blocks, G.p(out["part"]), G.p(grad), G.p(out["ssep"]),
                                         G.p(sse), None))
    return G.host(grad), float(G.host(sse)[0])


def _split(cs, g):
    H, F = cs["H"], cs["F"]
    return g[: H * F].reshape(H, F), g[H * F: H * F + H * H].reshape(H, H), g[H * F + H * H:].reshape(16, H)


def _fp64(oracle, cs, lod_args):
    """FP64 gather (with the step's noise) and forward activations."""
    fd, origins, T, M = cs["fd"], cs["origins"], cs["T"], cs["M"]
    a = cs["a"]
    x = oracle.draw_features(fd, origins, T, M, a.tiles.norm_lod, cs["g0"].astype(np.float64),
                             cs["g1"].astype(np.float64), seed=a.noise.seed, step=a.noise.step,
                             level=a.noise.level, noise_on=True, d0=a.noise.delta0, d1=a.noise.delta1,
                             dtype=np.float64)
    H, F, Od = cs["H"], cs["F"], cs["Od"]
    dec = cs["dec"].astype(np.float64)
    w0, w1, wo = dec[: H * F].reshape(H, F), dec[H * F: H * F + H * H].reshape(H, H), dec[H * F + H * H:].reshape(16, H)
    z0 = w0 @ x
    a0 = act_f(cs["act"], z0)
    z1 = w1 @ a0
    a1 = act_f(cs["act"], z1)
    return x, (w0, w1, wo[:Od]), (z0, a0, z1, a1), wo[:Od] @ a1


def _forward(cs, blocks, with_targets=True):
    lib = G.lib()
    yhat = G.empty((cs["Od"], cs["B"]), np.float32)
    ssep = G.empty((blocks,), np.float64)
    a = cs["a"]
    a.sse_partials, a.blocks = G.p(ssep), blocks
    if not with_targets:
        saved, a.targets = a.targets, None
    G.ok(lib.ntc_step_forward(G.byref(a), G.p(yhat), None), "ntc_step_forward")
    if not with_targets:
        a.targets = saved
    return G.host(yhat), ssep


CASES = [("c3_bilinear_h128", 0, 2), ("c3_bilinear_h128", 3, 4), ("c1_bilinear_relu", 1, 1),
         ("ref_bpp_2_25", 2, 3), ("ref_bpp_0_5", 4, 5), ("c2_bilinear", 6, 7)]


@pytest.mark.parametrize("kind,lod,n_tiles", CASES)
def test_forward_outputs(oracle, kind, lod, n_tiles):
    cs = _step_case(oracle, kind, lod, n_tiles)
    blocks = _blocks(cs)
    yhat, ssep = _forward(cs, blocks)
    # the same forward with the kernel's BF16 rounding points
    H, F, Od, act = cs["H"], cs["F"], cs["Od"], cs["act"]
    x, (w0, w1, wo), _, y64 = _fp64(oracle, cs, None)
    fd, a = cs["fd"], cs["a"]
    x32 = oracle.draw_features(fd, cs["origins"], cs["T"], cs["M"], a.tiles.norm_lod, cs["g0"], cs["g1"],
                               seed=a.noise.seed, step=a.noise.step, level=a.noise.level, noise_on=True,
                               d0=a.noise.delta0, d1=a.noise.delta1)
    a0 = bf16(act_f(act, (bf16(w0.astype(np.float32)) @ bf16(x32)).astype(np.float32)))
    a1 = bf16(act_f(act, (bf16(w1.astype(np.float32)) @ a0).astype(np.float32)))
    yb = bf16(wo.astype(np.float32)) @ a1
    assert rel(yhat, yb) < 2e-3, rel(yhat, yb)
    assert rel(yhat, y64) < 2e-2, rel(yhat, y64)
    # SSE partials: the SSE against the targets, as the training step's
    yt = oracle.draw_targets(cs["origins"], cs["T"], cs["M"], Od, cs["pkg"], np.float32)
    # residual in FP32, squares summed in FP64, as the kernel forms them
    want = float(np.sum((yt - yhat).astype(np.float64) ** 2))
    got = float(np.sum(G.host(ssep)))
    assert abs(got - want) <= 1e-11 * want
    # without targets the outputs are the same
    y2, _ = _forward(cs, blocks, with_targets=False)
    assert np.array_equal(y2, yhat)


@pytest.mark.parametrize("kind,lod,n_tiles", CASES)
def test_pair_equals_training_step(oracle, kind, lod, n_tiles):
    lib = G.lib()
    cs = _step_case(oracle, kind, lod, n_tiles)
    blocks = _blocks(cs)
    a = cs["a"]
    # the training step
    ref = _outputs(cs, blocks)
    a.dg0, a.dg1, a.dw_partials, a.sse_partials, a.blocks = (G.p(ref["dg0"]), G.p(ref["dg1"]), G.p(ref["part"]),
                                                             G.p(ref["ssep"]), blocks)
    G.ok(lib.ntc_fused_train_step(G.byref(a), None), "ntc_fused_train_step")
    # the pair, with dL/dYhat formed in FP32 from the forward's outputs
    yhat, _ = _forward(cs, blocks)
    yt = oracle.draw_targets(cs["origins"], cs["T"], cs["M"], cs["Od"], cs["pkg"], np.float32)
    dy = (np.float32(a.loss_scale) * (yt - yhat)).astype(np.float32)
    out = _outputs(cs, blocks)
    a.dg0, a.dg1, a.dw_partials, a.sse_partials = G.p(out["dg0"]), G.p(out["dg1"]), G.p(out["part"]), G.p(out["ssep"])
    G.ok(lib.ntc_step_backward(G.byref(a), G.p(G.dev(dy)), None), "ntc_step_backward")
    assert np.array_equal(G.host(out["part"]), G.host(ref["part"]))
    for k in ("dg0", "dg1"):
        got, want = G.host(out[k]), G.host(ref[k])
        assert rel(got, want) < 1e-6, (k, rel(got, want))
    # the backward writes zero SSE partials (it reads no targets)
    assert not np.any(G.host(out["ssep"]))


@pytest.mark.parametrize("kind,lod,n_tiles", CASES)
def test_backward_of_random_upstream_gradient(oracle, kind, lod, n_tiles):
    lib = G.lib()
    cs = _step_case(oracle, kind, lod, n_tiles)
    blocks = _blocks(cs)
    rng = np.random.default_rng(lod * 31 + n_tiles)
    Od, B, act = cs["Od"], cs["B"], cs["act"]
    dy = (rng.standard_normal((Od, B)) / B).astype(np.float32)
    out = _outputs(cs, blocks)
    a = cs["a"]
    a.dg0, a.dg1, a.dw_partials, a.sse_partials, a.blocks = (G.p(out["dg0"]), G.p(out["dg1"]), G.p(out["part"]),
                                                             G.p(out["ssep"]), blocks)
    G.ok(lib.ntc_step_backward(G.byref(a), G.p(G.dev(dy)), None), "ntc_step_backward")
    g, _ = _reduce(cs, out, blocks)
    gw0, gw1, gwo = _split(cs, g)
    x, (w0, w1, wo), (z0, a0, z1, a1), _ = _fp64(oracle, cs, None)
    d = dy.astype(np.float64)
    dz1 = act_b(act, z1, wo.T @ d)
    dz0 = act_b(act, z0, w1.T @ dz1)
    dx = w0.T @ dz0
    rdg0, rdg1 = oracle.scatter(cs["fd"], cs["origins"], cs["T"], cs["M"], dx, cs["g0"].shape, cs["g1"].shape)
    # (1) the kernel's BF16 rounding points (features, weights, activations,
    # dYhat and the hidden-unit gradients rounded where the kernel rounds
    # them), products summed in FP64: FP32 accumulation order only
    fd, a = cs["fd"], cs["a"]
    x32 = oracle.draw_features(fd, cs["origins"], cs["T"], cs["M"], a.tiles.norm_lod, cs["g0"], cs["g1"],
                               seed=a.noise.seed, step=a.noise.step, level=a.noise.level, noise_on=True,
                               d0=a.noise.delta0, d1=a.noise.delta1)
    xb = bf16(x32)
    w0b, w1b, wob = bf16(w0.astype(np.float32)), bf16(w1.astype(np.float32)), bf16(wo.astype(np.float32))
    bz0 = w0b @ xb
    ba0 = bf16(act_f(act, bz0.astype(np.float32)))
    bz1 = w1b @ ba0
    ba1 = bf16(act_f(act, bz1.astype(np.float32)))
    db = bf16(dy)
    bdz1 = bf16(act_b(act, bz1, wob.T @ db).astype(np.float32))
    bdz0 = bf16(act_b(act, bz0, w1b.T @ bdz1).astype(np.float32))
    bdg0, bdg1 = oracle.scatter(fd, cs["origins"], cs["T"], cs["M"], np.ascontiguousarray(w0b.T @ bdz0),
                                cs["g0"].shape, cs["g1"].shape)
    for got, want, name in ((gw0, bdz0 @ xb.T, "dW0"), (gw1, bdz1 @ ba0.T, "dW1"), (gwo[:Od], db @ ba1.T, "dWout"),
                            (G.host(out["dg0"]), bdg0, "dG0"), (G.host(out["dg1"]), bdg1, "dG1")):
        assert rel(got, want) < 2e-3, (name, rel(got, want))
    assert not gwo[Od:].any()
    # (2) the FP64 chain rule. An upstream gradient of random signs makes
    # every sum over texels cancel (unlike the loss gradient, which follows
    # the residuals), so the BF16 rounding of the operands weighs about twice
    # as much as in the training step's check
    tol, tol_lat = (1e-1, 2e-1) if B >= 1024 else (0.25, 0.25)
    assert rel(gw0, dz0 @ x.T) < tol, rel(gw0, dz0 @ x.T)
    assert rel(gw1, dz1 @ a0.T) < tol, rel(gw1, dz1 @ a0.T)
    assert rel(gwo[:Od], d @ a1.T) < tol, rel(gwo[:Od], d @ a1.T)
    assert rel(G.host(out["dg0"]), rdg0) < tol_lat
    assert rel(G.host(out["dg1"]), rdg1) < tol_lat


def test_backward_vs_finite_differences_at_4096_widths(oracle):
    """64 texels (one 8x8 tile of the 8x8 mip) at the 40 -> 128 -> 128 -> 9
    widths: the BF16 backward of L = <dYhat, Yhat> against central
    differences of L in FP64 (h = 1e-6) for 48 weights and 32 latent values."""
    lib = G.lib()
    cs = _step_case(oracle, "c3_bilinear_h128", 5, 1)
    assert cs["B"] == 64 and cs["H"] == 128 and cs["F"] == 40
    blocks = _blocks(cs)
    rng = np.random.default_rng(64)
    Od = cs["Od"]
    dy = rng.standard_normal((Od, 64)).astype(np.float32)
    out = _outputs(cs, blocks)
    a = cs["a"]
    a.dg0, a.dg1, a.dw_partials, a.sse_partials, a.blocks = (G.p(out["dg0"]), G.p(out["dg1"]), G.p(out["part"]),
                                                             G.p(out["ssep"]), blocks)
    G.ok(lib.ntc_step_backward(G.byref(a), G.p(G.dev(dy)), None), "ntc_step_backward")
    g, _ = _reduce(cs, out, blocks)
    dg0, dg1 = G.host(out["dg0"]), G.host(out["dg1"])
    d64 = dy.astype(np.float64)
    base_dec, base_g0, base_g1 = cs["dec"].copy(), cs["g0"].copy(), cs["g1"].copy()

    def loss(dec=None, g0=None, g1=None):
        cs2 = dict(cs)
        cs2["dec"] = base_dec if dec is None else dec
        cs2["g0"] = base_g0 if g0 is None else g0
        cs2["g1"] = base_g1 if g1 is None else g1
        return float(np.sum(d64 * _fp64(oracle, cs2, None)[3]))

    h = 1e-6
    H, F = cs["H"], cs["F"]
    got, want = [], []
    n_dec = H * F + H * H + Od * H
    for e in rng.choice(n_dec, 48, replace=False):
        dp, dm = base_dec.astype(np.float64), base_dec.astype(np.float64)
        dp[e] += h
        dm[e] -= h
        want.append((loss(dec=dp) - loss(dec=dm)) / (2 * h))
        got.append(g[e])
    # latent values inside the batch's footprint
    for grid, gd, base in ((0, dg0, base_g0), (1, dg1, base_g1)):
        idx = np.flatnonzero(gd)
        for e in rng.choice(idx, 16, replace=False):
            gp, gm = base.astype(np.float64).ravel(), base.astype(np.float64).ravel()
            gp[e] += h
            gm[e] -= h
            kw = "g0" if grid == 0 else "g1"
            want.append((loss(**{kw: gp.reshape(base.shape)}) - loss(**{kw: gm.reshape(base.shape)})) / (2 * h))
            got.append(gd.ravel()[e])
    got, want = np.array(got), np.array(want)
    assert rel(got, want) < 0.25, rel(got, want)
    assert np.corrcoef(got, want)[0, 1] > 0.97
