"""The hand-written backward run on the GPU in FP64 (the plain path's
kernels: gather, GEMMs, activations, loss, latent scatter) against central
finite differences of the FP64 loss and against the CPU oracle's FP64
gradients, for 64 texels at the widths of the 4096^2 configuration
(F = 40 features from 16 + 16 latent channels and 8 triangle waves,
H = 128, O = 9): every checked entry of dW0, dW1, dWout, dG0 and dG1 within
1e-3 relative of its finite difference, and the whole gradient within 1e-10
relative of the oracle's. The tile wraps the mip edge."""
import numpy as np
import pytest

from openntc_amd import _lib as L

pytestmark = pytest.mark.gpu

from tests import gpu_util as G  # noqa: E402

H, O, T = 128, 9, 8


def act_f(act, z):
    if act == 1:
        return np.maximum(z, 0.0)
    return np.where(z < -1.5, 0.0, np.where(z < 1.5, z / 3 * (z + 1.5), z))


def gemm(lib, op, n, m, k, a, b):
    c = G.empty((n, m), np.float64)
    wsb = lib.ntc_gemm_workspace_bytes(op, n, m, k, 8)
    ws = G.empty((max(wsb, 8),), np.uint8)
    G.ok(lib.ntc_gemm_f64(op, n, m, k, G.p(a), G.p(b), G.p(c), G.p(ws), wsb, None), "ntc_gemm_f64")
    return c


@pytest.mark.parametrize("act", [1, 0])
def test_fp64_backward_vs_finite_differences(oracle, act):
    Ora = oracle
    Lay = Ora.layout_custom(4, 4, 16, 16, 2, H, 1, act, 64)
    fd = Ora.feature_desc(Lay, 0)
    F, M, B = fd["F"], Lay.mip_dim[0], T * T
    assert F == 40
    rng = np.random.default_rng(40 + act)
    origins = np.array([M - 3, 21], np.int32)
    g0 = rng.uniform(-1, 1, (fd["g0_dim"], fd["g0_dim"], fd["c0"]))
    g1 = rng.uniform(-1, 1, (fd["g1_dim"], fd["g1_dim"], fd["c1"]))
    w0 = rng.uniform(-1, 1, (H, F)) / np.sqrt(F)
    w1 = rng.uniform(-1, 1, (H, H)) / np.sqrt(H)
    wout = rng.uniform(-1, 1, (O, H)) / np.sqrt(H)
    pkg = rng.integers(0, 256, (M, M, O)).astype(np.uint8)
    scale = -2.0 / (O * B)
    lib = G.lib()

    # ---- GPU, FP64 ----
    fdc = G.feature_desc(fd)
    tiles = G.tiles(G.dev(origins), 1, T, M, 0.0)
    x = G.empty((F, B), np.float64)
    G.ok(lib.ntc_draw_features_f64(G.byref(fdc), G.byref(tiles), G.byref(L.Noise(123, 0, 0, 0, 0.0, 0.0)),
                                   G.p(G.dev(g0)), G.p(G.dev(g1)), G.p(x), None))
    yt = G.empty((O, B), np.float64)
    G.ok(lib.ntc_draw_targets_f64(G.byref(tiles), O, G.p(G.dev(pkg)), G.p(yt), None))
    d_w0, d_w1, d_wo = G.dev(w0), G.dev(w1), G.dev(wout)
    z0 = gemm(lib, 0, H, B, F, d_w0, x)
    a0 = G.empty((H, B), np.float64)
    G.ok(lib.ntc_act_forward_f64(act, H * B, G.p(z0), G.p(a0), None))
    z1 = gemm(lib, 0, H, B, H, d_w1, a0)
    a1 = G.empty((H, B), np.float64)
    G.ok(lib.ntc_act_forward_f64(act, H * B, G.p(z1), G.p(a1), None))
    yh = gemm(lib, 0, O, B, H, d_wo, a1)
    dy = G.empty((O, B), np.float64)
    sse = G.empty((1,), np.float64)
    ws = G.empty((lib.ntc_loss_workspace_bytes(O * B),), np.uint8)
    G.ok(lib.ntc_loss_f64(O * B, scale, G.p(yt), G.p(yh), G.p(dy), G.p(sse), G.p(ws), None))
    dwo = gemm(lib, 2, O, H, B, dy, a1)
    da1 = gemm(lib, 1, H, B, O, d_wo, dy)
    dz1 = G.empty((H, B), np.float64)
    G.ok(lib.ntc_act_backward_f64(act, H * B, G.p(z1), G.p(da1), G.p(dz1), None))
    dw1 = gemm(lib, 2, H, H, B, dz1, a0)
    da0 = gemm(lib, 1, H, B, H, d_w1, dz1)
    dz0 = G.empty((H, B), np.float64)
    G.ok(lib.ntc_act_backward_f64(act, H * B, G.p(z0), G.p(da0), G.p(dz0), None))
    dw0 = gemm(lib, 2, H, F, B, dz0, x)
    dx = gemm(lib, 1, F, B, H, d_w0, dz0)
    dg0 = G.empty(g0.shape, np.float64)
    dg1 = G.empty(g1.shape, np.float64)
    G.ok(lib.ntc_accumulate_grid_gradients_f64(G.byref(fdc), G.byref(tiles), G.p(dx), G.p(dg0), G.p(dg1), None))
    got = {"W0": G.host(dw0), "W1": G.host(dw1), "Wout": G.host(dwo), "G0": G.host(dg0), "G1": G.host(dg1)}

    # ---- the CPU oracle's FP64 gradients ----
    xr = Ora.draw_features(fd, origins, T, M, 0.0, g0, g1, dtype=np.float64)
    ytr = Ora.draw_targets(origins, T, M, O, pkg, np.float64)
    wout16 = np.zeros((16, H))
    wout16[:O] = wout
    _, rw0, rw1, rwo, rdx = Ora.mlp_fwd_bwd(F, H, O, act, xr, ytr, w0, w1, wout16)
    rg0, rg1 = Ora.scatter(fd, origins, T, M, rdx, g0.shape, g1.shape)
    ref = {"W0": rw0, "W1": rw1, "Wout": rwo[:O], "G0": rg0, "G1": rg1}
    for k in got:
        err = np.linalg.norm(got[k] - ref[k]) / np.linalg.norm(ref[k])
        assert err < 1e-10, (k, err)

    # ---- central finite differences of the FP64 loss ----
    def loss(w0_, w1_, wo_, g0_, g1_):
        xx = Ora.draw_features(fd, origins, T, M, 0.0, g0_, g1_, dtype=np.float64)
        y = wo_ @ act_f(act, w1_ @ act_f(act, w0_ @ xx))
        return float(np.sum((ytr - y) ** 2)) / (O * B)

    params = {"W0": w0, "W1": w1, "Wout": wout, "G0": g0, "G1": g1}
    h = 1e-6
    for name, arr in params.items():
        g = got[name]
        flat = np.abs(g).ravel()
        # the largest entries plus random non-zero ones
        nz = np.flatnonzero(flat)
        picks = list(np.argsort(flat)[-8:]) + list(rng.choice(nz, 8, replace=False))
        for idx in picks:
            pi = np.unravel_index(idx, arr.shape)
            vals = []
            for sgn in (1, -1):
                p = {k: v.copy() for k, v in params.items()}
                p[name][pi] += sgn * h
                vals.append(loss(p["W0"], p["W1"], p["Wout"], p["G0"], p["G1"]))
            fdv = (vals[0] - vals[1]) / (2 * h)
            rel = abs(g[pi] - fdv) / max(abs(fdv), 1e-3 * np.abs(g).max())
            assert rel < 1e-3, (name, pi, g[pi], fdv, rel)
