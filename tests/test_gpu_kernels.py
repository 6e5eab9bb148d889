"""Per-stage parity of the gfx950 kernels against the CPU oracle, through the
C ABI. Bit-exact where the operation order is fixed (gather, targets, GEMMs,
activations, loss gradient, Adam, quantise, pack, Lanczos); a stated
tolerance where it is not (FP64 sums, atomic scatter order)."""
import numpy as np
import pytest

from openntc_amd import _lib as L

pytestmark = pytest.mark.gpu

from tests import gpu_util as G  # noqa: E402


def _layouts(O):
    return {
        "ref_bpp_2_25": O.layout_profile(3, 256, 9),
        "ref_bpp_1_0_c10": O.layout_profile(2, 256, 9),
        "bilinear_c1": O.layout_custom(4, 4, 8, 16, 2, 64, 1, 1, 512, 9),
    }


@pytest.mark.parametrize("name", ["ref_bpp_2_25", "ref_bpp_1_0_c10", "bilinear_c1"])
@pytest.mark.parametrize("lod", [0, 1, 3])
@pytest.mark.parametrize("noise", [False, True])
def test_gather_bitexact(oracle, name, lod, noise):
    O = oracle
    Lay = _layouts(O)[name]
    rng = np.random.default_rng(lod * 7 + noise)
    level = O.level_for_lod(Lay, lod)
    fd = O.feature_desc(Lay, level)
    M = Lay.mip_dim[lod]
    T = min(M, 256)
    n_tiles = 2
    origins = rng.integers(0, M, 2 * n_tiles).astype(np.int32)
    g0 = rng.uniform(-1, 1, (fd["g0_dim"], fd["g0_dim"], fd["c0"])).astype(np.float32)
    g1 = rng.uniform(-1, 1, (fd["g1_dim"], fd["g1_dim"], fd["c1"])).astype(np.float32)
    norm_lod = np.float32(lod / (Lay.mip_count - 1))
    args = dict(seed=123, step=77, level=level, noise_on=noise, d0=Lay.g0_delta, d1=Lay.g1_delta)
    ref = O.draw_features(fd, origins, T, M, norm_lod, g0, g1, **args)

    B = n_tiles * T * T
    x = G.empty((fd["F"], B), np.float32)
    og = G.dev(origins)
    nz = L.Noise(123, 77, level, int(noise), Lay.g0_delta, Lay.g1_delta)
    G.ok(G.lib().ntc_draw_features_f32(G.byref(G.feature_desc(fd)), G.byref(G.tiles(og, n_tiles, T, M, norm_lod)),
                                       G.byref(nz), G.p(G.dev(g0)), G.p(G.dev(g1)), G.p(x), None))
    got = G.host(x)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), \
        f"max diff {np.max(np.abs(got - ref))}"


def test_gather_f64_matches_oracle(oracle):
    O = oracle
    Lay = _layouts(O)["ref_bpp_2_25"]
    rng = np.random.default_rng(3)
    fd = O.feature_desc(Lay, 0)
    M, T = 256, 256
    origins = np.array([17, 200], np.int32)
    g0 = rng.uniform(-1, 1, (fd["g0_dim"], fd["g0_dim"], fd["c0"]))
    g1 = rng.uniform(-1, 1, (fd["g1_dim"], fd["g1_dim"], fd["c1"]))
    ref = O.draw_features(fd, origins, T, M, 0.0, g0, g1, noise_on=True, d0=0.125, d1=0.125,
                          dtype=np.float64)
    x = G.empty((fd["F"], T * T), np.float64)
    nz = L.Noise(123, 0, 0, 1, 0.125, 0.125)
    G.ok(G.lib().ntc_draw_features_f64(G.byref(G.feature_desc(fd)), G.byref(G.tiles(G.dev(origins), 1, T, M, 0.0)),
                                       G.byref(nz), G.p(G.dev(g0)), G.p(G.dev(g1)), G.p(x), None))
    assert np.array_equal(G.host(x), ref)


@pytest.mark.parametrize("M,T,n_tiles", [(256, 256, 3), (64, 64, 1), (512, 256, 8)])
def test_targets_bitexact(oracle, M, T, n_tiles):
    rng = np.random.default_rng(M)
    O_ = 9
    pkg = rng.integers(0, 256, (M, M, O_)).astype(np.uint8)
    origins = rng.integers(0, M, 2 * n_tiles).astype(np.int32)
    ref = oracle.draw_targets(origins, T, M, O_, pkg)
    y = G.empty((O_, n_tiles * T * T), np.float32)
    G.ok(G.lib().ntc_draw_targets_f32(G.byref(G.tiles(G.dev(origins), n_tiles, T, M, 0.0)), O_, G.p(G.dev(pkg)),
                                      G.p(y), None))
    assert np.array_equal(G.host(y), ref)


@pytest.mark.parametrize("op,n,m,k", [
    (0, 64, 5000, 40), (0, 128, 3001, 128), (0, 9, 4096, 64),
    (1, 64, 5000, 9), (1, 92, 2500, 64),
    (2, 9, 64, 5000), (2, 128, 128, 4097), (2, 64, 92, 1024),
])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_gemm_bitexact(oracle, op, n, m, k, dtype):
    rng = np.random.default_rng(n * m + k)
    if op == 0:
        a, b = rng.standard_normal((n, k)), rng.standard_normal((k, m))
    elif op == 1:
        a, b = rng.standard_normal((k, n)), rng.standard_normal((k, m))
    else:
        a, b = rng.standard_normal((n, k)), rng.standard_normal((m, k))
    a, b = a.astype(dtype), b.astype(dtype)
    ref = oracle.gemm(op, n, m, k, a, b, dtype)
    c = G.empty((n, m), dtype)
    es = np.dtype(dtype).itemsize
    wsb = G.lib().ntc_gemm_workspace_bytes(op, n, m, k, es)
    ws = G.empty((max(wsb, 8),), np.uint8)
    fn = G.lib().ntc_gemm_f32 if dtype == np.float32 else G.lib().ntc_gemm_f64
    G.ok(fn(op, n, m, k, G.p(G.dev(a)), G.p(G.dev(b)), G.p(c), G.p(ws), wsb, None))
    got = G.host(c)
    assert np.array_equal(got, ref), f"max diff {np.max(np.abs(got - ref))}"
    # and against numpy in double (the oracle's own pin)
    full = (a.astype(np.float64) @ b.astype(np.float64)) if op == 0 else (
        a.T.astype(np.float64) @ b.astype(np.float64) if op == 1 else a.astype(np.float64) @ b.T.astype(np.float64))
    tol = 1e-4 if dtype == np.float32 else 1e-11
    assert np.allclose(got, full, rtol=tol, atol=tol * np.sqrt(k))


@pytest.mark.parametrize("act", [0, 1])
def test_activations_bitexact(oracle, act):
    rng = np.random.default_rng(act)
    z = np.concatenate([rng.uniform(-3, 3, 100000), [-1.5, 1.5, 0.0, -0.0, 1e-30, -1e-30]]).astype(np.float32)
    g = rng.standard_normal(z.size).astype(np.float32)
    a = G.empty(z.shape, np.float32)
    d = G.empty(z.shape, np.float32)
    G.ok(G.lib().ntc_act_forward_f32(act, z.size, G.p(G.dev(z)), G.p(a), None))
    G.ok(G.lib().ntc_act_backward_f32(act, z.size, G.p(G.dev(z)), G.p(G.dev(g)), G.p(d), None))
    assert np.array_equal(G.host(a).view(np.uint32), oracle.act_forward(act, z).view(np.uint32))
    assert np.array_equal(G.host(d).view(np.uint32), oracle.act_backward(act, z, g).view(np.uint32))


def test_loss(oracle):
    rng = np.random.default_rng(5)
    n = 9 * 70000
    yt = rng.uniform(0, 1, n).astype(np.float32)
    yp = rng.uniform(-0.2, 1.2, n).astype(np.float32)
    scale = np.float32(-2.0) / (np.float32(9) * np.float32(70000))
    sse_ref, dy_ref = oracle.loss(scale, yt, yp)
    dy = G.empty((n,), np.float32)
    sse = G.empty((1,), np.float64)
    ws = G.empty((G.lib().ntc_loss_workspace_bytes(n),), np.uint8)
    G.ok(G.lib().ntc_loss_f32(n, float(scale), G.p(G.dev(yt)), G.p(G.dev(yp)), G.p(dy), G.p(sse), G.p(ws), None))
    assert np.array_equal(G.host(dy), dy_ref)
    assert abs(G.host(sse)[0] - sse_ref) <= 1e-12 * sse_ref


@pytest.mark.parametrize("name", ["ref_bpp_2_25", "bilinear_c1"])
@pytest.mark.parametrize("lod", [0, 2])
def test_scatter_matches_oracle(oracle, name, lod):
    O = oracle
    Lay = _layouts(O)[name]
    level = O.level_for_lod(Lay, lod)
    fd = O.feature_desc(Lay, level)
    M = Lay.mip_dim[lod]
    T = min(M, 256)
    rng = np.random.default_rng(11 + lod)
    n_tiles = 3
    origins = rng.integers(0, M, 2 * n_tiles).astype(np.int32)
    dx = rng.standard_normal((fd["F"], n_tiles * T * T)).astype(np.float32)
    s0 = (fd["g0_dim"], fd["g0_dim"], fd["c0"])
    s1 = (fd["g1_dim"], fd["g1_dim"], fd["c1"])
    r0, r1 = O.scatter(fd, origins, T, M, dx, s0, s1)
    dg0 = G.empty(s0, np.float32)
    dg1 = G.empty(s1, np.float32)
    G.ok(G.lib().ntc_accumulate_grid_gradients_f32(
        G.byref(G.feature_desc(fd)), G.byref(G.tiles(G.dev(origins), n_tiles, T, M, 0.0)), G.p(G.dev(dx)),
        G.p(dg0), G.p(dg1), None))
    # atomics add in a different order than the oracle's texel order
    for got, ref in ((G.host(dg0), r0), (G.host(dg1), r1)):
        scale = np.abs(ref).max()
        assert np.max(np.abs(got - ref)) <= 1e-5 * scale


def test_adam_and_grid_update_bitexact(oracle):
    rng = np.random.default_rng(9)
    n = 1_000_003
    g = (rng.standard_normal(n) * 1e-3).astype(np.float32)
    m = (rng.standard_normal(n) * 1e-3).astype(np.float32)
    v = (rng.uniform(0, 1e-6, n)).astype(np.float32)
    p = rng.uniform(-1.2, 1.2, n).astype(np.float32)
    lr, b1, b2, c1, c2 = 0.01, 0.9, 0.999, 1.7, 3.1
    mr, vr, pr = m.copy(), v.copy(), p.copy()
    oracle.adam(lr, b1, b2, c1, c2, g, mr, vr, pr)
    dm, dv, dp, dg = G.dev(m), G.dev(v), G.dev(p), G.dev(g)
    G.ok(G.lib().ntc_adam_f32(n, lr, b1, b2, c1, c2, G.p(dg), G.p(dm), G.p(dv), G.p(dp), None))
    assert np.array_equal(G.host(dm), mr) and np.array_equal(G.host(dv), vr) and np.array_equal(G.host(dp), pr)

    delta = 0.125
    oracle.clamp_grid(delta, pr)
    dm, dv, dp = G.dev(m), G.dev(v), G.dev(p)
    G.ok(G.lib().ntc_grid_update_f32(n, lr, b1, b2, c1, c2, -1.0, 1.0 - delta, G.p(dg), G.p(dm), G.p(dv), G.p(dp),
                                     None))
    assert np.array_equal(G.host(dp), pr) and np.array_equal(G.host(dm), mr)
    assert not G.host(dg).any(), "grid update must leave the gradient zeroed"


@pytest.mark.parametrize("bits", [1, 2, 4, 8])
def test_quantize_and_pack_bitexact(oracle, bits):
    rng = np.random.default_rng(bits)
    delta = np.float32(2.0 / 2 ** bits)
    half = 2 ** (bits - 1)
    # every code, the half-way points between codes, and out-of-range values
    ks = np.arange(-half - 2, half + 2)
    x = np.concatenate([ks * delta, (ks + 0.5) * delta, rng.uniform(-1.3, 1.3, 4096)]).astype(np.float32)
    x = np.concatenate([x, np.zeros((-x.size) % 32, np.float32)])
    q = G.dev(x)
    G.ok(G.lib().ntc_quantize_grid_f32(x.size, bits, float(delta), G.p(q), None))
    ref = x.copy()
    oracle.quantize_grid(bits, delta, ref)
    assert np.array_equal(G.host(q), ref)
    out = G.empty((x.size * bits // 32,), np.uint32)
    G.ok(G.lib().ntc_quantize_pack_f32(x.size, bits, G.p(G.dev(x)), G.p(out), None))
    assert np.array_equal(G.host(out, np.uint32), oracle.quantize_pack(bits, x))


@pytest.mark.parametrize("dim,c", [(64, 3), (32, 1), (16, 4), (8, 3)])
def test_lanczos_bitexact(oracle, dim, c):
    rng = np.random.default_rng(dim + c)
    src = rng.integers(0, 256, (dim, dim, c)).astype(np.uint8)
    ref = oracle.filter_lanczos(src)
    tmp = G.empty((dim * dim * c,), np.float32)
    dst = G.empty((dim // 2, dim // 2, c), np.uint8)
    G.ok(G.lib().ntc_filter_lanczos(dim, c, G.p(G.dev(src)), G.p(tmp), G.p(dst), None))
    assert np.array_equal(G.host(dst), ref)


def test_max_abs(oracle):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(300001).astype(np.float32)
    r = G.empty((1,), np.float32)
    G.ok(G.lib().ntc_max_abs_f32(x.size, G.p(G.dev(x)), G.p(r), None))
    assert G.host(r)[0] == np.abs(x).max()


def test_invalid_arguments_rejected(oracle):
    lib = G.lib()
    fd = L.FeatureDesc(8, 4, 40, 4, 0, 8, 3, 1, 92)  # 40 channels > 32
    t = L.Tiles(0, 1, 8, 8, 0.0)
    assert lib.ntc_draw_features_f32(G.byref(fd), G.byref(t), None, 0, 0, 0, None) == L.NTC_E_INVALID
    assert lib.ntc_quantize_pack_f32(33, 4, 1, 1, None) == L.NTC_E_INVALID
    assert lib.ntc_gemm_f32(2, 4, 4, 4096, 1, 1, 1, None, 0, None) == L.NTC_E_WORKSPACE
