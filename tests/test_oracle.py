"""Pins of the CPU oracle against known answers, closed forms, library
routines and the reference's second statement of the formats (the HLSL
decoder). The reference ships no tests or golden vectors; each pin here is
chosen so that a plausible restatement mistake (dropped term, wrong index or
sign, swapped operand) fails it."""
import math

import numpy as np
import pytest


# ---- RNG ------------------------------------------------------------------------
def test_mt19937_known_answer(oracle):
    # ISO C++ [rand.predef]: the 10000th output of a default-seeded
    # std::mt19937 (seed 5489) is 4123659995.
    m = oracle.MT(5489)
    for _ in range(9999):
        m.next()
    assert m.next() == 4123659995


@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox_known_answers(oracle, ctr, key, expect):
    # Random123 known-answer vectors for Philox4x32-10
    assert oracle.philox(ctr, key) == expect


def test_noise_range_and_moments(oracle):
    delta = 0.125
    vals = np.array([oracle.noise(123, 5, g, 1, cell, ch, delta)
                     for g in (0, 1) for cell in range(0, 4000, 7) for ch in range(12)], np.float64)
    assert vals.min() > -delta / 2 and vals.max() <= delta / 2
    assert abs(vals.mean()) < 0.01 * delta
    assert abs(vals.var() - delta ** 2 / 12) < 0.05 * delta ** 2 / 12
    # values are k * 2^-24 * delta - delta/2: on the 24-bit lattice
    k = (vals + delta / 2) / delta * 2 ** 24
    assert np.array_equal(k, np.rint(k))
    # consistent within a step, different across steps
    assert oracle.noise(1, 2, 0, 0, 10, 3, delta) == oracle.noise(1, 2, 0, 0, 10, 3, delta)
    assert oracle.noise(1, 2, 0, 0, 10, 3, delta) != oracle.noise(1, 3, 0, 0, 10, 3, delta)


def test_uniform_int_unbiased_and_in_range(oracle):
    m = oracle.MT(7)
    draws = np.array([m.uniform_int(0, 10) for _ in range(55000)])
    assert draws.min() == 0 and draws.max() == 10
    counts = np.bincount(draws, minlength=11)
    assert np.all(np.abs(counts - 5000) < 5 * math.sqrt(5000))


def test_schedule_distribution(oracle):
    L = oracle.layout_profile(3, 4096)
    m = oracle.MT(123)
    n = 40000
    steps = oracle.schedule(m, L, n, 2)
    lods = np.array([s.lod for s in steps])
    mips = L.mip_count
    for k in range(mips - 1):
        p = 0.95 * 0.75 * 4.0 ** -k + 0.05 / mips
        assert abs((lods == k).mean() - p) < 5 * math.sqrt(p * (1 - p) / n) + 1e-4, k
    # tile origins uniform over the sampled mip
    s0 = [s for s in steps if s.lod == 0]
    xs = np.array([s.origins[0] for s in s0])
    assert xs.min() >= 0 and xs.max() < 4096 and abs(xs.mean() - 2047.5) < 60
    # lock at floor(95 n / 100); grid updates stop there
    lock = (95 * n) // 100
    assert steps[lock].lock == 1 and sum(s.lock for s in steps) == 1
    assert all(s.grid_update == (i < lock) for i, s in enumerate(steps[lock - 2: lock + 2], lock - 2))
    assert all(s.level == oracle.level_for_lod(L, s.lod) for s in steps[:100])


def test_cosine_annealing_and_bias(oracle):
    L = oracle.layout_profile(1, 1024)
    steps = oracle.schedule(oracle.MT(1), L, 1000, 1, 0.01, 0.005)
    assert steps[0].lr_grid == np.float32(0.01) and steps[0].lr_decoder == np.float32(0.005)
    assert abs(steps[500].lr_grid - 0.005) < 1e-7
    assert steps[0].bias1 == np.float32(1 / np.float32(1 - np.float32(0.9)))
    assert abs(steps[9].bias2 - 1 / (1 - 0.999 ** 10)) < 1e-4 * 100.45  # float32 cancellation


# ---- layout ----------------------------------------------------------------------
@pytest.mark.parametrize("profile,size,F", [(0, 3924472, 60), (1, 9518920, 84), (2, 19024976, 72),
                                            (3, 42514080, 92)])
def test_readme_sizes(oracle, profile, size, F):
    # README.md:48 compressed sizes at 4096^2 (3.93 / 9.52 / 19.0 / 42.5 MB)
    L = oracle.layout_profile(profile, 4096)
    assert oracle.compressed_bytes(L) == size
    assert L.feature_dim == F and L.mip_count == 11 and L.level_count == 5
    # HLSL FEATURE_DIM_PADDED (ntc_decode.hlsli:93)
    assert L.feature_dim_padded == ((4 * L.g0_channels + L.g1_channels + 13 + 15) // 16) * 16


def test_uncompressed_set_size(oracle):
    # README.md:44: 9 channels at 4096^2 with mips down to 4x4 = 201 326 544 B
    L = oracle.layout_profile(3, 4096)
    assert sum(L.mip_dim[i] ** 2 * 9 for i in range(L.mip_count)) == 201326544


@pytest.mark.parametrize("profile,level0_mips", [(0, 4), (1, 4), (2, 3), (3, 3)])
def test_level_for_lod_matches_hlsl(oracle, profile, level0_mips):
    # ntc_decode.hlsli:127-132 (no clamp there; the reference host clamps)
    L = oracle.layout_profile(profile, 4096)
    for lod in range(L.mip_count):
        hl = 0 if lod < level0_mips else 1 + (lod - level0_mips) // 2
        assert oracle.level_for_lod(L, lod) == min(hl, L.level_count - 1)


# ---- gather ----------------------------------------------------------------------
def _fd(oracle, bilinear=False, dim=64):
    if bilinear:
        L = oracle.layout_custom(4, 4, 4, 4, 2, 16, 1, 1, dim)
    else:
        L = oracle.layout_custom(4, 4, 4, 4, 2, 16, 0, 0, dim)
    return L, oracle.feature_desc(L, 0)


def test_gather_closed_form_cases(oracle):
    L, fd = _fd(oracle)
    M, T = 64, 64
    D0, D1 = fd["g0_dim"], fd["g1_dim"]
    rng = np.random.default_rng(0)
    g0 = rng.uniform(-1, 1, (D0, D0, 4)).astype(np.float32)
    g1 = rng.uniform(-1, 1, (D1, D1, 4)).astype(np.float32)
    x = oracle.draw_features(fd, np.array([0, 0], np.int32), T, M, 0.25, g0, g1)
    # texel (0,0): g = 0.5/64*32 - 0.5 = -0.25 -> x0 = -1, both x corners clamp to 0
    t = 0
    for c in range(4):
        assert np.array_equal(x[c * 4: c * 4 + 4, t], g0[0, 0])
    # PE: P=8, phase 0 at coordinate 0 -> 1 - 4|0.0625 - 0.5| = -0.75
    base = 16 + 4
    assert x[base, t] == np.float32(-0.75)
    # LOD slot then zero padding
    assert x[base + 12, t] == np.float32(0.25)
    assert np.all(x[base + 13:, t] == 0)
    # interior texel (5, 9): g0 x = 5.5/2 - 0.5 = 2.25 -> corners 2,3; y = 9.5/2-.5 = 4.25 -> 4,5
    t = 9 * 64 + 5
    assert np.array_equal(x[0:4, t], g0[4, 2]) and np.array_equal(x[4:8, t], g0[4, 3])
    assert np.array_equal(x[8:12, t], g0[5, 2]) and np.array_equal(x[12:16, t], g0[5, 3])
    # g1 (D1 = 16): gx = 5.5/4 - 0.5 = 0.875, gy = 9.5/4 - 0.5 = 1.875 -> bilinear
    fx, fy = 0.875, 0.875
    want = ((1 - fx) * (1 - fy) * g1[1, 0] + fx * (1 - fy) * g1[1, 1] + (1 - fx) * fy * g1[2, 0]
            + fx * fy * g1[2, 1])
    assert np.allclose(x[16:20, t], want, atol=1e-6)


def test_gather_matches_hlsl_formulas(oracle):
    # Recompute features the way the HLSL decoder does (ntc_decode.hlsli:
    # 218-224 corners via floor(uv*dim - 0.5) with max/min clamps, 259-267
    # bilinear g1 via frac(), 315-341 PE via floor(uv * (dim >> lod))) on
    # float grids and compare with the oracle's training gather.
    L, fd = _fd(oracle, dim=128)
    lod, M = 1, 64
    D0, D1 = fd["g0_dim"], fd["g1_dim"]
    rng = np.random.default_rng(4)
    g0 = rng.uniform(-1, 1, (D0, D0, 4)).astype(np.float64)
    g1 = rng.uniform(-1, 1, (D1, D1, 4)).astype(np.float64)
    norm_lod = lod / (L.mip_count - 1)
    x = oracle.draw_features(fd, np.array([13, 40], np.int32), 64, M, norm_lod, g0, g1, dtype=np.float64)
    for t in (0, 63, 64 * 30 + 17, 64 * 63 + 63):
        px, py = (t % 64 + 13) % M, (t // 64 + 40) % M
        uv = np.array([(px + 0.5) / M, (py + 0.5) / M])
        feats = []
        xy = np.floor(uv * D0 - 0.5).astype(int)
        xs = [max(xy[0], 0), min(xy[0] + 1, D0 - 1)]
        ys = [max(xy[1], 0), min(xy[1] + 1, D0 - 1)]
        for i in range(2):
            for j in range(2):
                feats.extend(g0[ys[i], xs[j]])
        xy1 = np.floor(uv * D1 - 0.5).astype(int)
        xs1 = [max(xy1[0], 0), min(xy1[0] + 1, D1 - 1)]
        ys1 = [max(xy1[1], 0), min(xy1[1] + 1, D1 - 1)]
        fr = uv * D1 - 0.5 - np.floor(uv * D1 - 0.5)
        mult = [(1 - fr[0]) * (1 - fr[1]), fr[0] * (1 - fr[1]), (1 - fr[0]) * fr[1], fr[0] * fr[1]]
        blend = np.zeros(4)
        for i in range(2):
            for j in range(2):
                blend += mult[i * 2 + j] * g1[ys1[i], xs1[j]]
        feats.extend(blend)
        cpos = uv * (L.dim >> lod)
        for P in (8, 4, 2):
            for a in range(2):
                for h in range(2):
                    c = math.floor(cpos[a])
                    tt = (c + 0.5 + (0 if h == 0 else 0.25 * P)) / P
                    feats.append(1 - 4 * abs(tt - math.floor(tt) - 0.5))
        feats.append(norm_lod)
        assert np.allclose(x[: len(feats), t], feats, atol=1e-12), t


def test_gather_scatter_adjoint(oracle):
    # <dX, J dG> == <J^T dX, dG>: the scatter is the exact adjoint of the
    # grid part of the gather (FP64, no noise), for both g0 modes and with
    # wrapping tiles on a coarser mip.
    for bilinear in (False, True):
        L, fd = _fd(oracle, bilinear, dim=128)
        rng = np.random.default_rng(1 + bilinear)
        M, T = 32, 32
        origins = np.array([20, 7, 31, 31], np.int32)
        D0, D1 = fd["g0_dim"], fd["g1_dim"]
        dg0 = rng.standard_normal((D0, D0, 4))
        dg1 = rng.standard_normal((D1, D1, 4))
        zero0, zero1 = np.zeros_like(dg0), np.zeros_like(dg1)
        jx = (oracle.draw_features(fd, origins, T, M, 0.0, dg0, dg1, dtype=np.float64)
              - oracle.draw_features(fd, origins, T, M, 0.0, zero0, zero1, dtype=np.float64))
        dx = rng.standard_normal(jx.shape)
        s0, s1 = oracle.scatter(fd, origins, T, M, dx, dg0.shape, dg1.shape)
        lhs = float(np.sum(dx * jx))
        rhs = float(np.sum(s0 * dg0) + np.sum(s1 * dg1))
        assert abs(lhs - rhs) <= 1e-10 * abs(lhs), (lhs, rhs)


# ---- GEMM / MLP --------------------------------------------------------------------
@pytest.mark.parametrize("op", [0, 1, 2])
def test_gemm_matches_numpy(oracle, op):
    rng = np.random.default_rng(op)
    n, m, k = 7, 33, 2500
    a = rng.standard_normal((n, k) if op != 1 else (k, n))
    b = rng.standard_normal((k, m) if op != 2 else (m, k))
    c = oracle.gemm(op, n, m, k, a, b, np.float64)
    want = (a if op != 1 else a.T) @ (b if op != 2 else b.T)
    assert np.allclose(c, want, rtol=1e-12, atol=1e-11)


def _mlp_setup(rng, F=12, H=16, O=3, B=40):
    x = rng.standard_normal((F, B))
    yt = rng.uniform(0, 1, (O, B))
    w0 = rng.standard_normal((H, F)) * 0.5
    w1 = rng.standard_normal((H, H)) * 0.4
    wout = rng.standard_normal((16, H)) * 0.4
    return x, yt, w0, w1, wout


@pytest.mark.parametrize("act", [0, 1])
def test_mlp_gradients_vs_finite_differences(oracle, act):
    # FP64 central differences of MSE = SSE/(O*B) against the hand-written
    # backward (A.4-A.6). Entries whose pre-activations sit within the step of
    # an activation kink are skipped by construction (random data, tiny h).
    rng = np.random.default_rng(10 + act)
    F, H, O, B = 12, 16, 3, 40
    x, yt, w0, w1, wout = _mlp_setup(rng, F, H, O, B)

    def mse(w0_, w1_, wout_, x_):
        ws = np.zeros(4 * H * B + O * B)
        sse = oracle.lib().o_mlp_forward_f64(F, H, O, act, B, oracle.ptr(np.ascontiguousarray(x_)),
                                            oracle.ptr(yt), oracle.ptr(np.ascontiguousarray(w0_)),
                                            oracle.ptr(np.ascontiguousarray(w1_)),
                                            oracle.ptr(np.ascontiguousarray(wout_)), oracle.ptr(ws), None, None)
        return sse / (O * B)

    _, dw0, dw1, dwout, dx = oracle.mlp_fwd_bwd(F, H, O, act, x, yt, w0, w1, wout)
    h = 1e-6
    for (arr, grad, name) in ((w0, dw0, "w0"), (w1, dw1, "w1"), (wout[:O], dwout, "wout"), (x, dx, "x")):
        for _ in range(12):
            idx = tuple(rng.integers(0, s) for s in grad.shape)
            args = [w0.copy(), w1.copy(), wout.copy(), x.copy()]
            which = {"w0": 0, "w1": 1, "wout": 2, "x": 3}[name]
            args[which][idx] += h
            fp = mse(*args)
            args[which][idx] -= 2 * h
            fm = mse(*args)
            fd = (fp - fm) / (2 * h)
            assert abs(fd - grad[idx]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (name, idx, fd, grad[idx])


def test_end_to_end_latent_gradient_vs_finite_differences(oracle):
    # Gather -> MLP -> loss -> backward -> scatter, FP64, against central
    # differences of the MSE with respect to individual latent values.
    L, fd = _fd(oracle, dim=32)
    rng = np.random.default_rng(21)
    M, T = 32, 8
    origins = np.array([3, 29], np.int32)
    F, H, O = fd["F"], 16, 3
    D0, D1 = fd["g0_dim"], fd["g1_dim"]
    g0 = rng.uniform(-1, 1, (D0, D0, 4))
    g1 = rng.uniform(-1, 1, (D1, D1, 4))
    yt = rng.uniform(0, 1, (O, T * T))
    w0 = rng.standard_normal((H, F)) * 0.3
    w1 = rng.standard_normal((H, H)) * 0.3
    wout = rng.standard_normal((16, H)) * 0.3

    def loss(g0_, g1_):
        x = oracle.draw_features(fd, origins, T, M, 0.5, g0_, g1_, dtype=np.float64)
        sse, *_ = oracle.mlp_fwd_bwd(F, H, O, 0, x, yt, w0, w1, wout)
        return sse / (O * T * T)

    x = oracle.draw_features(fd, origins, T, M, 0.5, g0, g1, dtype=np.float64)
    _, _, _, _, dx = oracle.mlp_fwd_bwd(F, H, O, 0, x, yt, w0, w1, wout)
    s0, s1 = oracle.scatter(fd, origins, T, M, dx, g0.shape, g1.shape)
    h = 1e-6
    touched0 = np.argwhere(np.abs(s0) > 0)
    touched1 = np.argwhere(np.abs(s1) > 0)
    for grid, grad, cells in ((0, s0, touched0), (1, s1, touched1)):
        for idx in cells[rng.choice(len(cells), 8, replace=False)]:
            idx = tuple(idx)
            a0, a1 = g0.copy(), g1.copy()
            (a0 if grid == 0 else a1)[idx] += h
            fp = loss(a0, a1)
            (a0 if grid == 0 else a1)[idx] -= 2 * h
            fm = loss(a0, a1)
            fdv = (fp - fm) / (2 * h)
            assert abs(fdv - grad[idx]) <= 1e-6 * abs(fdv) + 1e-10, (grid, idx, fdv, grad[idx])


def test_hardgelu_closed_form(oracle):
    z = np.array([-2.0, -1.5, -1.0, 0.0, 0.75, 1.5, 3.0], np.float32)
    a = oracle.act_forward(0, z)
    want = np.where(z < -1.5, 0, np.where(z < 1.5, z / 3 * (z + 1.5), z))
    assert np.allclose(a, want, rtol=1e-7)
    g = np.ones_like(z)
    d = oracle.act_backward(0, z, g)
    # derivative jumps at +-1.5: -0.5 -> 0 at -1.5 and 1.5 -> 1 at +1.5
    assert np.allclose(d, [0, -0.5, -1 / 6, 0.5, 1.0, 1.0, 1.0], rtol=1e-6)


# ---- optimiser / quantisation ------------------------------------------------------------
def test_adam_first_step_is_signed_lr(oracle):
    g = np.array([3e-3, -2e-5, 0.7, -4.0], np.float32)
    m = np.zeros(4, np.float32)
    v = np.zeros(4, np.float32)
    p = np.zeros(4, np.float32)
    b1 = np.float32(1) / (np.float32(1) - np.float32(0.9))
    b2 = np.float32(1) / (np.float32(1) - np.float32(0.999))
    oracle.adam(0.01, 0.9, 0.999, b1, b2, g, m, v, p)
    # -lr * g / (|g| + 1e-8): the epsilon shows at |g| = 2e-5 (ratio 0.9995)
    assert np.allclose(p, -0.01 * g / (np.abs(g) + 1e-8), rtol=1e-5)
    assert np.allclose(p, -0.01 * np.sign(g), rtol=1e-3)
    # 1 - 0.999f = 0.00099998713 in float32
    assert np.allclose(m, 0.1 * g, rtol=1e-6) and np.allclose(v, 0.001 * g * g, rtol=2e-5)


@pytest.mark.parametrize("bits", [1, 2, 4, 8])
def test_pack_exhaustive_and_hlsl_unpack(oracle, bits):
    # Every code once, packed LSB-first; unpacked the way the HLSL decoder
    # reads g0 ((v << (8-b)) ^ 0x80 as int8, ntc_decode.hlsli:251-253) and g1
    # ((v - 2^(b-1)) / 2^(b-1), :298-301).
    half = 2 ** (bits - 1)
    delta = 1.0 / half
    codes = np.arange(-half, half)
    per = 32 // bits
    vals = np.resize(codes, max(per, codes.size)) * delta
    vals = np.concatenate([vals, np.zeros((-vals.size) % per)]).astype(np.float32)
    words = oracle.quantize_pack(bits, vals)
    for i, val in enumerate(vals):
        w, s = divmod(i, per)
        v = (int(words[w]) >> (s * bits)) & ((1 << bits) - 1)
        assert v == int(round(val * half)) + half
        g0 = ((v << (8 - bits)) ^ 0x80) & 0xFF
        g0 = g0 - 256 if g0 >= 128 else g0
        assert g0 == int(round(val * 128))
        assert (v - half) / half == val


def test_quantize_grid_rounds_half_even_and_clamps(oracle):
    delta = np.float32(0.25)  # 3 bits
    x = np.array([0.125, 0.375, -0.125, 2.0, -2.0, 0.8], np.float32)
    oracle.quantize_grid(3, delta, x)
    assert np.array_equal(x, np.array([0.0, 0.5, -0.0, 0.75, -1.0, 0.75], np.float32))


def test_quantize_weights_closed_form(oracle):
    w = np.array([[0.5, -1.0, 0.25, 0.0], [0, 0, 0, 0], [127 / 254, 1.0, -0.5, 0.125]], np.float32)
    q, s = oracle.quantize_weights(w, 2.0)
    b = q.view(np.uint8).reshape(3, 4).view(np.int8)
    assert list(b[0]) == [64, -127, 32, 0]  # 63.5 -> 64 (lroundf half away)
    assert list(b[1]) == [0, 0, 0, 0] and s[1] == 2.0
    assert s[0] == np.float32(np.float32(1.0) / np.float32(127.0)) * np.float32(2.0)
    assert b[2][1] == 127


# ---- Lanczos -------------------------------------------------------------------------------
def test_lanczos_weights_symmetric_and_reference_form(oracle):
    w = oracle.lanczos_weights()
    assert np.array_equal(w, w[::-1])
    # the reference evaluates them in float (ntc_kernel.cu:569-579)
    pi = np.float32(math.pi)
    for i in range(-6, 6):
        x = np.float32((0.5 + i) * 0.5)
        a = np.float32(pi * x)
        b = np.float32(pi * x / np.float32(3))
        ref = np.float32(np.sin(a) / a) * np.float32(np.sin(b) / b)
        assert abs(w[i + 6] - ref) <= 4 * np.spacing(abs(ref)) + 1e-7


@pytest.mark.parametrize("value", [0, 1, 128, 200, 255])
def test_lanczos_constant_image_within_one_lsb(oracle, value):
    img = np.full((32, 32, 3), value, np.uint8)
    out = oracle.filter_lanczos(img).astype(int)
    assert np.all(np.abs(out - value) <= 1)


def test_lanczos_matches_separable_float64_filter(oracle):
    rng = np.random.default_rng(2)
    img = rng.integers(0, 256, (16, 16, 2)).astype(np.uint8)
    w = oracle.lanczos_weights().astype(np.float64)
    dim = 16

    def down(a, axis):
        a = np.moveaxis(a, axis, 0)
        out = np.zeros((dim // 2,) + a.shape[1:])
        for t in range(dim // 2):
            tot, tw = 0.0, 0.0
            for i in range(-6, 6):
                s = 2 * t + 1 + i
                if 0 <= s < dim:
                    tot = tot + a[s] * w[i + 6]
                    tw += w[i + 6]
            out[t] = tot / tw
        return np.moveaxis(out, 0, axis)

    ref = down(down(img / 255.0, 1), 0) * 255
    got = oracle.filter_lanczos(img).astype(np.float64)
    # the oracle truncates like the reference: result in (ref - 1 - eps, ref + eps]
    assert np.all(got <= np.clip(ref, 0, 255) + 1e-3) and np.all(got > np.clip(ref, 0, 255) - 1 - 1e-3)


# ---- evaluation decode at reduced precision --------------------------------------------------
def _e4m3_value(c):
    """OCP FP8 E4M3 by its definition: bias 7, 3 mantissa bits, subnormals at
    exponent 0."""
    s = -1.0 if c & 0x80 else 1.0
    e, m = (c >> 3) & 15, c & 7
    return s * (m / 8 * 2.0 ** -6 if e == 0 else (1 + m / 8) * 2.0 ** (e - 7))


def test_e4m3_codes_round_trip_and_round_to_nearest_even(oracle):
    codes = [c for c in range(256) if (c & 0x7F) != 0x7F]
    vals = {c: _e4m3_value(c) for c in codes}
    for c in codes:
        assert oracle.e4m3_decode(c) == vals[c], c
        assert oracle.e4m3_encode(vals[c]) == c, c
    assert vals[0x7E] == 448.0 and oracle.e4m3_decode(0x08) == 2.0 ** -6
    # brute force: nearest representable value, ties to the even code, saturation
    pos = sorted((vals[c], c) for c in codes if c < 0x80)
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-470, 470, 3000), rng.uniform(-0.03, 0.03, 3000),
                         [(a[0] + b[0]) / 2 for a, b in zip(pos, pos[1:])]]).astype(np.float32)
    for x in xs:
        a = min(abs(float(x)), 448.0)
        best = min(pos, key=lambda vc: (abs(vc[0] - a), vc[1] & 1))
        want = best[1] | (0x80 if np.signbit(x) else 0)
        assert oracle.e4m3_encode(float(x)) == want, (x, want, oracle.e4m3_encode(float(x)))
    assert oracle.e4m3_encode(1e9) == 0x7E and oracle.e4m3_encode(-1e9) == 0xFE


def test_bf16_round_is_nearest_even(oracle):
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.standard_normal(2000) * 10 ** rng.uniform(-6, 6, 2000),
                         [1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -(1.0 + 2 ** -8)]]).astype(np.float32)
    for x in xs:
        u = int(np.float32(x).view(np.uint32))
        lo = np.uint32(u & 0xFFFF0000).view(np.float32)
        hi = np.uint32((u & 0xFFFF0000) + 0x10000).view(np.float32)
        dlo, dhi = abs(float(x) - float(lo)), abs(float(hi) - float(x))
        want = lo if dlo < dhi else (hi if dhi < dlo else (lo if ((u >> 16) & 1) == 0 else hi))
        assert oracle.bf16_round(float(x)) == float(want), x


def test_clamp_s8_rounds_half_to_even(oracle):
    for x, want in ((2.5, 2), (3.5, 4), (-2.5, -2), (-0.5, 0), (127.4, 127), (127.5, 127), (300.0, 127),
                    (-128.5, -128), (-1000.0, -128), (0.49, 0)):
        assert oracle.clamp_s8(x) == want, x


def _routing_decoder(F, H, O, rng):
    """Each hidden unit and output reads one input with weight +-2^-k: every
    chain holds a single product."""
    w0 = np.zeros((H, F), np.float32)
    w1 = np.zeros((H, H), np.float32)
    wo = np.zeros((16, H), np.float32)
    p0, p1, po = rng.integers(0, F, H), rng.permutation(H), rng.permutation(H)[:O]
    for h in range(H):
        w0[h, p0[h]] = rng.choice([1.0, -0.5, 0.25, 2.0])
        w1[h, p1[h]] = rng.choice([1.0, 0.5, -1.0, 0.125])
    for o in range(O):
        wo[o, po[o]] = rng.choice([1.0, -0.25, 0.75])
    return w0, w1, wo, p0, p1, po


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_decode_routing_network_closed_form(oracle, precision):
    """Sparse decoders whose chains have one product each: the decode equals
    the precision's rounding applied at the stated points (computed here with
    single float32 operations), for ReLU."""
    F, H, O, B = 40, 64, 9, 257
    rng = np.random.default_rng(7 + precision)
    w0, w1, wo, p0, p1, po = _routing_decoder(F, H, O, rng)
    x = rng.uniform(-1, 1, (F, B)).astype(np.float32)
    dec = np.concatenate([w0.ravel(), w1.ravel(), wo.ravel()])
    f32 = np.float32
    if precision == 0:
        q = lambda v, r=None: v  # noqa: E731
    elif precision == 1:
        q = lambda v, r=None: np.vectorize(oracle.bf16_round, otypes=[np.float32])(v)  # noqa: E731
    else:
        q = lambda v, r: np.vectorize(lambda t: oracle.e4m3_decode(oracle.e4m3_encode(float(t))),  # noqa: E731
                                      otypes=[np.float32])(np.asarray(v, np.float32) * r)
    relu = lambda v: np.maximum(v, f32(0))  # noqa: E731
    if precision == 2:
        s = oracle.fp8_scales(F, H, O, dec, 0.9, 1.1)
        want_s = [f32(1) / f32(256), np.abs(w0).max() / f32(448), np.abs(w1).max() / f32(448),
                  np.abs(wo[:O]).max() / f32(448), f32(0.9) / f32(448), f32(1.1) / f32(448)]
        assert np.array_equal(s, np.array(want_s, np.float32))
        r = [f32(1) / f32(v) for v in s]
        xq = q(x, r[0])
        z0 = (q(w0[np.arange(H), p0], r[1])[:, None] * xq[p0]) * (s[1] * s[0])
        a0 = q(relu(z0), r[4])
        z1 = (q(w1[np.arange(H), p1], r[2])[:, None] * a0[p1]) * (s[2] * s[4])
        a1 = q(relu(z1), r[5])
        y = (q(wo[np.arange(O), po], r[3])[:, None] * a1[po]) * (s[3] * s[5])
        got = oracle.decode(precision, 1, F, 48, H, O, x, dec=dec, s8=s)
    else:
        xq = q(x)
        a0 = q(relu(q(w0[np.arange(H), p0])[:, None] * xq[p0]))
        a1 = q(relu(q(w1[np.arange(H), p1])[:, None] * a0[p1]))
        y = q(wo[np.arange(O), po])[:, None] * a1[po]
        got = oracle.decode(precision, 1, F, 48, H, O, x, dec=dec)
    assert np.array_equal(got, y.astype(np.float32))


def _int8_blob(oracle, w0, w1, wo, Fp, s_a1, s_a2):
    H, F = w0.shape
    O = wo.shape[0]
    w0p = np.zeros((H, Fp), np.float32)
    w0p[:, :F] = w0
    q0, s0 = oracle.quantize_weights(w0p, 1.0 / 128.0)
    q1, s1 = oracle.quantize_weights(w1, s_a1)
    q2, s2 = oracle.quantize_weights(wo, s_a2)
    wo16 = np.zeros(16 * H, np.uint8)
    wo16[: q2.nbytes] = q2.view(np.uint8)
    sc = np.zeros(2 * H + 16, np.float32)
    sc[:H], sc[H:2 * H], sc[2 * H:2 * H + O] = s0, s1, s2
    return np.concatenate([q0.view(np.uint8), q1.view(np.uint8), wo16, sc.view(np.uint8)])


@pytest.mark.parametrize("act", [0, 1])
def test_decode_int8_matches_hlsl_restatement(oracle, act):
    """The INT8 decode against a vectorised restatement of the deployed HLSL
    decoder (ntc_decode.hlsli:112-115, 400-502): int8 features, int32 dot
    products, float(acc) * scale, activation, * rcp_s_a, round half to even,
    clamp; output float(acc) * scale."""
    F, Fp, H, O, B = 40, 48, 64, 9, 333
    rng = np.random.default_rng(21 + act)
    w0 = (rng.uniform(-1, 1, (H, F)) / np.sqrt(F)).astype(np.float32)
    w1 = (rng.uniform(-1, 1, (H, H)) / np.sqrt(H)).astype(np.float32)
    wo = (rng.uniform(-1, 1, (O, H)) / np.sqrt(H)).astype(np.float32)
    x = rng.uniform(-1, 1, (F, B)).astype(np.float32)
    s_a1, s_a2 = np.float32(1.1 * 1.3 / 127), np.float32(1.1 * 0.9 / 127)
    blob = _int8_blob(oracle, w0, w1, wo, Fp, s_a1, s_a2)
    rcp = np.array([np.float32(1) / s_a1, np.float32(1) / s_a2], np.float32)
    got = oracle.decode(3, act, F, Fp, H, O, x, blob=blob, rcp=rcp)
    q0 = blob[: H * Fp].view(np.int8).reshape(H, Fp).astype(np.int64)
    q1 = blob[H * Fp: H * Fp + H * H].view(np.int8).reshape(H, H).astype(np.int64)
    qo = blob[H * Fp + H * H: H * Fp + H * H + 16 * H].view(np.int8).reshape(16, H)[:O].astype(np.int64)
    sc = blob[H * Fp + H * H + 16 * H:].view(np.float32)
    f32 = np.float32

    def act_f(z):
        if act == 1:
            return np.maximum(z, f32(0))
        return np.where(z < f32(-1.5), f32(0), np.where(z < f32(1.5), (z / f32(3)) * (z + f32(1.5)), z)).astype(f32)

    xi = np.zeros((Fp, B), np.int64)
    xi[:F] = np.clip(np.rint(x * f32(128)), -128, 127)
    a0 = np.clip(np.rint(act_f((q0 @ xi).astype(f32) * sc[:H, None]) * rcp[0]), -128, 127).astype(np.int64)
    a1 = np.clip(np.rint(act_f((q1 @ a0).astype(f32) * sc[H:2 * H, None]) * rcp[1]), -128, 127).astype(np.int64)
    y = (qo @ a1).astype(f32) * sc[2 * H:2 * H + O, None]
    assert np.array_equal(got, y)
    # and within the quantisation error of the FP32 decode
    ref = oracle.decode(0, act, F, Fp, H, O, x, dec=np.concatenate([w0.ravel(), w1.ravel(), np.concatenate(
        [wo, np.zeros((16 - O, H), np.float32)]).ravel()]))
    assert np.max(np.abs(got - ref)) < 0.1 * np.max(np.abs(ref))


# ---- parameter init (libopenntc.cpp:60-78, 132-152, 483-501) --------------------
class _PyMT:
    """std::mt19937 written out here from the published algorithm
    (Matsumoto-Nishimura 1998; init_genrand seeding), independent of the
    oracle's C generator."""

    def __init__(self, seed):
        self.mt = [0] * 624
        self.mt[0] = seed & 0xFFFFFFFF
        for i in range(1, 624):
            self.mt[i] = (1812433253 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
        self.i = 624

    def next(self):
        if self.i >= 624:
            for k in range(624):
                y = (self.mt[k] & 0x80000000) | (self.mt[(k + 1) % 624] & 0x7FFFFFFF)
                self.mt[k] = self.mt[(k + 397) % 624] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
            self.i = 0
        y = self.mt[self.i]
        self.i += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y


def test_python_mt19937_known_answer():
    m = _PyMT(5489)
    for _ in range(9999):
        m.next()
    assert m.next() == 4123659995


def _draws(m, n, lo, hi):
    # std::uniform_real_distribution<float>(lo, hi) over 24-bit uniforms:
    # lo + (hi - lo) * u in float32 (the oracle's stated definition of an
    # implementation-defined distribution)
    f = np.float32
    u = np.array([(m.next() >> 8) for _ in range(n)], np.float64) * 2.0 ** -24
    return (f(lo) + (f(hi) - f(lo)) * u.astype(np.float32)).astype(np.float32)


def test_init_params_draw_order_ranges_and_fan_in(oracle):
    """Replays BeginTraining's draws with an independent generator: per level
    g0 then g1 (libopenntc.cpp:485-493), then W0, W1, Wout (:496-504), each
    Kaiming-uniform with bound 1/sqrt(row length) (:60-78). A swapped order,
    levels drawn grid-major, or a fan-in taken from the wrong axis fails."""
    L = oracle.layout_custom(4, 2, 4, 8, 2, 64, 1, 1, 64, out_dim=7, mip_min_dim=4)
    assert L.level_count >= 2
    H, F = L.hidden_dim, L.feature_dim
    P = oracle.init_params(oracle.MT(99), L, -0.5, 0.5)
    m = _PyMT(99)
    for lvl in range(L.level_count):
        n0 = L.g0_dim[lvl] ** 2 * L.g0_channels
        n1 = L.g1_dim[lvl] ** 2 * L.g1_channels
        assert np.array_equal(P.g0[lvl].ravel(), _draws(m, n0, -0.5, 0.5)), f"g0 level {lvl}"
        assert np.array_equal(P.g1[lvl].ravel(), _draws(m, n1, -0.5, 0.5)), f"g1 level {lvl}"
    b0 = np.float32(1.0) / np.sqrt(np.float32(F))
    b1 = np.float32(1.0) / np.sqrt(np.float32(H))
    w0 = P.dec[: H * F]
    w1 = P.dec[H * F: H * F + H * H]
    wo = P.dec[H * F + H * H:]
    assert np.array_equal(w0, _draws(m, H * F, -b0, b0))
    assert np.array_equal(w1, _draws(m, H * H, -b1, b1))
    assert np.array_equal(wo, _draws(m, 16 * H, -b1, b1))
    # the bounds are the fan-in ones: each tensor fills its own range
    for w, b in ((w0, b0), (w1, b1), (wo, b1)):
        assert np.all(np.abs(w) <= b) and np.max(np.abs(w)) > 0.97 * b
    assert F != H and abs(b0 - b1) > 0.05  # the two bounds are told apart
    # Adam moments start at zero
    assert not np.any(P.dec_m) and not np.any(P.dec_v)
    assert all(not np.any(a) for a in P.g0m + P.g0v + P.g1m + P.g1v)
    # and the next draw after the init is the generator's next output
    m2 = oracle.MT(99)
    oracle.init_params(m2, L, -0.5, 0.5)
    assert m2.next() == m.next()


# ---- calibration (libopenntc.cpp:898-970) ----------------------------------------
def _hardgelu64(z):
    return np.where(z < -1.5, 0.0, np.where(z < 1.5, (z / 3.0) * (z + 1.5), z))


def _act64(act, z):
    return np.maximum(z, 0.0) if act == 1 else _hardgelu64(z)


@pytest.mark.parametrize("act", [0, 1])
def test_calibrate_max_abs_vs_float64_walk(oracle, act):
    """max |A0| and max |A1| over every texel of every mip against a numpy
    FP64 forward over each whole mip (gathered as one mip-sized tile). The
    decoder is built so that each maximum sits where a plausible walk mistake
    misses it: max |A0| in the last tile of mip 0's batch (a latent bump read
    only there at full strength), max |A1| in the last mip (a unit driven by
    the lod slot), and a unit whose pre-activation reaches -2x the maximum
    (the maximum of the pre-activations would differ)."""
    L = oracle.layout_custom(4, 4, 4, 8, 2, 64, 0, act, 512, mip_min_dim=4)
    H, F = L.hidden_dim, L.feature_dim
    assert L.lod_slot == 1 and L.mip_count == 8
    lod = 4 * L.g0_channels + L.g1_channels + 12
    P = oracle.init_params(oracle.MT(5), L, -0.5, 0.5)
    P.g0[0][200, 200, 0] = 30.0  # corner 0 of texel (401, 401): mip 0, last tile
    rng = np.random.default_rng(6)
    w0 = (rng.uniform(-1, 1, (H, F)) * 0.05).astype(np.float32)
    w0[0, 0], w0[0, lod] = 1.0, -100.0   # +30 at mip 0, at most ~16 at mip 1
    w0[1, lod] = 8.0                     # largest at the last mip (lod feature 1)
    w0[2, 0] = -2.0                      # pre-activation -60 at the bump
    w1 = (rng.uniform(-1, 1, (H, H)) * 0.01).astype(np.float32)
    w1[0, 1] = 5.0                       # A1 unit 0 follows A0 unit 1
    P.dec[: H * F] = w0.ravel()
    P.dec[H * F: H * F + H * H] = w1.ravel()
    tr = oracle.Trainer(L, P, [np.zeros((L.mip_dim[m], L.mip_dim[m], L.out_dim), np.uint8)
                               for m in range(L.mip_count)], max_batch=4 * 65536)
    got = tr.calibrate()
    w0d, w1d = w0.astype(np.float64), w1.astype(np.float64)
    mx = np.zeros((L.mip_count, 2))
    zmax = 0.0
    for mip in range(L.mip_count):
        M = L.mip_dim[mip]
        lvl = oracle.level_for_lod(L, mip)
        fd = oracle.feature_desc(L, lvl)
        x = oracle.draw_features(fd, np.array([0, 0], np.int32), M, M, mip / (L.mip_count - 1),
                                 P.g0[lvl], P.g1[lvl], dtype=np.float64)
        z0 = w0d @ x
        a0 = _act64(act, z0)
        a1 = _act64(act, w1d @ a0)
        mx[mip] = np.abs(a0).max(), np.abs(a1).max()
        zmax = max(zmax, np.abs(z0).max())
    want = mx.max(axis=0)
    assert np.argmax(mx[:, 0]) == 0 and np.argmax(mx[:, 1]) == L.mip_count - 1
    assert mx[1:, 0].max() < 0.6 * want[0] and mx[:-1, 1].max() < 0.9 * want[1] and zmax > 1.9 * want[0]
    assert np.allclose(got, want, rtol=2e-5, atol=0), (got, want)


# ---- texture packaging (ntc_kernel.cu:533-567, 665-686; libopenntc.cpp:1195-1199, 1324-1362)
def test_build_package_channel_map_sort_and_interleave(oracle):
    """Hand-built sources whose RGBA bytes name their source, channel and
    position. The package must hold, per texel, the selected channels of each
    source in channel-string order, the sources stably sorted by semantic
    (Albedo 1 < Metallic 6 < Normal 7 < Roughness 9; the two Metallic sources
    keep their manifest order), and each coarser mip the per-channel Lanczos
    filter of the one above."""
    L = oracle.layout_custom(4, 4, 4, 8, 2, 64, 1, 1, 16, out_dim=9, mip_min_dim=4)
    dim = 16
    yy, xx = np.meshgrid(np.arange(dim), np.arange(dim), indexing="ij")

    def src(sid):
        img = np.zeros((dim, dim, 4), np.uint8)
        for c in range(4):
            img[:, :, c] = (sid * 64 + c * 16 + (xx + 3 * yy) % 16).astype(np.uint8)
        return img

    # (rgba, semantic, channel string) in manifest order
    sources = [(src(0), 7, "RGB"),   # Normal
               (src(1), 6, "B"),     # Metallic #1
               (src(2), 1, "BGR"),   # Albedo, reversed channel order
               (src(3), 9, "A"),     # Roughness
               (src(1), 6, "R")]     # Metallic #2 (same image, other channel)
    pk = oracle.build_package(sources, L)
    # expected out channels: (source id, RGBA index)
    order = [(2, 2), (2, 1), (2, 0), (1, 2), (1, 0), (0, 0), (0, 1), (0, 2), (3, 3)]
    assert pk[0].shape == (dim, dim, len(order))
    for o, (sid, c) in enumerate(order):
        want = (sid * 64 + c * 16 + (xx + 3 * yy) % 16).astype(np.uint8)
        assert np.array_equal(pk[0][:, :, o], want), o
    assert len(pk) == L.mip_count
    for m in range(1, L.mip_count):
        M = L.mip_dim[m]
        assert pk[m].shape == (M, M, len(order))
        for o, (sid, c) in enumerate(order):
            plane = pk[m - 1][:, :, o:o + 1]
            assert np.array_equal(pk[m][:, :, o], oracle.filter_lanczos(plane)[:, :, 0]), (m, o)


# ---- FP64 decode (the error oracle of every evaluation precision) ---------------------
@pytest.mark.parametrize("act", [0, 1])
def test_decode_f64_vs_numpy_chain(oracle, act):
    """Wout[:O] act(W1 act(W0 x)) in numpy float64 (HardGELU from
    ntc_kernel.cu:13-20). Rows of Wout from O on must not reach the output;
    pre-activations straddle the HardGELU knees at +-1.5."""
    F, H, O, B = 36, 64, 9, 301
    rng = np.random.default_rng(40 + act)
    w0 = rng.uniform(-1, 1, (H, F)) * 0.6
    w1 = rng.uniform(-1, 1, (H, H)) * 0.35
    wo = rng.uniform(-1, 1, (16, H)) * 0.3
    wo[O:] = 1e6  # never read
    x = rng.uniform(-1, 1, (F, B))
    dec = np.concatenate([w0.ravel(), w1.ravel(), wo.ravel()])
    got = oracle.decode_f64(act, F, H, O, x, dec)
    z0 = w0 @ x
    assert np.mean(np.abs(z0) > 1.5) > 0.05 and np.mean(np.abs(z0) < 1.5) > 0.3
    want = wo[:O] @ _act64(act, w1 @ _act64(act, z0))
    assert got.shape == (O, B)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)
