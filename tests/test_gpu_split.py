"""The training step in three launches (ntc_split_train_step: gather, MLP on
BF16 MFMA, latent scatter) against the same references as the one-launch
kernel (tests/test_gpu_fused.py): the step with the kernel's BF16 rounding
points to 2e-3 and the FP64 oracle within BF16's precision loss; its feature
rows bit for bit against the oracle's gather rounded to BF16; and against the
one-launch kernel element by element (the same BF16 products per texel, FP32
sums grouped differently)."""
import ctypes as C

import numpy as np
import pytest

from openntc_amd import _lib as L
from tests import gpu_util as G
from tests.test_gpu_fused import REL_TOL_BF16, _check_step, _outputs, _step_case, bf16, rel

pytestmark = pytest.mark.gpu

PASS = 128


def pass_rows(B, T):
    """Batch index of every pass-order row of the split step's workspace
    (-1: past the batch): 16 x 8 texel blocks of each tile for T >= 16, tile
    after tile, else the batch in order."""
    npass = -(-B // PASS)
    if T < 16:
        t = np.arange(npass * PASS)
        return np.where(t < B, t, -1)
    lt = T.bit_length() - 1
    lbh = 3
    lrow = lt - 4
    ltile = lrow + lt - lbh
    p = np.arange(npass)[:, None]
    tl = np.arange(PASS)[None, :]
    b = p >> ltile
    q = p & ((1 << ltile) - 1)
    lx = ((q & ((1 << lrow) - 1)) << 4) + (tl & 15)
    ly = ((q >> lrow) << lbh) + (tl >> 4)
    return ((b * T + ly) * T + lx).ravel()


def _max_blocks(lib):
    maxb = C.c_int32()
    G.ok(lib.ntc_fused_max_blocks(C.byref(maxb)))
    return maxb.value


def _split_step(lib, cs, max_blocks=0):
    """One split step on cs's inputs; returns its outputs, workgroup count and
    workspace."""
    fdc, H, act, B = cs["fdc"], cs["H"], cs["act"], cs["B"]
    assert lib.ntc_split_supported(G.byref(fdc), H, act) == 1
    blocks = lib.ntc_split_blocks_for(G.byref(fdc), H, act, B, max_blocks or _max_blocks(lib))
    assert blocks == min(-(-B // PASS), max_blocks or _max_blocks(lib))
    out = _outputs(cs, blocks)
    nbytes = lib.ntc_split_workspace_bytes(G.byref(fdc), H, act, B)
    ws = G.empty((nbytes,), np.uint8)
    a = cs["a"]
    a.dg0, a.dg1, a.dw_partials, a.sse_partials, a.blocks = (G.p(out["dg0"]), G.p(out["dg1"]), G.p(out["part"]),
                                                             G.p(out["ssep"]), blocks)
    G.ok(lib.ntc_split_train_step(G.byref(a), G.p(ws), nbytes, None), "ntc_split_train_step")
    return out, blocks, ws


def _fused_step(lib, cs):
    blocks = lib.ntc_fused_blocks_for(G.byref(cs["fdc"]), cs["H"], cs["act"], cs["B"], _max_blocks(lib))
    out = _outputs(cs, blocks)
    a = cs["a"]
    a.dg0, a.dg1, a.dw_partials, a.sse_partials, a.blocks = (G.p(out["dg0"]), G.p(out["dg1"]), G.p(out["part"]),
                                                             G.p(out["ssep"]), blocks)
    G.ok(lib.ntc_fused_train_step(G.byref(a), None), "ntc_fused_train_step")
    return out, blocks


def _reduced(lib, cs, out, blocks):
    """Decoder gradient (ordered sum of the partials, lr = 0), latent
    gradients and SSE of one step's outputs."""
    n_dec = cs["dec"].size
    grad = G.empty((n_dec,), np.float32)
    m = G.empty((n_dec,), np.float32)
    v = G.empty((n_dec,), np.float32)
    sse = G.empty((1,), np.float64)
    G.ok(lib.ntc_decoder_update(cs["H"], cs["F"], cs["Od"], blocks, G.p(out["part"]), 0.0, 0.9, 0.999, 1.0, 1.0,
                                G.p(grad), G.p(m), G.p(v), G.p(G.dev(cs["dec"])), G.p(out["ssep"]), G.p(sse), None))
    return G.host(grad), G.host(out["dg0"]), G.host(out["dg1"]), float(G.host(sse)[0])


LODS = [(0, 2), (1, 1), (2, 3), (3, 4), (4, 5), (5, 2), (6, 7)]


@pytest.mark.parametrize("lod,n_tiles", LODS)
def test_split_step_vs_references(oracle, lod, n_tiles):
    lib = G.lib()
    cs = _step_case(oracle, "c3_bilinear_h128", lod, n_tiles)
    out, blocks, _ = _split_step(lib, cs)
    _check_step(cs, out, blocks)


@pytest.mark.parametrize("lod,n_tiles", [(7, 1), (8, 1), (7, 3), (8, 2)])
def test_split_smallest_mips(oracle, lod, n_tiles):
    """Batches of 1-12 texels: one partly filled linear pass."""
    lib = G.lib()
    cs = _step_case(oracle, "c3_bilinear_h128", lod, n_tiles, 256, 1)
    assert cs["B"] == n_tiles * (256 >> lod) ** 2
    out, blocks, _ = _split_step(lib, cs)
    assert blocks == 1
    _check_step(cs, out, blocks)


@pytest.mark.parametrize("dim,lod,n_tiles", [(256, 0, 2), (256, 3, 4), (256, 5, 2), (1024, 0, 1), (1024, 1, 2),
                                             (2048, 0, 2), (2048, 2, 1)])
def test_split_rows_bitexact(oracle, dim, lod, n_tiles):
    """The workspace rows: features bit for bit the oracle's gather (FP32,
    with its noise) rounded to BF16, zero past the batch; dX of the grid slots
    within FP32 accumulation of the BF16-rounding reference."""
    lib = G.lib()
    cs = _step_case(oracle, "c3_bilinear_h128", lod, n_tiles, dim)
    _, _, ws = _split_step(lib, cs)
    B, T, F = cs["B"], cs["T"], cs["F"]
    fxs = (F + 7) // 8 * 8
    rows = pass_rows(B, T)
    nx = rows.size * fxs * 2
    nx_al = (nx + 255) // 256 * 256
    raw = G.host(ws)
    x = raw[:nx].view(np.uint16).reshape(rows.size, fxs)
    valid = rows >= 0
    want = bf16(cs["x32"]).astype(np.float32).view(np.uint32) >> 16
    got = x[valid, :F].astype(np.uint32)
    exp = want[:, rows[valid]].T
    bad = int(np.count_nonzero(got != exp))
    assert bad == 0, (bad, got.size)
    assert not x[~valid].any()
    assert not x[:, F:].any()
    gs = 32  # grid slots of the layout (16 g0 + 16 g1 channels), one dX row each
    dxr = raw[nx_al: nx_al + rows.size * gs * 4].view(np.float32).reshape(rows.size, gs)
    want_dx = cs["bdx"][:gs, rows[valid]].T
    got_dx = dxr[valid]
    assert rel(got_dx, want_dx) < REL_TOL_BF16, rel(got_dx, want_dx)


@pytest.mark.parametrize("dim,lod,n_tiles", [(256, 0, 2), (256, 1, 1), (256, 2, 3), (256, 4, 5), (256, 6, 7),
                                             (1024, 0, 1), (1024, 0, 3), (1024, 2, 1), (2048, 0, 1), (2048, 1, 1),
                                             (2048, 3, 1)])
def test_split_agrees_with_one_launch(oracle, dim, lod, n_tiles):
    """Split and one-launch steps on the same inputs form the same BF16
    products per texel; every decoder- and latent-gradient element agrees to
    FP32 rounding of differently grouped sums (relative to the largest)."""
    lib = G.lib()
    cs = _step_case(oracle, "c3_bilinear_h128", lod, n_tiles, dim)
    so, sb, _ = _split_step(lib, cs)
    rs = _reduced(lib, cs, so, sb)
    fo, fb = _fused_step(lib, cs)
    rf = _reduced(lib, cs, fo, fb)
    assert abs(rs[3] - rf[3]) <= 1e-9 * rf[3]
    for k, name in ((0, "dW"), (1, "dG0"), (2, "dG1")):
        scale = float(np.max(np.abs(rf[k]))) + 1e-30
        err = np.abs(rs[k] - rf[k]) / scale
        worst = int(np.argmax(err))
        assert float(err.max()) < 1e-5, (name, float(err.max()), worst)


@pytest.mark.parametrize("dim,lod,n_tiles", [(256, 0, 1), (2048, 0, 1), (256, 3, 4)])
@pytest.mark.parametrize("workgroups", [16, 5, 1])
def test_split_many_passes_per_workgroup(oracle, dim, lod, n_tiles, workgroups):
    """Few MLP workgroups, each walking many passes through its feature-row
    prefetch: the same results as one pass per workgroup up to FP32 grouping."""
    lib = G.lib()
    cs = _step_case(oracle, "c3_bilinear_h128", lod, n_tiles, dim)
    so, sb, _ = _split_step(lib, cs, workgroups)
    assert sb == min(workgroups, -(-cs["B"] // PASS))
    r = _reduced(lib, cs, so, sb)
    fo, fb, _ = _split_step(lib, cs)
    full = _reduced(lib, cs, fo, fb)
    assert abs(r[3] - full[3]) <= 1e-9 * full[3]
    for k, name in ((0, "dW"), (1, "dG0"), (2, "dG1")):
        scale = float(np.max(np.abs(full[k]))) + 1e-30
        err = float(np.max(np.abs(r[k] - full[k]))) / scale
        assert err < 1e-5, (name, err)


def test_split_rejects_bad_arguments(oracle):
    """Workspace too small or misaligned, blocks not the MLP launch's
    workgroup count, a layout without the split form: error codes, nothing
    launched."""
    lib = G.lib()
    cs = _step_case(oracle, "c3_bilinear_h128", 2, 1)
    fdc, H, act, B = cs["fdc"], cs["H"], cs["act"], cs["B"]
    blocks = lib.ntc_split_blocks_for(G.byref(fdc), H, act, B, _max_blocks(lib))
    out = _outputs(cs, blocks)
    nbytes = lib.ntc_split_workspace_bytes(G.byref(fdc), H, act, B)
    ws = G.empty((nbytes + 512,), np.uint8)
    a = cs["a"]
    a.dg0, a.dg1, a.dw_partials, a.sse_partials, a.blocks = (G.p(out["dg0"]), G.p(out["dg1"]), G.p(out["part"]),
                                                             G.p(out["ssep"]), blocks)
    assert lib.ntc_split_train_step(G.byref(a), G.p(ws), nbytes - 1, None) == -2
    assert lib.ntc_split_train_step(G.byref(a), None, nbytes, None) == -2
    assert lib.ntc_split_train_step(G.byref(a), G.p(ws) + 16, nbytes, None) == -1
    a.blocks = blocks + 1
    assert lib.ntc_split_train_step(G.byref(a), G.p(ws), nbytes, None) == -1
    a.blocks = blocks
    a.hidden_dim = 64
    assert lib.ntc_split_train_step(G.byref(a), G.p(ws), nbytes, None) == -1
    assert lib.ntc_split_supported(G.byref(fdc), 64, act) == 0
    assert lib.ntc_split_workspace_bytes(G.byref(fdc), 64, act, B) == 0
    a.hidden_dim = H
    G.ok(lib.ntc_split_train_step(G.byref(a), G.p(ws), nbytes, None))
    _check_step(cs, out, blocks)


def test_context_uses_split_form_and_matches_one_launch():
    """A Context with the H = 128 bilinear ReLU layout trains through the
    three launches by default (ntc_ctx_fused_form 0) and through the
    one-launch kernel when asked (form 1); on the same seed both reach the
    same PSNR (1024^2, one 256^2 tile per step: deferred grid steps; 600
    steps)."""
    import openntc_amd as ntc

    lib = G.lib()
    tex = ntc.synthetic_pbr(1024, 4242)
    psnr = {}
    for form in (0, 1):
        ctx = ntc.Context()
        lay = ntc.LayoutInfo(4, 4, 16, 16, 2, 128, ntc.FeatureLayout.Bilinear, ntc.Activation.ReLU)
        assert ctx.InitEx(layout=lay, min_dimension=64) == ntc.Result.Success
        assert ctx.LoadTextures(tex, 1024) == ntc.Result.Success
        assert lib.ntc_ctx_set_fused_form(ctx._h, 2) != 0
        assert lib.ntc_ctx_set_fused_form(ctx._h, form) == 0
        assert ctx.BeginTrainingEx(600, 1, 1e-2, 1e-3, -0.5, 0.5, ntc.TrainPath.Fused) == ntc.Result.Success
        assert lib.ntc_ctx_fused_form(ctx._h) == form
        prog = ctx.TrainUntilComplete()
        assert prog.phase == ntc.TrainPhase.TrainComplete, prog
        psnr[form] = ctx.Eval().psnr
        print(f"form {form}: {psnr[form]:.4f} dB", flush=True)
    assert psnr[0] > 20.0
    assert abs(psnr[0] - psnr[1]) < 0.1, psnr


@pytest.mark.parametrize("dim,lod,n_tiles,n_next", [(1024, 0, 1, 3), (1024, 1, 2, 2), (2048, 1, 1, 4)])
def test_split_tail_matches_separate_launches(oracle, dim, lod, n_tiles, n_next):
    """ntc_split_tail (the scatter, the next step's catch-up and the decoder
    update in one launch) against the three as separate launches
    (ntc_split_scatter, ntc_grid_lazy_catchup_f32 with this step's tiles as
    its skip, ntc_decoder_update_bf16) after the same gather and MLP launch:
    the catch-up's p, m, v and step counters and the decoder's gradient,
    moments, parameters, BF16 table and SSE bit for bit; the latent gradients
    to FP32 atomic order. Mips larger than a tile, so that the next step's
    rectangles reach cells outside this step's."""
    lib = G.lib()
    cs = _step_case(oracle, "c3_bilinear_h128", lod, n_tiles, dim)
    assert cs["M"] > cs["T"]
    fdc, H, act, B, Od, a = cs["fdc"], cs["H"], cs["act"], cs["B"], cs["Od"], cs["a"]
    blocks = lib.ntc_split_blocks_for(G.byref(fdc), H, act, B, _max_blocks(lib))
    out = _outputs(cs, blocks)
    nbytes = lib.ntc_split_workspace_bytes(G.byref(fdc), H, act, B)
    ws = G.empty((nbytes,), np.uint8)
    a.dw_partials, a.sse_partials, a.blocks = G.p(out["part"]), G.p(out["ssep"]), blocks
    a.dg0, a.dg1 = G.p(out["dg0"]), G.p(out["dg1"])  # set per run below; the gather checks them
    G.ok(lib.ntc_split_gather(G.byref(a), G.p(ws), nbytes, None), "gather")
    G.ok(lib.ntc_split_mlp(G.byref(a), G.p(ws), nbytes, None), "mlp")

    rng = np.random.default_rng(100 + lod)
    shapes = [cs["g0"].shape, cs["g1"].shape]
    grids = []
    for k, shp in enumerate(shapes):
        p0 = rng.uniform(-1, 1, shp).astype(np.float32)
        m0 = (rng.standard_normal(shp) * 1e-3).astype(np.float32)
        v0 = rng.uniform(0, 1e-6, shp).astype(np.float32)
        D = shp[0]
        m0[: D // 4] = 0.0  # cells no gradient has reached yet
        v0[: D // 4] = 0.0
        grids.append((p0, m0, v0))
    K, upto = 20, 13
    consts = np.stack([rng.uniform(0.002, 0.01, K), 1.0 / (1 - 0.9 ** np.arange(1, K + 1)),
                       1.0 / (1 - 0.999 ** np.arange(1, K + 1))], 1).astype(np.float32)
    last = [rng.integers(0, upto + 1, shp[0] * shp[1]).astype(np.int32) for shp in shapes]
    n_dec = cs["dec"].size
    dec_m = (rng.standard_normal(n_dec) * 1e-3).astype(np.float32)
    dec_v = rng.uniform(0, 1e-6, n_dec).astype(np.float32)
    table0 = G.host(_table_copy(lib, cs))
    M, T = cs["M"], cs["T"]
    nxt = G.dev(rng.integers(0, M, 2 * n_next).astype(np.int32))
    tn = G.tiles(nxt, n_next, T, M, 0.0)
    bounds = (-1.0, 1.0 - 2.0 / 16, -1.0, 1.0 - 2.0 / 16)
    tab = G.dev(consts.ravel())
    lr, b1c, b2c = 0.004, 1.0 / (1 - 0.9 ** 6), 1.0 / (1 - 0.999 ** 6)

    def state():
        st = dict(g=[[G.dev(x) for x in g] for g in grids], last=[G.dev(x) for x in last],
                  dg=[G.dev(np.zeros(shp, np.float32)) for shp in shapes], grad=G.dev(np.zeros(n_dec, np.float32)),
                  m=G.dev(dec_m), v=G.dev(dec_v), p=G.dev(cs["dec"]), table=G.dev(table0),
                  sse=G.dev(np.zeros(1, np.float64)))
        st["adam"] = L.GridAdam(G.p(st["g"][0][1]), G.p(st["g"][0][2]), G.p(st["g"][1][1]), G.p(st["g"][1][2]),
                                0.0, 0.9, 0.999, 0.0, 0.0, *bounds)
        return st

    def result(st):
        r = {f"{n}{k}": G.host(st["g"][k][j]) for k in range(2) for j, n in enumerate("pmv")}
        r.update({f"last{k}": G.host(st["last"][k]) for k in range(2)})
        r.update({n: G.host(st[n]) for n in ("grad", "m", "v", "p", "table", "sse")})
        return r, [G.host(d) for d in st["dg"]]

    sep, tail = state(), state()
    # separate launches
    a.dg0, a.dg1 = G.p(sep["dg"][0]), G.p(sep["dg"][1])
    G.ok(lib.ntc_split_scatter(G.byref(a), G.p(ws), nbytes, None), "scatter")
    G.ok(lib.ntc_grid_lazy_catchup_f32(G.byref(fdc), G.byref(tn), G.byref(a.tiles), G.p(sep["g"][0][0]),
                                       G.p(sep["g"][1][0]), G.byref(sep["adam"]), G.p(tab), upto,
                                       G.p(sep["last"][0]), G.p(sep["last"][1]), None), "catch-up")
    G.ok(lib.ntc_decoder_update_bf16(G.byref(fdc), H, Od, blocks, G.p(out["part"]), lr, 0.9, 0.999, b1c, b2c,
                                     G.p(sep["grad"]), G.p(sep["m"]), G.p(sep["v"]), G.p(sep["p"]),
                                     G.p(sep["table"]), G.p(out["ssep"]), G.p(sep["sse"]), None), "decoder update")
    # one launch
    a.dg0, a.dg1 = G.p(tail["dg"][0]), G.p(tail["dg"][1])
    t = L.SplitTailArgs()
    t.catch_feat, t.catch_tiles, t.catch_skip = C.addressof(fdc), C.addressof(tn), C.addressof(a.tiles)
    t.catch_g0, t.catch_g1 = G.p(tail["g"][0][0]), G.p(tail["g"][1][0])
    t.catch_adam, t.catch_steps, t.catch_upto = C.addressof(tail["adam"]), G.p(tab), upto
    t.catch_last0, t.catch_last1 = G.p(tail["last"][0]), G.p(tail["last"][1])
    t.dec_feat = C.addressof(fdc)
    t.dec_lr, t.dec_beta1, t.dec_beta2, t.dec_bias1, t.dec_bias2 = lr, 0.9, 0.999, b1c, b2c
    t.dec_grad, t.dec_m, t.dec_v, t.dec_p = G.p(tail["grad"]), G.p(tail["m"]), G.p(tail["v"]), G.p(tail["p"])
    t.dec_table, t.sse_out = G.p(tail["table"]), G.p(tail["sse"])
    G.ok(lib.ntc_split_tail(G.byref(a), G.p(ws), nbytes, G.byref(t), None), "ntc_split_tail")
    rs, dgs = result(sep)
    rt, dgt = result(tail)
    for name in rs:
        assert np.array_equal(rs[name], rt[name]), (name, int(np.count_nonzero(rs[name] != rt[name])))
    # the catch-up did work: counters moved to upto inside the next rectangles
    assert any((rt[f"last{k}"] == upto).sum() > (last[k] == upto).sum() for k in range(2))
    for k in range(2):
        scale = float(np.abs(dgs[k]).max())
        assert scale > 0
        np.testing.assert_allclose(dgt[k], dgs[k], rtol=0, atol=1e-5 * scale)


def _table_copy(lib, cs):
    """A device copy of the case's packed BF16 decoder table."""
    n = lib.ntc_fused_table_elems(G.byref(cs["fdc"]), cs["H"])
    d_dec = G.dev(cs["dec"])
    tab = G.empty((n,), np.uint16)
    G.ok(lib.ntc_pack_decoder_bf16(G.byref(cs["fdc"]), cs["H"], cs["Od"], G.p(d_dec), G.p(tab), None))
    return tab
