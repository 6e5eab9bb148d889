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
