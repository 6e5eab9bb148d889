"""ntc_fused_train_step_noalloc (the training step with its partials in the
library's own device array, no workspace argument, no allocation) against
ntc_fused_train_step + ntc_reduce_decoder_partials with caller-owned
partials: the decoder gradient and SSE bit for bit, the latent gradients to
atomic-order rounding; two calls in a row, the second on a smaller batch,
then the size limit."""
import ctypes as C

import numpy as np
import pytest

from openntc_amd import _lib as L

pytestmark = pytest.mark.gpu

from tests import gpu_util as G  # noqa: E402
from tests.test_gpu_fused import _outputs, _step_case, rel  # noqa: E402


def _blocks(cs):
    lib = G.lib()
    maxb = C.c_int32()
    G.ok(lib.ntc_fused_max_blocks(C.byref(maxb)))
    return lib.ntc_fused_blocks_for(G.byref(cs["fdc"]), cs["H"], cs["act"], cs["B"], maxb.value)


def _reference(cs, blocks):
    lib = G.lib()
    out = _outputs(cs, blocks)
    a = cs["a"]
    a.dg0, a.dg1, a.dw_partials, a.sse_partials, a.blocks = (G.p(out["dg0"]), G.p(out["dg1"]), G.p(out["part"]),
                                                             G.p(out["ssep"]), blocks)
    G.ok(lib.ntc_fused_train_step(G.byref(a), None), "ntc_fused_train_step")
    grad = G.empty((cs["dec"].size,), np.float32)
    sse = G.empty((1,), np.float64)
    G.ok(lib.ntc_reduce_decoder_partials(cs["H"], cs["F"], blocks, G.p(out["part"]), G.p(grad), G.p(out["ssep"]),
                                         G.p(sse), None))
    return G.host(grad), G.host(sse)[0], G.host(out["dg0"]), G.host(out["dg1"])


def _noalloc(cs, blocks):
    lib = G.lib()
    out = _outputs(cs, 1)
    a = cs["a"]
    a.dg0, a.dg1, a.dw_partials, a.sse_partials, a.blocks = G.p(out["dg0"]), G.p(out["dg1"]), None, None, blocks
    grad = G.empty((cs["dec"].size,), np.float32)
    sse = G.empty((1,), np.float64)
    G.ok(lib.ntc_fused_train_step_noalloc(G.byref(a), G.p(grad), G.p(sse), None), "ntc_fused_train_step_noalloc")
    return G.host(grad), G.host(sse)[0], G.host(out["dg0"]), G.host(out["dg1"])


@pytest.mark.parametrize("first,second", [(("c3_bilinear_h128", 0, 8), ("c3_bilinear_h128", 2, 3)),
                                          (("ref_bpp_2_25", 0, 4), ("c1_bilinear_relu", 4, 2))])
def test_noalloc_step_matches_workspace_step(oracle, first, second):
    sizes = []
    for kind, lod, n_tiles in (first, second):
        cs = _step_case(oracle, kind, lod, n_tiles)
        blocks = _blocks(cs)
        sizes.append(cs["B"])
        g_ref, s_ref, d0_ref, d1_ref = _reference(cs, blocks)
        g, s, d0, d1 = _noalloc(cs, blocks)
        assert np.array_equal(g, g_ref), kind
        assert s == s_ref
        assert rel(d0, d0_ref) < 1e-6 and rel(d1, d1_ref) < 1e-6
        G.release()
    assert sizes[1] < sizes[0]


def test_noalloc_rejects_workspace_arguments_and_oversize(oracle):
    lib = G.lib()
    cs = _step_case(oracle, "c3_bilinear_h128", 2, 3)
    blocks = _blocks(cs)
    out = _outputs(cs, blocks)
    grad = G.empty((cs["dec"].size,), np.float32)
    a = cs["a"]
    a.dg0, a.dg1, a.blocks = G.p(out["dg0"]), G.p(out["dg1"]), blocks
    a.dw_partials, a.sse_partials = G.p(out["part"]), G.p(out["ssep"])
    assert lib.ntc_fused_train_step_noalloc(G.byref(a), G.p(grad), None, None) == L.NTC_E_INVALID
    a.dw_partials, a.sse_partials = None, None
    assert lib.ntc_fused_train_step_noalloc_max() == 256 * 32768
    a.blocks = 257
    assert lib.ntc_fused_train_step_noalloc(G.byref(a), G.p(grad), None, None) == L.NTC_E_WORKSPACE
