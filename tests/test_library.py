"""CPU checks of the native library: it loads, exports every entry point the
public C header declares, rejects bad arguments without a GPU, and the
Context reports a GPU failure cleanly on a host without one."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import openntc_amd as ntc
from openntc_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ntc_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = ntc.load()
    names = _declared("ntc_hip.h")
    assert len(names) > 60
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and every declared symbol has a ctypes signature
    assert not [n for n in names if n not in L.SIGNATURES]


def test_public_cpp_headers_compile_standalone(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text("#include <libopenntc/libopenntc.h>\n#include <ntc_hip.h>\n"
                   "static_assert(sizeof(openntc::NTCConstants) == 240);\nint main(){return 0;}\n")
    r = os.system(f"g++ -std=c++20 -fsyntax-only -I{ROOT}/include {src}")
    assert r == 0


def test_version_and_argument_checks_without_gpu():
    lib = ntc.load()
    assert b"gfx950" in lib.ntc_version()
    fd = L.FeatureDesc(8, 4, 40, 4, 0, 8, 3, 1, 92)
    t = L.Tiles(0, 1, 8, 8, 0.0)
    assert lib.ntc_draw_features_f32(C.byref(fd), C.byref(t), None, 0, 0, 0, None) == L.NTC_E_INVALID
    assert lib.ntc_quantize_pack_f32(33, 4, 1, 1, None) == L.NTC_E_INVALID
    assert lib.ntc_gemm_f32(2, 4, 4, 4096, 1, 1, 1, None, 0, None) == L.NTC_E_WORKSPACE
    assert lib.ntc_gemm_workspace_bytes(2, 9, 64, 5000, 4) == 5 * 9 * 64 * 4
    assert lib.ntc_synthetic_pbr(100, 1, None, None, None) == L.NTC_E_INVALID


def test_fused_layout_support_table():
    lib = ntc.load()
    # reference profiles (corner features, HardGELU) and the configuration layouts
    for c0, c1, mode, pec, lod, F, H, act in [(16, 12, 0, 3, 1, 92, 64, 0), (8, 12, 0, 3, 1, 60, 64, 0),
                                              (12, 20, 0, 3, 1, 84, 64, 0), (12, 10, 0, 3, 1, 72, 64, 0),
                                              (8, 16, 1, 2, 0, 32, 64, 1), (12, 16, 1, 2, 0, 36, 64, 1),
                                              (16, 16, 1, 2, 0, 40, 128, 1)]:
        fd = L.FeatureDesc(64, 32, c0, c1, mode, 8 if pec == 3 else 4, pec, lod, F)
        assert lib.ntc_fused_supported(C.byref(fd), H, act) == 1, (c0, c1, H)
    fd = L.FeatureDesc(64, 32, 16, 16, 1, 4, 2, 0, 40)
    assert lib.ntc_fused_supported(C.byref(fd), 96, 1) == 0


def test_synthetic_set_is_seeded_and_in_range():
    a = ntc.synthetic_pbr(64, 7)
    b = ntc.synthetic_pbr(64, 7)
    c = ntc.synthetic_pbr(64, 8)
    assert all(np.array_equal(x[0], y[0]) for x, y in zip(a, b))
    assert not np.array_equal(a[0][0], c[0][0])
    albedo, normal = a[0][0], a[1][0]
    assert albedo[..., 3].min() == 255 and albedo[..., :3].std() > 5
    n = normal[..., :3].astype(np.float64) / 255 * 2 - 1
    assert np.allclose(np.linalg.norm(n, axis=-1), 1.0, atol=0.02)


def test_context_without_gpu_reports_failure():
    lib = ntc.load()
    n = C.c_int32()
    if lib.ntc_device_check(C.byref(n)) == 0:
        pytest.skip("a GPU is present")
    ctx = ntc.Context()
    assert ctx.Init(ntc.Profile.Bpp_0_5) == ntc.Result.CudaFailure
    prog = ctx.Train(1)
    assert prog.phase == ntc.TrainPhase.TrainError and prog.result == ntc.Result.InvalidState
    assert ctx.LoadManifest("/nonexistent.json") in (ntc.Result.FileNotFound, ntc.Result.InvalidState)


def test_fault_injection_switch_arguments():
    """The fault-injection switch takes the three sites and n >= 0 (n = 0
    disarms) and refuses anything else; arming touches no device."""
    lib = ntc.load()
    for site in (1, 2, 3):
        assert lib.ntc_debug_inject_fault(site, 3) == 0
        assert lib.ntc_debug_inject_fault(site, 0) == 0
    assert lib.ntc_debug_inject_fault(0, 1) == L.NTC_E_INVALID
    assert lib.ntc_debug_inject_fault(4, 1) == L.NTC_E_INVALID
    assert lib.ntc_debug_inject_fault(2, -1) == L.NTC_E_INVALID


def test_cpp_api_caller_links():
    """The C++ caller of the public API (tests/cpp/api_flow.cpp) is built
    against the library and its dynamic references resolve; without its
    arguments it prints its usage and touches no device."""
    import subprocess
    exe = os.path.join(ROOT, "tests", "cpp", "api_flow")
    if not os.path.exists(exe):
        pytest.skip("tests/cpp/api_flow not built (python -m openntc_amd.build)")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage: api_flow" in r.stderr
