"""The public C++ API driven by a C++ program linked against the library
(tests/cpp/api_flow.cpp): the Tensor wrappers, the README flow from a
1024^2 manifest of PNG and TGA files through training with the reference's
Low quality preset (30000 steps), evaluation, calibration, export and the
ONTC container, the housekeeping calls (GetTextureData, GetMipDim,
UnloadManifest, the Train(n) clamp and TrainComplete once no steps remain),
and the AllocationFailure / CudaFailure paths reached through the library's
fault-injection switch."""
import os
import subprocess

import numpy as np
import pytest

import openntc_amd as ntc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "api_flow")


@pytest.mark.timeout(300)
def test_cpp_api_flow(tmp_path):
    from tests.test_manifest import _manifest, write_png, write_tga
    assert os.path.exists(EXE), "tests/cpp/api_flow not built (python -m openntc_amd.build)"
    dim = 1024
    tex = ntc.synthetic_pbr(dim, 4242)
    albedo, normal, orm = tex[0][0], tex[1][0], tex[2][0]
    write_png(tmp_path / "albedo.png", albedo[..., :3], 2, 8)
    write_png(tmp_path / "normal.png", normal[..., :3], 2, 8)
    write_tga(tmp_path / "orm.tga", orm, rle=True)
    m = _manifest(tmp_path, [
        {"fileName": "orm.tga", "semantics": {"AO": "R", "Roughness": "G", "Metallic": "B"}},
        {"fileName": "albedo.png", "semantics": {"Albedo": "RGB"}},
        {"fileName": "normal.png", "semantics": {"Normal": "RGB"}}], width=dim, height=dim)
    raw = tmp_path / "albedo.raw"
    raw.write_bytes(np.ascontiguousarray(albedo[..., :3]).tobytes())
    print("[cpp api] running api_flow", flush=True)
    r = subprocess.run([EXE, str(m), str(tmp_path), str(raw)], capture_output=True, text=True, timeout=280)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stderr
    assert "checks passed" in r.stdout
