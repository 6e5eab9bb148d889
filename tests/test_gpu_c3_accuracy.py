"""The 4096^2 configuration's accuracy claim, run end to end: the fused BF16
path (MFMA, FP32 master state) and the plain FP32 path each train the 4096^2
9-channel set for the whole 8000-step schedule (the bench workload: g0 16 ch
at 2048^2 4-bit, g1 16 ch at 1024^2 4-bit, 40 -> 128 -> 128 -> 9, 8 tiles of
256^2 per step), each in a child process of its own, and their final PSNRs
(after the lock, quantised latents, as exported) agree within 0.05 dB."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _train(path):
    cmd = [sys.executable, os.path.join(ROOT, "bench.py"), "--time-to-psnr", "8000", "--ttp-path", path]
    print(f"[c3 accuracy] training the {path} path", flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=600)
    assert r.returncode == 0, f"{path} run failed ({r.returncode})"
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.timeout(900)
def test_c3_4096_fused_bf16_reaches_fp32_psnr():
    plain = _train("plain")
    fused = _train("fused")
    d = fused["final_psnr"] - plain["final_psnr"]
    print(f"plain FP32 {plain['final_psnr']:.4f} dB in {plain['train_s']:.2f} s; fused BF16 "
          f"{fused['final_psnr']:.4f} dB in {fused['train_s']:.2f} s; difference {d:+.4f} dB")
    assert abs(d) < 0.05, d
    assert plain["final_psnr"] > 30.0
