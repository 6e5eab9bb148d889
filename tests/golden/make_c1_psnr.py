"""Writes c1_1000_steps.json: the CPU oracle's rendition of the reference step
loop (oracle/, FP32, libopenntc.cpp:471-879) trained on the 512^2 9-channel
synthetic set (seed 20260) for the whole 1000-step schedule, with the
per-channel SSE and PSNR of its final evaluation and of the checkpoints at
250 and 500 steps. Layout: g0 8 ch @ 256^2 4-bit, g1 16 ch @ 128^2 4-bit,
bilinear features + 8 triangle waves, 32 -> 64 -> 64 -> 9 ReLU; one 256^2
tile (65 536 texels) per step; lr 1e-2 (grids) / 1e-3 (decoder), latents
U(-0.5, 0.5), schedule seed 123.

Usage: python tests/golden/make_c1_psnr.py   (several minutes of CPU)"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
import openntc_amd as ntc  # noqa: E402  (the seeded synthetic set only; no GPU)

DIM, SEED_SET, STEPS, TILES = 512, 20260, 1000, 1
LAYOUT = (4, 4, 8, 16, 2, 64)
LR = (1e-2, 1e-3)
INIT = (-0.5, 0.5)


def main():
    tex = ntc.synthetic_pbr(DIM, SEED_SET)
    L = O.layout_custom(*LAYOUT, 1, 1, DIM)
    mt = O.MT(123)
    P = O.init_params(mt, L, *INIT)
    sched = O.schedule(mt, L, STEPS, TILES, *LR)
    pk = O.build_package([(t[0], int(t[1]), t[2]) for t in tex], L)
    tr = O.Trainer(L, P, pk, seed=123, max_batch=8 * 65536)
    out = {"config": {"dim": DIM, "seed_set": SEED_SET, "steps": STEPS, "tiles": TILES, "layout": LAYOUT,
                      "feature_layout": "bilinear", "activation": "relu", "lr": LR, "init": INIT,
                      "mip_count": L.mip_count},
           "checkpoints": []}
    t0 = time.time()
    for t in range(STEPS):
        tr.step(sched[t], t)
        if t + 1 in (250, 500, STEPS):
            mse, cs = tr.eval(channel_sse=True)
            out["checkpoints"].append({"steps": t + 1, "mse": mse, "psnr": float(-10 * np.log10(mse)),
                                       "channel_sse": [float(x) for x in cs]})
            print(f"{t + 1} steps, {time.time() - t0:.0f} s: psnr {out['checkpoints'][-1]['psnr']:.4f}", flush=True)
    with open(os.path.join(HERE, "c1_1000_steps.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
