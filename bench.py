"""Training throughput of the openntc hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): a seeded
synthetic 9-channel 4096x4096 PBR set with Lanczos-3 mips down to 1x1;
latents g0 16 ch at 2048^2 (4-bit) and g1 16 ch at 1024^2 (4-bit); decoder
40 -> 128 -> 128 -> 9 (bilinear features + 8 triangle waves, ReLU); 8 wrapping
256^2 tiles = 524 288 texels per step; 8000-step schedule (cosine learning
rates 1e-2 / 1e-3), of which `--warmup` untimed and `--steps` timed steps run.
A step is the whole training step: gather, forward, loss, backward, latent
scatter, decoder Adam and the dense Adam + clamp over the sampled pyramid
level. Inputs are resident in HBM before the timed region.

Prints one JSON line. `roofline` covers the dominant kernel, timed live with
HIP events on the context's stream; `kernels` lists the others.
`cpu_baseline` is the CPU oracle's rendition of the same step (all host
threads) on a bounded sample of steps.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
FP32_PEAK_TFLOPS = 157.3

LAYOUT = dict(g0_bits=4, g1_bits=4, g0_channels=16, g1_channels=16, g0_scale=2, hidden=128)
DIM = 4096
TOTAL_STEPS = 8000
TILES = 8
TIMING_STRIDE = 8  # timed region: events around the MLP launch of every 8th step
SEED_SET = 20262


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_context(path, dim=DIM, seed=SEED_SET, layout=LAYOUT, steps=TOTAL_STEPS, tiles=TILES):
    import openntc_amd as ntc

    tex = ntc.synthetic_pbr(dim, seed)
    ctx = ntc.Context()
    lay = ntc.LayoutInfo(layout["g0_bits"], layout["g1_bits"], layout["g0_channels"], layout["g1_channels"],
                         layout["g0_scale"], layout["hidden"], ntc.FeatureLayout.Bilinear, ntc.Activation.ReLU)
    r = ctx.InitEx(layout=lay, min_dimension=512, mip_min_dim=1, seed=123)
    if r != ntc.Result.Success:
        raise SystemExit(f"InitEx failed: {r.name}")
    r = ctx.LoadTextures(tex, dim)
    if r != ntc.Result.Success:
        raise SystemExit(f"LoadTextures failed: {r.name}")
    r = ctx.BeginTrainingEx(steps, tiles, 1e-2, 1e-3, -0.5, 0.5, path)
    if r != ntc.Result.Success:
        raise SystemExit(f"BeginTrainingEx failed: {r.name} ({ctx.GetLastError().name})")
    return ctx, tex


# Evaluation workload: the 2048^2 configuration's layout (g0 12 ch at half
# resolution, 4-bit; g1 16 ch at quarter resolution, 2-bit; 36 -> 64 -> 64 -> 9)
# trained on a 4096^2 set (4 tiles per step, 4000 steps), then decoded over
# every mip down to 1x1 (22.37 M texels).
C2_LAYOUT = dict(g0_bits=4, g1_bits=2, g0_channels=12, g1_channels=16, g0_scale=2, hidden=64)
EVAL_WORKLOAD = dict(dim=4096, seed=20264, steps=4000, tiles=4, layout=C2_LAYOUT)
# The 2048^2 configuration itself: the same layout on a 2048^2 set, 4 tiles
# (262 144 texels) per step, 4000 steps.
C2_WORKLOAD = dict(dim=2048, seed=20261, steps=4000, tiles=4, layout=C2_LAYOUT)
# target channels after the semantic sort: albedo RGB, metallic, normal RGB, AO, roughness
CHANNEL_GROUPS = {"albedo": (0, 1, 2), "normal": (4, 5, 6), "rough_metal_ao": (3, 7, 8)}


# Batched configuration: 16 independent 9-channel sets with sides drawn
# (seeded) from {256, 512, 1024, 2048}; per set g0 8 ch at half and g1 16 ch at
# quarter resolution (4-bit), decoder 32 -> 64 -> 64 -> 9 ReLU, one 256^2 tile
# (65 536 texels) per set and step, 2000 steps, trained together.
SETS_WORKLOAD = dict(n_sets=16, sides=(256, 512, 1024, 2048), seed=20263, steps=2000,
                     layout=dict(g0_bits=4, g1_bits=4, g0_channels=8, g1_channels=16, g0_scale=2, hidden=64))


def make_sets(n_sets=16, sides=(256, 512, 1024, 2048), seed=20263, steps=2000, layout=None, dims=None):
    """The batched configuration's contexts, each ready to train on the fused
    path (own latents, decoder, targets and schedule); sides drawn from
    `sides` unless `dims` lists them."""
    import openntc_amd as ntc

    lw = layout or SETS_WORKLOAD["layout"]
    if dims is None:
        rng = np.random.default_rng(seed)
        dims = [int(d) for d in rng.choice(sides, n_sets)]
    sides = tuple(dims) + tuple(sides)
    ctxs = []
    for i, d in enumerate(dims):
        tex = ntc.synthetic_pbr(d, seed * 100 + i)
        ctx = ntc.Context()
        lay = ntc.LayoutInfo(lw["g0_bits"], lw["g1_bits"], lw["g0_channels"], lw["g1_channels"], lw["g0_scale"],
                             lw["hidden"], ntc.FeatureLayout.Bilinear, ntc.Activation.ReLU)
        for r in (ctx.InitEx(layout=lay, min_dimension=min(sides), mip_min_dim=1, seed=123 + i),
                  ctx.LoadTextures(tex, d),
                  ctx.BeginTrainingEx(steps, 1, 1e-2, 1e-3, -0.5, 0.5, ntc.TrainPath.Fused)):
            if r != ntc.Result.Success:
                raise SystemExit(f"set {i} ({d}^2) setup failed: {r.name}")
        ctxs.append(ctx)
    return ctxs, dims


def group_psnr(channel_sse, count):
    """PSNR over each channel group from per-channel FP64 squared errors."""
    return {g: float(-10.0 * np.log10(sum(channel_sse[c] for c in ch) / (count * len(ch))))
            for g, ch in CHANNEL_GROUPS.items()}


def step_texels(ctx, first, count):
    tot = 0
    infos = []
    for s in range(first, first + count):
        si = ctx.step_info(s)
        tot += si.n_tiles * si.tile_dim * si.tile_dim
        infos.append(si)
    return tot, infos


def _axis_corners(pu, M, D):
    """Unwrapped bilinear corner cells of unwrapped texel coordinate pu, in
    float32 arithmetic as the kernels compute them."""
    wrapped = pu >= M
    x = pu - M if wrapped else pu
    u = (np.float32(x) + np.float32(0.5)) / np.float32(M)
    g = u * np.float32(D) - np.float32(0.5)
    x0 = int(np.floor(g))
    add = D if wrapped else 0
    return min(max(x0, 0), D - 1) + add, min(max(x0 + 1, 0), D - 1) + add


def footprint_cells(si, M, D):
    """Cells of a D^2 grid inside the union of the step's tile footprints."""
    T = si.tile_dim
    if si.n_tiles == 1:
        n = 1
        for o in (si.origins[0] % M, si.origins[1] % M):
            a, _ = _axis_corners(o, M, D)
            _, d = _axis_corners(o + T - 1, M, D)
            n *= min(d - a + 1, D)
        return n
    mask = np.zeros((D, D), bool)
    for b in range(si.n_tiles):
        ox, oy = si.origins[2 * b] % M, si.origins[2 * b + 1] % M
        rx, ry = [], []
        for o, r in ((ox, rx), (oy, ry)):
            a, _ = _axis_corners(o, M, D)
            _, d = _axis_corners(o + T - 1, M, D)
            r.extend([a, min(d - a + 1, D)])
        xs = (rx[0] + np.arange(rx[1])) % D
        ys = (ry[0] + np.arange(ry[1])) % D
        mask[np.ix_(ys, xs)] = True
    return int(mask.sum())


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(ctx, first_step, n_steps, warm=5):
    """The CPU oracle (C, OpenMP, FP32) running the same schedule steps from
    the same initial parameters: `warm` untimed steps, then `n_steps` timed."""
    from oracle import oracle as O

    lay = ctx.layout()
    L = O.layout_custom(LAYOUT["g0_bits"], LAYOUT["g1_bits"], LAYOUT["g0_channels"], LAYOUT["g1_channels"],
                        LAYOUT["g0_scale"], LAYOUT["hidden"], 1, 1, DIM, 9, 1)
    mt = O.MT(123)
    P = O.init_params(mt, L, -0.5, 0.5)
    steps = O.schedule(mt, L, TOTAL_STEPS, TILES, 1e-2, 1e-3)
    pk = [ctx.read_package(m) for m in range(lay.mip_count)]
    tr = O.Trainer(L, P, pk, seed=123, max_batch=TILES * 65536)
    for s in range(max(0, first_step - warm), first_step):  # warm caches / page in
        tr.step(steps[s], s)
    texels = 0
    t0 = time.perf_counter()
    for s in range(first_step, first_step + n_steps):
        tr.step(steps[s], s)
        texels += steps[s].n_tiles * steps[s].tile_dim ** 2
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": texels / dt, "unit": "texel-samples/s", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"{n_steps} schedule steps ({texels} texels) of the same workload after {warm} warm-up "
                      f"steps, CPU oracle (C, OpenMP, FP32, {cores} threads), {dt:.1f} s"}


def lod0_micro(ntc, ctx, reps=50, seed=7):
    """The training step's kernels alone at a fixed lod 0 (8 random 256^2
    tiles of the 4096^2 mip, the context's current latents and decoder, noise
    on): wall time per step over `reps` back-to-back steps, of the one-launch
    kernel and of the three-launch form, whose MLP launch is also timed alone
    with an event pair right around it in every step."""
    from openntc_amd import _lib as L
    lib = ntc.load()
    lay = ctx.layout()
    H, F, O = lay.hidden_dim, lay.feature_dim, lay.out_dim
    fd = lay.feat  # level 0, the level lod 0 samples
    g0, g1 = ctx.read_grid(0, 0), ctx.read_grid(1, 0)
    dec = ctx.read_decoder()
    pkg = ctx.read_package(0)
    rng = np.random.default_rng(seed)
    origins = rng.integers(0, DIM, 2 * TILES).astype(np.int32)
    bufs = []

    def dev(a=None, nbytes=None):
        p = ctypes.c_void_p()
        n = a.nbytes if a is not None else nbytes
        ntc.check(lib.ntc_dev_alloc(n, ctypes.byref(p)), "ntc_dev_alloc")
        bufs.append(p)
        if a is not None:
            a = np.ascontiguousarray(a)
            ntc.check(lib.ntc_dev_upload(p, a.ctypes.data, a.nbytes), "upload")
        return p.value
    hip = ctypes.CDLL("libamdhip64.so")
    events = []
    flop = 6.0 * (F * H + H * H + H * O) * TILES * 256 * 256
    out = {"lod": 0, "texels": TILES * 256 * 256, "sample": f"{reps} back-to-back steps, wall clock"}
    try:
        d_dec = dev(dec)
        tab = dev(nbytes=2 * lib.ntc_fused_table_elems(ctypes.byref(fd), H))
        ntc.check(lib.ntc_pack_decoder_bf16(ctypes.byref(fd), H, O, d_dec, tab, None), "pack")
        maxb = ctypes.c_int32()
        lib.ntc_fused_max_blocks(ctypes.byref(maxb))
        B = TILES * 256 * 256
        a = L.FusedArgs()
        a.feat = fd
        a.tiles = L.Tiles(dev(origins), TILES, 256, DIM, 0.0)
        a.noise = L.Noise(123, 1, 0, 1, lay.g0_delta, lay.g1_delta)
        a.hidden_dim, a.out_dim, a.activation = H, O, lay.activation
        a.g0, a.g1, a.weights, a.targets = dev(g0), dev(g1), tab, dev(pkg)
        a.loss_scale = -2.0 / (O * B)
        a.scatter = 1
        a.dg0, a.dg1 = dev(nbytes=g0.nbytes), dev(nbytes=g1.nbytes)
        a.dw_partials, a.sse_partials = dev(nbytes=4 * maxb.value * dec.size), dev(nbytes=8 * maxb.value)

        def timed(call):
            for _ in range(3):
                ntc.check(call(None), "warm-up step")
            lib.ntc_dev_synchronize()
            t0 = time.perf_counter()
            for r in range(reps):
                call(r)
            ntc.check(lib.ntc_dev_synchronize(), "synchronize")
            return (time.perf_counter() - t0) * 1e3 / reps

        a.blocks = lib.ntc_fused_blocks_for(ctypes.byref(fd), H, lay.activation, B, maxb.value)
        ms = timed(lambda r: lib.ntc_fused_train_step(ctypes.byref(a), None))
        out["one_launch"] = {"kernel": "k_fused_train", "ms": ms, "texels_per_s": B / (ms * 1e-3),
                             "tflops": flop / (ms * 1e-3) / 1e12,
                             "mfma_frac": flop / (ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS}
        if lib.ntc_split_supported(ctypes.byref(fd), H, lay.activation):
            a.blocks = lib.ntc_split_blocks_for(ctypes.byref(fd), H, lay.activation, B, maxb.value)
            nbytes = lib.ntc_split_workspace_bytes(ctypes.byref(fd), H, lay.activation, B)
            ws = dev(nbytes=nbytes)
            for _ in range(2 * reps):
                e = ctypes.c_void_p()
                if hip.hipEventCreate(ctypes.byref(e)) != 0:
                    raise SystemExit("hipEventCreate failed")
                events.append(e)

            def step(r):
                st = lib.ntc_split_gather(ctypes.byref(a), ws, nbytes, None)
                if r is not None:
                    hip.hipEventRecord(events[2 * r], None)
                st = st or lib.ntc_split_mlp(ctypes.byref(a), ws, nbytes, None)
                if r is not None:
                    hip.hipEventRecord(events[2 * r + 1], None)
                return st or lib.ntc_split_scatter(ctypes.byref(a), ws, nbytes, None)
            ms = timed(step)
            mlp = []
            for r in range(reps):
                t = ctypes.c_float()
                if hip.hipEventElapsedTime(ctypes.byref(t), events[2 * r], events[2 * r + 1]) == 0:
                    mlp.append(t.value)
            mlp_ms = float(np.mean(mlp))
            out["three_launches"] = {"kernels": "k_split_gather + k_mlp_train + k_split_scatter", "ms": ms,
                                     "texels_per_s": B / (ms * 1e-3), "tflops": flop / (ms * 1e-3) / 1e12,
                                     "mfma_frac": flop / (ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS,
                                     "mlp_ms": mlp_ms, "mlp_tflops": flop / (mlp_ms * 1e-3) / 1e12,
                                     "mlp_mfma_frac": flop / (mlp_ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS}
    finally:
        for e in events:
            hip.hipEventDestroy(e)
        for p in bufs:
            lib.ntc_dev_free(p)
    return out


def pmc_traffic(kernel):
    """HBM bytes per launch of `kernel` as measured with rocprofv3 FETCH_SIZE /
    WRITE_SIZE (copy-calibrated) and recorded in the committed profile summary;
    None when the summary holds no row for it."""
    path = os.path.join(ROOT, "profiles", "train_step_4096_c3_hbm_traffic.md")
    try:
        lines = open(path).read().splitlines()
    except OSError:
        return None
    units = {"KB": 1e3, "MB": 1e6, "GB": 1e9}
    for ln in lines:
        cols = [c.strip() for c in ln.strip().strip("|").split("|")]
        if len(cols) >= 6 and cols[0].split(" ")[0] == kernel:
            m = re.match(r"([0-9.]+) (KB|MB|GB)$", cols[4])
            if m:
                return float(m.group(1)) * units[m.group(2)]
    return None


def time_to_psnr(ntc, chunk, workload=None, only=None, report=True):
    """Seconds of training for the fused BF16 path to reach the plain FP32
    path's final PSNR - 0.05 dB on the same workload and seed (default: the
    4096^2 bench workload). Both paths run the whole schedule; the fused one
    is evaluated every
    `chunk` steps (evaluation time excluded). Before the lock step (95 %)
    the evaluated latents are unquantised; the final figures of both paths
    are after the lock (quantised latents, as exported).

    Each path trains in a child process of its own (only = that path), as a
    training run normally owns its process: in one process, the fused path
    trained after a full plain-path schedule measured 1.09 instead of 0.79 ms
    per step (same clocks and temperatures), which a fresh process avoids."""
    w = workload or dict(dim=DIM, seed=SEED_SET, steps=TOTAL_STEPS, tiles=TILES, layout=LAYOUT)
    total = w["steps"]
    if only is None:
        import subprocess
        res = {}
        for path in ("plain", "fused"):
            cmd = [sys.executable, os.path.abspath(__file__), "--time-to-psnr", str(chunk), "--ttp-path", path,
                   "--workload", "2048" if workload is C2_WORKLOAD else "4096"]
            r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True)
            if r.returncode != 0:
                raise SystemExit(f"time-to-psnr {path} run failed ({r.returncode})")
            res[path] = json.loads(r.stdout.strip().splitlines()[-1])
        return time_to_psnr_report(res, w, total, report)
    lib = ntc.load()
    res = {}
    for path in (only,):
        log(f"time-to-psnr: {path} path, {total} steps")
        ctx, _ = make_context(ntc.TrainPath.Fused if path == "fused" else ntc.TrainPath.Plain, **w)
        lib.ntc_dev_synchronize()
        t_train, trace, done = 0.0, [], 0
        step = chunk if path == "fused" else 1000
        while done < total:
            n = min(step, total - done)
            t0 = time.perf_counter()
            prog = ctx.Train(n)
            lib.ntc_ctx_flush_grids(ctx._h)  # deferred grid steps counted as training
            lib.ntc_dev_synchronize()
            t_train += time.perf_counter() - t0
            if prog.result != ntc.Result.Success:
                raise SystemExit(f"{path} training failed: {prog}")
            done += n
            if path == "fused" or done == total:
                psnr = ctx.Eval().psnr
                trace.append({"steps": done, "train_s": round(t_train, 4), "psnr": round(psnr, 4)})
                log(f"  {path} {done} steps {t_train:.2f} s psnr {psnr:.3f}")
            else:
                log(f"  {path} {done} steps {t_train:.2f} s")
        res[path] = {"train_s": t_train, "final_psnr": trace[-1]["psnr"], "trace": trace}
    print(json.dumps(res[only]), flush=True)


def time_to_psnr_report(res, w, total, report=True):
    target = res["plain"]["final_psnr"] - 0.05
    hit = next((t for t in res["fused"]["trace"] if t["psnr"] >= target), None)
    lw = w["layout"]
    out = {"metric": f"seconds to FP32 PSNR (BF16 fused vs FP32 plain), {w['dim']}^2 9-ch, {total}-step schedule",
           "config": {"workload": f"{w['dim']}^2 9-ch synthetic set (seed {w['seed']}); g0 {lw['g0_channels']}ch@"
                                  f"{w['dim'] // lw['g0_scale']}^2 {lw['g0_bits']}-bit, g1 {lw['g1_channels']}ch@"
                                  f"{w['dim'] // lw['g0_scale'] // 2}^2 {lw['g1_bits']}-bit; hidden {lw['hidden']}; "
                                  f"{w['tiles']} x 256^2 tiles/step"},
           "plain_fp32": {"train_s": res["plain"]["train_s"], "final_psnr": res["plain"]["final_psnr"]},
           "fused_bf16": {"train_s": res["fused"]["train_s"], "final_psnr": res["fused"]["final_psnr"],
                          "trace": res["fused"]["trace"]},
           "target_psnr": target,
           "final_within_0.05dB": abs(res["fused"]["final_psnr"] - res["plain"]["final_psnr"]) <= 0.05,
           "seconds_to_target": hit["train_s"] if hit else None,
           "steps_to_target": hit["steps"] if hit else None,
           "speedup_full_schedule": res["plain"]["train_s"] / res["fused"]["train_s"]}
    if report:
        print(json.dumps(out), flush=True)
    return out


def eval_precisions(ntc, reps):
    """Full-chain evaluation of the trained evaluation workload at FP32, BF16,
    FP8 E4M3, INT8 (the exported decoder) and BF16 on the matrix cores:
    per-precision wall time of one
    Context::EvalEx walk (gather + targets + decode + FP64 SSE over all mips),
    decode throughput and PSNR overall and per channel group."""
    lib = ntc.load()
    w = EVAL_WORKLOAD
    log(f"eval workload: {w['dim']}^2 set, training {w['steps']} fused steps")
    ctx, _ = make_context(ntc.TrainPath.Fused, **w)
    lib.ntc_dev_synchronize()
    t0 = time.perf_counter()
    prog = ctx.TrainUntilComplete()
    lib.ntc_dev_synchronize()
    t_train = time.perf_counter() - t0
    if prog.phase != ntc.TrainPhase.TrainComplete:
        raise SystemExit(f"training failed: {prog}")
    log(f"trained in {t_train:.2f} s")
    g0b, g1b, blob, cal = ctx.GetCompressedData()
    lay = ctx.layout()
    F, H, O = lay.feature_dim, lay.hidden_dim, lay.out_dim
    texels = sum(lay.mip_dim[m] ** 2 for m in range(lay.mip_count))
    flop_per_texel = 2.0 * (F * H + H * H + H * O)
    res = {}
    for prec in (ntc.EvalPrecision.FP32, ntc.EvalPrecision.BF16, ntc.EvalPrecision.FP8, ntc.EvalPrecision.INT8,
                 ntc.EvalPrecision.BF16_MFMA, ntc.EvalPrecision.FP8_MFMA):
        ctx.EvalEx(prec)  # untimed: builds the precision's tables (FP8 runs its calibration walk)
        lib.ntc_dev_synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            ev = ctx.EvalEx(prec)
        lib.ntc_dev_synchronize()
        dt = (time.perf_counter() - t0) / reps
        res[prec.name] = {"eval_ms": dt * 1e3, "texels_per_s": texels / dt, "psnr": ev.psnr,
                          "psnr_groups": group_psnr(ev.channel_sse, texels)}
        log(f"  {prec.name}: {dt * 1e3:.1f} ms per walk, psnr {ev.psnr:.4f}")
    out = {"metric": "evaluation decode over all mips, 4096^2 9-ch (FP32 / BF16 / FP8 E4M3 / INT8 / BF16 MFMA)",
           "unit": "texels/s", "texels_per_walk": texels, "flop_per_texel": flop_per_texel, "reps": reps,
           "train_s": t_train,
           "config": {"workload": "4096^2 9-ch synthetic set; g0 12ch@2048^2 4-bit, g1 16ch@1024^2 2-bit; "
                                  "MLP 36->64->64->9 ReLU; 4 x 256^2 tiles/step, 4000 fused steps; "
                                  "decode of every mip down to 1x1"},
           "export_bytes": {"g0_packed": int(g0b.nbytes), "g1_packed": int(g1b.nbytes), "decoder_blob": int(blob.nbytes)},
           "precisions": res}
    print(json.dumps(out), flush=True)


def sets_bench(ntc, warmup, steps):
    """The batched configuration: 16 sets trained together, one launch per
    stage and step replayed from a captured graph, launch shape tuned at
    creation. Times `steps` schedule steps after `warmup` (the lock step and
    the completion are outside the timed span)."""
    w = SETS_WORKLOAD
    log(f"sets: building {w['n_sets']} sets")
    ctxs, dims = make_sets(n_sets=w["n_sets"], sides=w["sides"], seed=w["seed"], steps=w["steps"])
    lib = ntc.load()
    t0 = time.perf_counter()
    b = ntc.Batch(ctxs)
    t_tune = time.perf_counter() - t0
    st = b.stats()
    log(f"sets: launch shape pass {st.pass_texels} grid {st.fused_grid} (tuned in {t_tune:.2f} s)")
    lock = ctxs[0].layout().lock_step
    steps = min(steps, lock - warmup)
    if b.train(warmup) != ntc.Result.Success:
        raise SystemExit("sets warmup failed")
    lib.ntc_dev_synchronize()
    t0 = time.perf_counter()
    if b.train(steps) != ntc.Result.Success:
        raise SystemExit("sets timed steps failed")
    lib.ntc_dev_synchronize()
    dt = time.perf_counter() - t0
    texels = 0
    flop = 0.0
    for c, d in zip(ctxs, dims):
        lay = c.layout()
        F, H, O = lay.feature_dim, lay.hidden_dim, lay.out_dim
        for t in range(warmup, warmup + steps):
            si = c.step_info(t)
            B = si.n_tiles * si.tile_dim ** 2
            texels += B
            flop += 6.0 * (F * H + H * H + H * O) * B  # forward 2, backward 4 per multiply-add
    tflops = flop / dt / 1e12
    out = {"metric": "train texel-samples/s, 16 uneven sets in one launch per stage (graph replay)",
           "value": texels / dt, "unit": "texel-samples/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
           "ms_per_step": dt * 1e3 / steps, "higher_is_better": True, "dtype": "bf16",
           "data": "synthetic (seeded 9-channel PBR sets, random-init latents and decoders)",
           "config": {"workload": "16 sets, sides drawn from {256,512,1024,2048} (seed 20263): " + ",".join(map(str, dims))
                                  + "; per set g0 8ch@S/2 4-bit, g1 16ch@S/4 4-bit, MLP 32->64->64->9 ReLU, "
                                    "1 x 256^2 tile per set and step, 2000-step schedule"},
           "launch": {"pass_texels": st.pass_texels, "fused_grid": st.fused_grid,
                      "grid_update_grid": st.grid_update_grid, "tune_ms": list(st.tune_ms),
                      "tune_pass": list(st.tune_pass), "tune_grid": list(st.tune_grid), "tune_s": t_tune},
           "roofline": {"bound": "mfma", "kernel": "whole step (all stages) against the BF16 MFMA peak",
                        "achieved": tflops, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": tflops / BF16_PEAK_TFLOPS, "flop_per_step": flop / steps}}
    print(json.dumps(out), flush=True)
    b.close()


def copy_probe(ntc, mib):
    """A plain copy of known size: reads and writes mib MiB once each. Its
    FETCH_SIZE / WRITE_SIZE counter readings calibrate the traffic figures."""
    import ctypes

    lib = ntc.load()
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    ntc.check(lib.ntc_dev_alloc(n, ctypes.byref(a)), "alloc")
    ntc.check(lib.ntc_dev_alloc(n, ctypes.byref(b)), "alloc")
    for _ in range(3):
        ntc.check(lib.ntc_dev_copy(b.value, a.value, n), "copy")
    lib.ntc_dev_synchronize()
    t0 = time.perf_counter()
    ntc.check(lib.ntc_dev_copy(b.value, a.value, n), "copy")
    lib.ntc_dev_synchronize()
    dt = time.perf_counter() - t0
    print(json.dumps({"copy_probe_bytes": n, "seconds": dt, "GBps_rw": 2 * n / dt / 1e9}), flush=True)
    lib.ntc_dev_free(a.value)
    lib.ntc_dev_free(b.value)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--path", choices=["fused", "plain"], default="fused")
    ap.add_argument("--cpu-steps", type=int, default=50)
    ap.add_argument("--detail-steps", type=int, default=100,
                    help="steps after the timed region over which the step's other kernels are timed")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-time-to-psnr", action="store_true",
                    help="leave out the seconds-to-FP32-PSNR pair (two full schedules, about a minute)")
    ap.add_argument("--time-to-psnr", type=int, default=0, metavar="CHUNK",
                    help="only measure the seconds for the fused path to reach the plain path's final PSNR, "
                         "evaluating every CHUNK steps (runs both paths over the whole schedule)")
    ap.add_argument("--ttp-path", choices=["fused", "plain"], default=None, help=argparse.SUPPRESS)
    ap.add_argument("--workload", choices=["4096", "2048"], default="4096",
                    help="workload of --time-to-psnr: the 4096^2 bench workload or the 2048^2 configuration")
    ap.add_argument("--eval-precisions", type=int, default=0, metavar="REPS",
                    help="only train the evaluation workload and time REPS full evaluations per precision")
    ap.add_argument("--sets", action="store_true",
                    help="only measure the batched 16-set configuration (--steps after --warmup schedule steps)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (HBM counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every GPU box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import openntc_amd as ntc

    if a.copy_probe:
        copy_probe(ntc, a.copy_probe)
        return
    if a.time_to_psnr:
        time_to_psnr(ntc, a.time_to_psnr, C2_WORKLOAD if a.workload == "2048" else None, a.ttp_path)
        return
    if a.eval_precisions:
        eval_precisions(ntc, a.eval_precisions)
        return
    if a.sets:
        sets_bench(ntc, a.warmup, a.steps)
        return
    ttp = None
    if not a.no_time_to_psnr and a.path == "fused":
        # the second half of the metric: each path in a child process started
        # before this process does any GPU work
        log("time to FP32 PSNR: the plain and the fused path over the whole schedule, one child each")
        ttp = time_to_psnr(ntc, 500, report=False)
    path = ntc.TrainPath.Fused if a.path == "fused" else ntc.TrainPath.Plain
    if a.warmup + a.steps + (a.detail_steps if a.path == "fused" else 0) > TOTAL_STEPS:
        raise SystemExit(f"warmup + steps + detail steps exceed the {TOTAL_STEPS}-step schedule")
    if a.path == "fused" and a.detail_steps < 1:
        raise SystemExit("--detail-steps must be at least 1")
    log(f"setting up {DIM}^2 synthetic set, path={a.path}")
    t_setup = time.perf_counter()
    ctx, _ = make_context(path)
    log(f"setup done in {time.perf_counter() - t_setup:.1f} s")
    lib = ntc.load()

    prog = ctx.Train(a.warmup)
    if prog.result != ntc.Result.Success:
        raise SystemExit(f"warmup failed: {prog}")
    log(f"warmup {a.warmup} steps done")
    # events around the fused kernel (the MLP launch of the three-launch
    # step) only, and only in every TIMING_STRIDE-th step inside the timed
    # region: each event between two of the step's kernels leaves the GPU idle
    # for a few microseconds; the other kernels are timed over a sample after it
    lib.ntc_ctx_set_timing(ctx._h, 2)
    lib.ntc_ctx_set_timing_stride(ctx._h, TIMING_STRIDE)
    lib.ntc_dev_synchronize()
    t0 = time.perf_counter()
    prog = ctx.Train(a.steps)
    # the timed work ends with every latent cell current: the Adam steps the
    # fused path defers (zero-gradient steps outside the tile footprints) are
    # applied inside the timed region
    lib.ntc_ctx_flush_grids(ctx._h)
    lib.ntc_dev_synchronize()
    dt = time.perf_counter() - t0
    if prog.result != ntc.Result.Success:
        raise SystemExit(f"timed steps failed: {prog}")
    log(f"timed {a.steps} steps in {dt * 1e3:.1f} ms")
    texels, infos = step_texels(ctx, a.warmup, a.steps)

    lay = ctx.layout()
    H, F, O = lay.hidden_dim, lay.feature_dim, lay.out_dim
    out = {
        "metric": "train texel-samples/s (fwd+bwd+Adam) @4096² 9-ch; s to FP32 PSNR, BF16 vs FP32",
        "value": texels / dt,
        "unit": "texel-samples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16" if a.path == "fused" else "fp32",
        "data": "synthetic (seeded 9-channel PBR set, random-init latents and decoder)",
        "config": {"workload": "4096^2 9-ch synthetic set; g0 16ch@2048^2 4-bit, g1 16ch@1024^2 4-bit; "
                               "MLP 40->128->128->9 ReLU; 8 x 256^2 tiles = 524288 texels/step; "
                               "8000-step schedule",
                   "path": a.path, "global_batch": TILES * 65536, "parallelism": "single GPU"},
    }

    if a.path == "fused":
        times = (ctypes.c_float * 10)()
        lib.ntc_ctx_kernel_times(ctx._h, times, 10)
        fused_ms = float(times[0])
        # the three-launch step (gather, MLP, scatter): its MLP launch alone,
        # events right around it in the same timed steps
        mlp_ms = float(times[9])
        split = mlp_ms > 0
        # the step's other kernels: an event around each, over the next
        # `detail_steps` schedule steps (after the timed region)
        lib.ntc_ctx_set_timing(ctx._h, 1)
        prog = ctx.Train(a.detail_steps)
        lib.ntc_ctx_flush_grids(ctx._h)
        lib.ntc_dev_synchronize()
        if prog.result != ntc.Result.Success:
            raise SystemExit(f"kernel-time sample failed: {prog}")
        lib.ntc_ctx_kernel_times(ctx._h, times, 10)
        core_ms, dec_ms, grid_ms, n_t, n_grid, n_tiles_upd, catch_ms, hot_ms, n_lazy, _ = list(times)
        if split:
            # the three launches together: events around them in the sample
            # (the timed region has events around the MLP launch only)
            fused_ms = core_ms
        _, dinfos = step_texels(ctx, a.warmup + a.steps, a.detail_steps)
        tiles_upd = n_grid > 0 and n_tiles_upd == n_grid
        lazy = n_grid > 0 and n_lazy == n_grid
        if n_grid > 0 and 0 < n_tiles_upd < n_grid:
            raise SystemExit("timed steps mix the two grid-update kernels")
        if n_grid > 0 and 0 < n_lazy < n_grid:
            raise SystemExit("timed steps mix the deferred and one-pass grid updates")
        # algorithmic work per launch, averaged over the timed steps:
        #   fused launch: 6 (F H + H H + H O) flop per texel; HBM bytes = the
        #   footprint cells read once (4 B/param), targets (O B/texel) and the
        #   decoder-gradient partials;
        #   grid update: 32 B/param (p, m, v, g read; p, m, v, g = 0 written)
        #   inside the tile footprints and 24 B/param (g untouched: zero)
        #   outside them, or 32 B/param over the level (k_grid_update2); with
        #   the zero-gradient steps deferred (the default), the catch-up reads
        #   and writes p, m, v of the footprint cells (24 B/param) and the hot
        #   part 32 B/param there; the deferred steps of the other cells cost
        #   their catch-up later, when a footprint or a flush reaches them
        #   (fused figures over the timed steps, grid figures over the sample)
        n_dec = H * F + H * H + 16 * H

        def cells(si):
            M = DIM >> si.lod
            D0, D1 = lay.g0_dim[si.level], lay.g1_dim[si.level]
            pl = D0 * D0 * lay.g0_channels + D1 * D1 * lay.g1_channels
            pf = footprint_cells(si, M, D0) * lay.g0_channels + footprint_cells(si, M, D1) * lay.g1_channels
            return pl, pf

        def grid_bytes(si, pl, pf):
            if lazy:
                return 24.0 * pf + 32.0 * pf, 24.0 * (pl - pf) + 32.0 * pf
            if tiles_upd:
                return 24.0 * (pl - pf) + 32.0 * pf, 24.0 * (pl - pf) + 32.0 * pf
            return 32.0 * pl, 32.0 * pl

        # three-launch step: the MLP launch reads the feature rows (2 B per
        # feature slot, F rounded up to 8) and the targets, writes the dX rows
        # (4 B per grid slot, G of them) and the partials; the gather reads
        # the footprint cells and writes the feature rows, the scatter reads
        # the dX rows and adds into the footprint cells' gradients (8 B/param)
        fxs = (F + 7) // 8 * 8
        gsl = -(-(lay.g0_channels + lay.g1_channels) // 16) * 16
        def launch_work(steps):
            """Per step: flops, bytes of the fused launch (the three launches),
            bytes of the MLP launch, its workgroups, dense-equivalent grid bytes."""
            flops, fbytes, mbytes, blks, dense = [], [], [], [], []
            for si in steps:
                B = si.n_tiles * si.tile_dim ** 2
                pl, pf = cells(si)
                blocks = min(256, (B + (127 if split else 63)) // (128 if split else 64))
                flops.append(6.0 * (F * H + H * H + H * O) * B)
                blks.append(blocks)
                if split:
                    mbytes.append(2.0 * fxs * B + O * B + 4.0 * gsl * B + 4.0 * blocks * n_dec)
                    fbytes.append(4.0 * pf + 2.0 * fxs * B + mbytes[-1] + 4.0 * gsl * B + 8.0 * pf)
                else:
                    fbytes.append(4.0 * pf + O * B + 4.0 * blocks * n_dec)
                if si.grid_update:
                    dense.append(grid_bytes(si, pl, pf)[1])
            return flops, fbytes, mbytes, blks, dense

        # the events of the timed region: every TIMING_STRIDE-th timed step
        flops, fbytes, mbytes, blks, _ = launch_work(infos[::TIMING_STRIDE])
        if split:
            # the three launches together were timed over the sample after it
            flops, fbytes, _, blks, _ = launch_work(dinfos)
            mflops, _, mbytes, mblks, _ = launch_work(infos[::TIMING_STRIDE])
        dense = launch_work(infos)[4]
        gbytes, cbytes, hbytes = [], [], []
        for si in dinfos:
            if not si.grid_update:
                continue
            pl, pf = cells(si)
            gbytes.append(grid_bytes(si, pl, pf)[0])
            if lazy:
                cbytes.append(24.0 * pf)
                hbytes.append(32.0 * pf)
        fused_tflops = np.mean(flops) / (fused_ms * 1e-3) / 1e12
        fused_gbs = np.mean(fbytes) / (fused_ms * 1e-3) / 1e9
        kern = {"fused_train_step": {"kernels": ("k_split_gather + k_mlp_train + k_split_scatter" if split
                                                 else "k_fused_train"),
                                     "ms": fused_ms, "tflops": fused_tflops, "mfma_frac": fused_tflops / BF16_PEAK_TFLOPS,
                                     "GBps": fused_gbs, "hbm_frac": fused_gbs / HBM_PEAK_GBS,
                                     "flop_per_launch": float(np.mean(flops)),
                                     "bytes_per_launch": float(np.mean(fbytes)),
                                     "workgroups": float(np.mean(blks))},
                "decoder_update": {"ms": dec_ms},
                "sample": (f"MLP launch: every {TIMING_STRIDE}th of the {a.steps} timed steps; the three launches "
                           f"together, decoder update and grid kernels: the {a.detail_steps} steps after them, an "
                           f"event around each" if split else
                           f"fused kernel: every {TIMING_STRIDE}th of the {a.steps} timed steps; decoder update and "
                           f"grid kernels: the {a.detail_steps} steps after them, an event around each kernel")}
        if gbytes and lazy:
            c_gbs = float(np.mean(cbytes)) / (catch_ms * 1e-3) / 1e9
            hot_gbs = float(np.mean(hbytes)) / (hot_ms * 1e-3) / 1e9
            kern["grid_update"] = {"kernel": "k_grid_lazy_catchup_rects + k_grid_lazy_hot", "ms": grid_ms,
                                   "bytes_per_launch": float(np.mean(gbytes)),
                                   "catch_up": {"kernel": "k_grid_lazy_catchup_rects", "ms": catch_ms,
                                                "achieved": c_gbs, "unit": "GB/s", "frac": c_gbs / HBM_PEAK_GBS,
                                                "bytes_per_launch": float(np.mean(cbytes))},
                                   "hot": {"kernel": "k_grid_lazy_hot", "ms": hot_ms, "achieved": hot_gbs,
                                           "unit": "GB/s", "frac": hot_gbs / HBM_PEAK_GBS,
                                           "bytes_per_launch": float(np.mean(hbytes))}}
        elif gbytes:
            grid_gbs = float(np.mean(gbytes)) / (grid_ms * 1e-3) / 1e9
            kern["grid_update"] = {"kernel": "k_grid_update_tiles" if tiles_upd else "k_grid_update2",
                                   "ms": grid_ms, "achieved": grid_gbs, "unit": "GB/s", "peak": HBM_PEAK_GBS,
                                   "frac": grid_gbs / HBM_PEAK_GBS, "bytes_per_launch": float(np.mean(gbytes))}
        # the whole step against HBM: the bytes an eager dense update of the
        # level plus the fused launch would move, over the measured step time
        # (with deferred steps the grid part is not moved every step: this is
        # the equivalent rate, not measured traffic)
        step_bytes = float(np.mean(launch_work(infos)[1])) + (float(np.sum(dense)) / len(infos) if dense else 0.0)
        kern["step"] = {"ms": dt * 1e3 / a.steps, "dense_equivalent_bytes": step_bytes,
                        "achieved": step_bytes / (dt / a.steps) / 1e9, "unit": "GB/s",
                        "frac": step_bytes / (dt / a.steps) / 1e9 / HBM_PEAK_GBS}
        if split:
            mlp_tflops = np.mean(mflops) / (mlp_ms * 1e-3) / 1e12
            mlp_gbs = np.mean(mbytes) / (mlp_ms * 1e-3) / 1e9
            kern["mlp_train"] = {"kernel": "k_mlp_train", "ms": mlp_ms, "tflops": mlp_tflops,
                                 "mfma_frac": mlp_tflops / BF16_PEAK_TFLOPS, "GBps": mlp_gbs,
                                 "hbm_frac": mlp_gbs / HBM_PEAK_GBS, "flop_per_launch": float(np.mean(mflops)),
                                 "bytes_per_launch": float(np.mean(mbytes)), "workgroups": float(np.mean(mblks)),
                                 "sample": f"every {TIMING_STRIDE}th of the {a.steps} timed steps "
                                           f"({len(infos[::TIMING_STRIDE])} launches), events right around the launch"}
        out["kernels"] = kern
        # the roofline names the longest-running kernel of the step
        if split:
            m = kern["mlp_train"]
            out["roofline"] = {"bound": "mfma", "kernel": "k_mlp_train", "achieved": m["tflops"],
                               "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": m["mfma_frac"],
                               "traffic": pmc_traffic("k_mlp_train")}
        elif gbytes and not lazy and grid_ms >= fused_ms:
            g = kern["grid_update"]
            out["roofline"] = {"bound": "hbm", "kernel": g["kernel"], "achieved": g["achieved"],
                               "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": g["frac"],
                               "traffic": pmc_traffic(g["kernel"])}
        else:
            f = kern["fused_train_step"]
            if f["hbm_frac"] >= f["mfma_frac"]:
                out["roofline"] = {"bound": "hbm", "kernel": "k_fused_train", "achieved": fused_gbs,
                                   "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": f["hbm_frac"],
                                   "traffic": pmc_traffic("k_fused_train")}
            else:
                out["roofline"] = {"bound": "mfma", "kernel": "k_fused_train", "achieved": fused_tflops,
                                   "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": f["mfma_frac"],
                                   "traffic": pmc_traffic("k_fused_train")}

    if a.path == "fused":
        out["kernels"]["fused_train_step_lod0"] = lod0_micro(ntc, ctx)
        l0 = out["kernels"]["fused_train_step_lod0"]
        log(f"lod-0 step kernels: one launch {l0['one_launch']['ms']:.4f} ms"
            + (f", three launches {l0['three_launches']['ms']:.4f} ms (MLP {l0['three_launches']['mlp_ms']:.4f} ms)"
               if "three_launches" in l0 else ""))
    if ttp is not None:
        out["time_to_psnr"] = {k: ttp[k] for k in ("metric", "plain_fp32", "target_psnr", "final_within_0.05dB",
                                                    "seconds_to_target", "steps_to_target",
                                                    "speedup_full_schedule")}
        out["time_to_psnr"]["fused_bf16"] = {k: ttp["fused_bf16"][k] for k in ("train_s", "final_psnr", "trace")}
    if not a.no_cpu_baseline:
        log(f"cpu baseline: {a.cpu_steps} oracle steps")
        out["cpu_baseline"] = cpu_baseline(ctx, a.warmup, a.cpu_steps)
        log("cpu baseline done")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
