"""Python mirror of openntc::Context over the flat C shim (include/ntc_hip.h).

Method names, arguments and return values follow the C++ API
(include/libopenntc/libopenntc.h, itself the reference's libopenntc.h:168-323):
results are ``Result`` values, Train returns a ``TrainProgress``, Eval an
``EvalResults``. All work happens in the native library.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass

import numpy as np

from . import _lib as L


class Profile(enum.IntEnum):
    Bpp_0_2 = 0
    Bpp_0_5 = 1
    Bpp_1_0 = 2
    Bpp_2_25 = 3


class Result(enum.IntEnum):
    Success = 0
    InvalidDimension = 1
    AllocationFailure = 2
    InvalidManifest = 3
    FileNotFound = 4
    InvalidFile = 5
    FileWriteFailure = 6
    InvalidState = 7
    CudaFailure = 8


class Semantic(enum.IntEnum):
    none = 0
    Albedo = 1
    Alpha = 2
    Displacement = 3
    Emissive = 4
    Gloss = 5
    Metallic = 6
    Normal = 7
    AO = 8
    Roughness = 9
    Specular = 10
    Transmission = 11


class Quality(enum.IntEnum):
    Low = 0
    Medium = 1
    High = 2
    Ultra = 3


class TrainPhase(enum.IntEnum):
    ManifestLoaded = 0
    TrainInProgress = 1
    TrainComplete = 2
    TrainError = -1


class FeatureLayout(enum.IntEnum):
    Reference = 0
    Bilinear = 1


class Activation(enum.IntEnum):
    HardGELU = 0
    ReLU = 1


class TrainPath(enum.IntEnum):
    Fused = 0
    Plain = 1


class EvalPrecision(enum.IntEnum):
    FP32 = 0
    BF16 = 1
    FP8 = 2
    INT8 = 3
    BF16_MFMA = 4
    FP8_MFMA = 5


@dataclass
class TrainProgress:
    phase: TrainPhase
    result: Result
    batches_complete: int
    total_batches: int


@dataclass
class EvalResults:
    mse: float
    psnr: float
    channel_sse: np.ndarray | None = None


@dataclass
class LayoutInfo:
    g0_bits: int
    g1_bits: int
    g0_channels: int
    g1_channels: int
    g0_scale: int
    hidden_dim: int
    feature_layout: FeatureLayout = FeatureLayout.Reference
    activation: Activation = Activation.HardGELU


class Context:
    """openntc::Context. Non-copyable; release with Destroy() or del."""

    def __init__(self):
        self._lib = L.load()
        self._h = self._lib.ntc_ctx_create()
        if not self._h:
            raise RuntimeError("ntc_ctx_create failed")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.ntc_ctx_destroy(h)
            self._h = None

    # ---- reference API ----
    def Init(self, profile: Profile) -> Result:
        return Result(self._lib.ntc_ctx_init(self._h, int(profile)))

    def Destroy(self):
        if self._h:
            self._lib.ntc_ctx_destroy(self._h)
            self._h = None

    def LoadManifest(self, path: str) -> Result:
        return Result(self._lib.ntc_ctx_load_manifest(self._h, path.encode()))

    def UnloadManifest(self):
        self._lib.ntc_ctx_unload_manifest(self._h)

    def BeginTraining(self, quality: Quality, grids_per_batch: int) -> Result:
        return Result(self._lib.ntc_ctx_begin_training(self._h, int(quality), grids_per_batch))

    def Train(self, num_batches: int) -> TrainProgress:
        p = L.Progress()
        self._lib.ntc_ctx_train(self._h, num_batches, C.byref(p))
        return TrainProgress(TrainPhase(p.phase), Result(p.result), p.batches_complete, p.total_batches)

    def TrainUntilComplete(self) -> TrainProgress:
        p = L.Progress()
        self._lib.ntc_ctx_train_until_complete(self._h, C.byref(p))
        return TrainProgress(TrainPhase(p.phase), Result(p.result), p.batches_complete, p.total_batches)

    def Eval(self) -> EvalResults:
        return self.EvalEx(EvalPrecision.FP32)

    def GetLastError(self) -> Result:
        return Result(self._lib.ntc_ctx_get_last_error(self._h))

    def Calibrate(self, headroom: float = 1.1):
        out = (C.c_float * 4)()
        r = self._lib.ntc_ctx_calibrate(self._h, headroom, out)
        return Result(r), list(out)

    def FeatureLevelForLod(self, lod: int) -> int:
        return self._lib.ntc_ctx_feature_level_for_lod(self._h, lod)

    def GetMipDim(self, mip: int) -> int:
        lay = self.layout()
        return lay.mip_dim[mip] if 0 <= mip < lay.mip_count else 0

    def GetCompressedData(self):
        """Returns (g0 bytes, g1 bytes, decoder bytes, caldata[4]) or raises."""
        sizes = [C.c_size_t(), C.c_size_t(), C.c_size_t()]
        r = self._lib.ntc_ctx_get_compressed(self._h, None, C.byref(sizes[0]), None, C.byref(sizes[1]),
                                             None, C.byref(sizes[2]), None)
        if r != 0:
            raise RuntimeError(f"GetCompressedData: {Result(r).name}")
        g0 = np.zeros(sizes[0].value, np.uint8)
        g1 = np.zeros(sizes[1].value, np.uint8)
        dec = np.zeros(sizes[2].value, np.uint8)
        cal = np.zeros(4, np.float32)
        r = self._lib.ntc_ctx_get_compressed(self._h, g0.ctypes.data, None, g1.ctypes.data, None,
                                             dec.ctypes.data, None, cal.ctypes.data)
        if r != 0:
            raise RuntimeError(f"GetCompressedData: {Result(r).name}")
        return g0, g1, dec, cal

    def Dump(self, path: str) -> Result:
        return Result(self._lib.ntc_ctx_dump(self._h, path.encode()))

    # ---- extensions ----
    def InitEx(self, profile: Profile = Profile.Bpp_0_2, layout: LayoutInfo | None = None,
               min_dimension: int = 0, mip_min_dim: int = 0, seed: int = 0) -> Result:
        info = L.InitEx()
        info.profile = int(profile)
        info.use_layout = 1 if layout is not None else 0
        if layout is not None:
            li = info.layout
            li.g0_bits, li.g1_bits = layout.g0_bits, layout.g1_bits
            li.g0_channels, li.g1_channels = layout.g0_channels, layout.g1_channels
            li.g0_scale, li.hidden_dim = layout.g0_scale, layout.hidden_dim
            li.feature_layout, li.activation = int(layout.feature_layout), int(layout.activation)
        info.min_dimension = min_dimension
        info.mip_min_dim = mip_min_dim
        info.seed = seed
        return Result(self._lib.ntc_ctx_init_ex(self._h, C.byref(info)))

    def LoadTextures(self, textures, dim: int) -> Result:
        """textures: list of (rgba8 ndarray [dim,dim,4] uint8, Semantic, channels str)."""
        n = len(textures)
        self._tex_keep = [np.ascontiguousarray(t[0], np.uint8) for t in textures]
        ptrs = (C.c_void_p * n)(*[a.ctypes.data for a in self._tex_keep])
        sems = (C.c_int32 * n)(*[int(t[1]) for t in textures])
        chans = (C.c_char_p * n)(*[t[2].encode() for t in textures])
        return Result(self._lib.ntc_ctx_load_textures(self._h, n, ptrs, sems, chans, dim))

    def BeginTrainingEx(self, steps: int, grids_per_batch: int, lr_grid=0.01, lr_decoder=0.005,
                        init_lo=-0.1, init_hi=0.1, path: TrainPath = TrainPath.Fused) -> Result:
        info = L.TrainEx(steps, grids_per_batch, lr_grid, lr_decoder, init_lo, init_hi, int(path))
        return Result(self._lib.ntc_ctx_begin_training_ex(self._h, C.byref(info)))

    def EvalEx(self, precision: EvalPrecision = EvalPrecision.FP32) -> EvalResults:
        mse, psnr = C.c_double(), C.c_double()
        ch = np.zeros(16, np.float64)
        r = self._lib.ntc_ctx_eval(self._h, int(precision), C.byref(mse), C.byref(psnr), ch.ctypes.data)
        if r != 0:
            raise RuntimeError(f"Eval failed: {Result(r).name}")
        return EvalResults(mse.value, psnr.value, ch)

    def layout(self) -> L.CtxLayout:
        out = L.CtxLayout()
        self._lib.ntc_ctx_get_layout(self._h, C.byref(out))
        return out

    def step_info(self, step: int) -> L.StepInfo:
        out = L.StepInfo()
        r = self._lib.ntc_ctx_step_info(self._h, step, C.byref(out))
        if r != 0:
            raise IndexError(step)
        return out

    def _grid_shape(self, which: int, level: int):
        lay = self.layout()
        if which in (0, 2, 3, 6):
            d, c = lay.g0_dim[level], lay.g0_channels
        else:
            d, c = lay.g1_dim[level], lay.g1_channels
        return (d, d, c)

    def read_grid(self, which: int, level: int) -> np.ndarray:
        a = np.zeros(self._grid_shape(which, level), np.float32)
        L.check(self._lib.ntc_ctx_read_grid(self._h, which, level, a.ctypes.data), "read_grid")
        return a

    def write_grid(self, which: int, level: int, a: np.ndarray):
        a = np.ascontiguousarray(a, np.float32)
        assert a.shape == self._grid_shape(which, level)
        L.check(self._lib.ntc_ctx_write_grid(self._h, which, level, a.ctypes.data), "write_grid")

    def decoder_elems(self) -> int:
        lay = self.layout()
        H, F = lay.hidden_dim, lay.feature_dim
        return H * F + H * H + 16 * H

    def read_decoder(self, which: int = 0) -> np.ndarray:
        a = np.zeros(self.decoder_elems(), np.float32)
        L.check(self._lib.ntc_ctx_read_decoder(self._h, which, a.ctypes.data), "read_decoder")
        return a

    def write_decoder(self, a: np.ndarray, which: int = 0):
        a = np.ascontiguousarray(a, np.float32)
        assert a.size == self.decoder_elems()
        L.check(self._lib.ntc_ctx_write_decoder(self._h, which, a.ctypes.data), "write_decoder")

    def read_package(self, mip: int) -> np.ndarray:
        lay = self.layout()
        M = lay.mip_dim[mip]
        a = np.zeros((M, M, lay.out_dim), np.uint8)
        L.check(self._lib.ntc_ctx_read_package(self._h, mip, a.ctypes.data), "read_package")
        return a

    def stream(self) -> int:
        return self._lib.ntc_ctx_stream(self._h) or 0


class Batch:
    """Several contexts (each after BeginTrainingEx on the fused path, same
    decoder shape and step count) trained together: one launch per stage and
    step for all of them, optionally replayed from one captured hipGraph.
    pass_texels / fused_grid 0: tuned by timing the first steps."""

    def __init__(self, contexts, pass_texels: int = 0, fused_grid: int = 0):
        self._lib = L.load()
        self._contexts = list(contexts)  # keeps the contexts alive
        arr = (C.c_void_p * len(self._contexts))(*[c._h for c in self._contexts])
        h = C.c_void_p()
        r = self._lib.ntc_batch_create(arr, len(self._contexts), pass_texels, fused_grid, C.byref(h))
        if r != 0:
            raise RuntimeError(f"ntc_batch_create failed: {Result(r).name}")
        self._h = h.value

    def train(self, n_steps: int, graph: bool = True) -> Result:
        return Result(self._lib.ntc_batch_train(self._h, n_steps, int(graph)))

    def stats(self) -> L.BatchStats:
        out = L.BatchStats()
        L.check(self._lib.ntc_batch_get_stats(self._h, C.byref(out)), "ntc_batch_get_stats")
        return out

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ntc_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def synthetic_pbr(dim: int, seed: int):
    """Seeded synthetic PBR set: returns the three RGBA8 images and the texture
    list for LoadTextures (9 channels: albedo RGB, metallic, normal RGB, AO,
    roughness after the semantic sort)."""
    lib = L.load()
    albedo = np.zeros((dim, dim, 4), np.uint8)
    normal = np.zeros((dim, dim, 4), np.uint8)
    orm = np.zeros((dim, dim, 4), np.uint8)
    L.check(lib.ntc_synthetic_pbr(dim, seed, albedo.ctypes.data, normal.ctypes.data, orm.ctypes.data),
            "ntc_synthetic_pbr")
    textures = [
        (albedo, Semantic.Albedo, "RGB"),
        (normal, Semantic.Normal, "RGB"),
        (orm, Semantic.AO, "R"),
        (orm, Semantic.Roughness, "G"),
        (orm, Semantic.Metallic, "B"),
    ]
    return textures
