"""ctypes binding of the C oracle (oracle/libntc_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker used by tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg. The product package never imports this.
Array helpers take and return numpy arrays; all computation is in C.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libntc_oracle.so")

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


class Layout(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("mip_count", C.c_int32), ("level_count", C.c_int32),
        ("out_dim", C.c_int32),
        ("mip_dim", C.c_int32 * 12),
        ("g0_dim", C.c_int32 * 6), ("g1_dim", C.c_int32 * 6),
        ("g0_channels", C.c_int32), ("g1_channels", C.c_int32),
        ("g0_bits", C.c_int32), ("g1_bits", C.c_int32), ("g0_scale", C.c_int32),
        ("feature_dim", C.c_int32), ("feature_dim_padded", C.c_int32),
        ("hidden_dim", C.c_int32), ("out_dim_padded", C.c_int32),
        ("g0_mode", C.c_int32), ("pe_first", C.c_int32), ("pe_count", C.c_int32),
        ("lod_slot", C.c_int32), ("activation", C.c_int32),
        ("g0_delta", C.c_float), ("g1_delta", C.c_float),
    ]


class Step(C.Structure):
    _fields_ = [
        ("lod", C.c_int32), ("level", C.c_int32), ("tile_dim", C.c_int32),
        ("n_tiles", C.c_int32), ("origins", C.c_int32 * 16),
        ("lr_grid", C.c_float), ("lr_decoder", C.c_float),
        ("bias1", C.c_float), ("bias2", C.c_float),
        ("grid_bias1", C.c_float), ("grid_bias2", C.c_float),
        ("grid_update", C.c_int32), ("lock", C.c_int32),
    ]


_P = C.c_void_p
_I = C.c_int32
_I64 = C.c_int64
_U = C.c_uint32
_F = C.c_float
_D = C.c_double


def _declare(L):
    L.o_mt_size.restype = C.c_size_t
    L.o_mt_seed.argtypes = [_P, _U]
    L.o_mt_next.argtypes = [_P]
    L.o_mt_next.restype = _U
    L.o_uniform01.argtypes = [_P]
    L.o_uniform01.restype = _F
    L.o_uniform01_open0.argtypes = [_P]
    L.o_uniform01_open0.restype = _F
    L.o_uniform_int.argtypes = [_P, _I, _I]
    L.o_uniform_int.restype = _I
    L.o_bernoulli.argtypes = [_P, _D]
    L.o_bernoulli.restype = _I
    L.o_layout_profile.argtypes = [_I, _I, _I, _I, C.POINTER(Layout)]
    L.o_layout_custom.argtypes = [_I] * 11 + [C.POINTER(Layout)]
    L.o_level_for_lod.argtypes = [C.POINTER(Layout), _I]
    L.o_compressed_bytes.argtypes = [C.POINTER(Layout)]
    L.o_compressed_bytes.restype = C.c_uint64
    L.o_decoder_bytes.argtypes = [C.POINTER(Layout)]
    L.o_decoder_bytes.restype = C.c_uint64
    L.o_init_params.argtypes = [_P, C.POINTER(Layout), _F, _F, _P, _P, _P, _P, _P]
    L.o_schedule.argtypes = [_P, C.POINTER(Layout), _I, _I, _F, _F, C.POINTER(Step)]
    L.o_cosine_annealing.argtypes = [_F, _F, _I, _I]
    L.o_cosine_annealing.restype = _F
    L.o_philox4x32_10.argtypes = [_P, _P, _P]
    L.o_noise.argtypes = [_U, _U, _I, _I, _U, _I, _F]
    L.o_noise.restype = _F
    for sfx in ("f32", "f64"):
        getattr(L, f"o_draw_features_{sfx}").argtypes = (
            [_I] * 9 + [_P, _I, _I, _I, _F, _U, _U, _I, _I, _F, _F, _P, _P, _P])
        getattr(L, f"o_draw_targets_{sfx}").argtypes = [_P, _I, _I, _I, _I, _P, _P]
        getattr(L, f"o_gemm_{sfx}").argtypes = [_I, _I, _I, _I, _P, _P, _P]
        getattr(L, f"o_act_forward_{sfx}").argtypes = [_I, _I64, _P, _P]
        getattr(L, f"o_act_backward_{sfx}").argtypes = [_I, _I64, _P, _P, _P]
        getattr(L, f"o_loss_{sfx}").argtypes = [_I64, _F if sfx == "f32" else _D, _P, _P, _P]
        getattr(L, f"o_loss_{sfx}").restype = _D
        getattr(L, f"o_scatter_{sfx}").argtypes = [_I] * 5 + [_P, _I, _I, _I, _P, _P, _P]
        getattr(L, f"o_mlp_fwd_bwd_{sfx}").argtypes = [_I, _I, _I, _I, _I64] + [_P] * 10
        getattr(L, f"o_mlp_fwd_bwd_{sfx}").restype = _D
        getattr(L, f"o_mlp_forward_{sfx}").argtypes = [_I, _I, _I, _I, _I64] + [_P] * 8
        getattr(L, f"o_mlp_forward_{sfx}").restype = _D
    L.o_adam_f32.argtypes = [_I64, _F, _F, _F, _F, _F, _P, _P, _P, _P]
    L.o_clamp_grid_f32.argtypes = [_I64, _F, _P]
    L.o_quantize_grid_f32.argtypes = [_I64, _I, _F, _P]
    L.o_quantize_pack_f32.argtypes = [_I64, _I, _P, _P]
    L.o_quantize_weights.argtypes = [_P, _I, _I, _F, _P, _P]
    L.o_lanczos_weights.argtypes = [_P]
    L.o_filter_lanczos.argtypes = [_I, _I, _P, _P]
    L.o_step_workspace_elems.argtypes = [C.POINTER(Layout), _I64]
    L.o_step_workspace_elems.restype = C.c_size_t
    L.o_train_step_f32.argtypes = [C.POINTER(Layout), C.POINTER(Step), _U, _U] + [_P] * 11
    L.o_train_step_f32.restype = _D
    L.o_lock_grids_f32.argtypes = [C.POINTER(Layout), _P, _P]
    L.o_eval_f32.argtypes = [C.POINTER(Layout), _P, _P, _P, _P, _P, _P]
    L.o_eval_f32.restype = _D
    L.o_calibrate_f32.argtypes = [C.POINTER(Layout), _P, _P, _P, _P, _P]
    L.o_bf16_round.argtypes = [_F]
    L.o_bf16_round.restype = _F
    L.o_e4m3_encode.argtypes = [_F]
    L.o_e4m3_encode.restype = C.c_uint8
    L.o_e4m3_decode.argtypes = [C.c_uint8]
    L.o_e4m3_decode.restype = _F
    L.o_clamp_s8.argtypes = [_F]
    L.o_clamp_s8.restype = _I
    L.o_fp8_scales.argtypes = [_I, _I, _I, _P, _F, _F, _P]
    L.o_decode_f32.argtypes = [_I, _I, _I, _I, _I, _I, _I64, _P, _P, _P, _P, _P, _P]
    L.o_decode_f64.argtypes = [_I, _I, _I, _I, _I64, _P, _P, _P]


def ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(C.c_void_p)


def _ptr_array(arrs):
    arr = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    return arr


# ---- RNG -------------------------------------------------------------------
class MT:
    def __init__(self, seed: int):
        self.buf = C.create_string_buffer(lib().o_mt_size())
        lib().o_mt_seed(self.buf, seed)

    def next(self) -> int:
        return lib().o_mt_next(self.buf)

    def uniform01(self) -> float:
        return lib().o_uniform01(self.buf)

    def uniform_int(self, lo, hi) -> int:
        return lib().o_uniform_int(self.buf, lo, hi)


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().o_philox4x32_10(c, k, o)
    return list(o)


def noise(seed, step, grid, level, cell, ch, delta):
    return lib().o_noise(seed, step, grid, level, cell, ch, delta)


# ---- layout ------------------------------------------------------------------
def layout_profile(profile: int, dim: int, out_dim: int = 9, mip_min_dim: int = 4) -> Layout:
    L = Layout()
    if lib().o_layout_profile(profile, dim, out_dim, mip_min_dim, C.byref(L)) != 0:
        raise ValueError("bad layout")
    return L


def layout_custom(g0_bits, g1_bits, g0_ch, g1_ch, g0_scale, hidden, feature_layout,
                  activation, dim, out_dim=9, mip_min_dim=4) -> Layout:
    L = Layout()
    if lib().o_layout_custom(g0_bits, g1_bits, g0_ch, g1_ch, g0_scale, hidden,
                             feature_layout, activation, dim, out_dim, mip_min_dim,
                             C.byref(L)) != 0:
        raise ValueError("bad layout")
    return L


def level_for_lod(L: Layout, lod: int) -> int:
    return lib().o_level_for_lod(C.byref(L), lod)


def compressed_bytes(L: Layout) -> int:
    return int(lib().o_compressed_bytes(C.byref(L)))


def decoder_bytes(L: Layout) -> int:
    return int(lib().o_decoder_bytes(C.byref(L)))


def grid_shapes(L: Layout):
    g0 = [(L.g0_dim[i], L.g0_dim[i], L.g0_channels) for i in range(L.level_count)]
    g1 = [(L.g1_dim[i], L.g1_dim[i], L.g1_channels) for i in range(L.level_count)]
    return g0, g1


def decoder_size(L: Layout) -> int:
    H, F = L.hidden_dim, L.feature_dim
    return H * F + H * H + 16 * H


def feature_desc(L: Layout, level: int = 0):
    return dict(g0_dim=L.g0_dim[level], g1_dim=L.g1_dim[level], c0=L.g0_channels,
                c1=L.g1_channels, g0_mode=L.g0_mode, pe_first=L.pe_first,
                pe_count=L.pe_count, lod_slot=L.lod_slot, F=L.feature_dim)


# ---- init / schedule -----------------------------------------------------------
class Params:
    """Grids per level, decoder W0|W1|Wout, and Adam state (FP32)."""

    def __init__(self, L: Layout):
        g0s, g1s = grid_shapes(L)
        self.g0 = [np.zeros(s, np.float32) for s in g0s]
        self.g1 = [np.zeros(s, np.float32) for s in g1s]
        self.g0m = [np.zeros(s, np.float32) for s in g0s]
        self.g0v = [np.zeros(s, np.float32) for s in g0s]
        self.g1m = [np.zeros(s, np.float32) for s in g1s]
        self.g1v = [np.zeros(s, np.float32) for s in g1s]
        n = decoder_size(L)
        self.dec = np.zeros(n, np.float32)
        self.dec_m = np.zeros(n, np.float32)
        self.dec_v = np.zeros(n, np.float32)

    def copy(self):
        c = Params.__new__(Params)
        for k, v in self.__dict__.items():
            c.__dict__[k] = [a.copy() for a in v] if isinstance(v, list) else v.copy()
        return c


def init_params(mt: MT, L: Layout, lo=-0.1, hi=0.1) -> Params:
    P = Params(L)
    H, F = L.hidden_dim, L.feature_dim
    w0 = P.dec[: H * F]
    w1 = P.dec[H * F: H * F + H * H]
    wout = P.dec[H * F + H * H:]
    lib().o_init_params(mt.buf, C.byref(L), lo, hi, _ptr_array(P.g0), _ptr_array(P.g1),
                        w0.ctypes.data, w1.ctypes.data, wout.ctypes.data)
    return P


def schedule(mt: MT, L: Layout, n_steps: int, grids_per_batch: int, lr_grid=0.01,
             lr_dec=0.005):
    steps = (Step * n_steps)()
    lib().o_schedule(mt.buf, C.byref(L), n_steps, grids_per_batch, lr_grid, lr_dec, steps)
    return steps


# ---- per-stage kernels -------------------------------------------------------------
def _dt(dtype):
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


def draw_features(fd: dict, origins, T, M, norm_lod, g0, g1, seed=123, step=0, level=0,
                  noise_on=False, d0=0.0, d1=0.0, dtype=np.float32):
    origins = np.ascontiguousarray(origins, np.int32)
    n_tiles = len(origins) // 2
    B = n_tiles * T * T
    x = np.zeros((fd["F"], B), dtype)
    g0 = np.ascontiguousarray(g0, dtype)
    g1 = np.ascontiguousarray(g1, dtype)
    getattr(lib(), f"o_draw_features_{_dt(dtype)}")(
        fd["g0_dim"], fd["g1_dim"], fd["c0"], fd["c1"], fd["g0_mode"], fd["pe_first"],
        fd["pe_count"], fd["lod_slot"], fd["F"], ptr(origins), n_tiles, T, M, norm_lod,
        seed, step, level, int(noise_on), d0, d1, ptr(g0), ptr(g1), ptr(x))
    return x


def draw_targets(origins, T, M, O, pkg, dtype=np.float32):
    origins = np.ascontiguousarray(origins, np.int32)
    n_tiles = len(origins) // 2
    y = np.zeros((O, n_tiles * T * T), dtype)
    pkg = np.ascontiguousarray(pkg, np.uint8)
    getattr(lib(), f"o_draw_targets_{_dt(dtype)}")(ptr(origins), n_tiles, T, M, O, ptr(pkg), ptr(y))
    return y


GEMM_AB, GEMM_ATB, GEMM_ABT = 0, 1, 2


def gemm(op, n, m, k, a, b, dtype=np.float32):
    a = np.ascontiguousarray(a, dtype)
    b = np.ascontiguousarray(b, dtype)
    c = np.zeros((n, m), dtype)
    getattr(lib(), f"o_gemm_{_dt(dtype)}")(op, n, m, k, ptr(a), ptr(b), ptr(c))
    return c


def act_forward(act, z):
    z = np.ascontiguousarray(z)
    a = np.zeros_like(z)
    getattr(lib(), f"o_act_forward_{_dt(z.dtype)}")(act, z.size, ptr(z), ptr(a))
    return a


def act_backward(act, z, g):
    z = np.ascontiguousarray(z)
    g = np.ascontiguousarray(g, z.dtype)
    o = np.zeros_like(z)
    getattr(lib(), f"o_act_backward_{_dt(z.dtype)}")(act, z.size, ptr(z), ptr(g), ptr(o))
    return o


def loss(scale, yt, yp):
    yt = np.ascontiguousarray(yt)
    yp = np.ascontiguousarray(yp, yt.dtype)
    dy = np.zeros_like(yt)
    sse = getattr(lib(), f"o_loss_{_dt(yt.dtype)}")(yt.size, scale, ptr(yt), ptr(yp), ptr(dy))
    return sse, dy


def scatter(fd: dict, origins, T, M, dx, dg0_shape, dg1_shape):
    dx = np.ascontiguousarray(dx)
    origins = np.ascontiguousarray(origins, np.int32)
    dg0 = np.zeros(dg0_shape, dx.dtype)
    dg1 = np.zeros(dg1_shape, dx.dtype)
    getattr(lib(), f"o_scatter_{_dt(dx.dtype)}")(
        fd["g0_dim"], fd["g1_dim"], fd["c0"], fd["c1"], fd["g0_mode"], ptr(origins),
        len(origins) // 2, T, M, ptr(dx), ptr(dg0), ptr(dg1))
    return dg0, dg1


def mlp_fwd_bwd(F, H, O, act, x, yt, w0, w1, wout):
    dtype = x.dtype
    B = x.shape[1]
    x = np.ascontiguousarray(x)
    yt = np.ascontiguousarray(yt, dtype)
    w0 = np.ascontiguousarray(w0, dtype)
    w1 = np.ascontiguousarray(w1, dtype)
    wout = np.ascontiguousarray(wout, dtype)
    dw0 = np.zeros((H, F), dtype)
    dw1 = np.zeros((H, H), dtype)
    dwout = np.zeros((O, H), dtype)
    dx = np.zeros((F, B), dtype)
    ws = np.zeros(8 * H * B + 2 * O * B, dtype)
    sse = getattr(lib(), f"o_mlp_fwd_bwd_{_dt(dtype)}")(
        F, H, O, act, B, ptr(x), ptr(yt), ptr(w0), ptr(w1), ptr(wout), ptr(dw0), ptr(dw1),
        ptr(dwout), ptr(dx), ptr(ws))
    return sse, dw0, dw1, dwout, dx


def adam(lr, b1, b2, bias1, bias2, g, m, v, p):
    """In place on m, v, p (float32 arrays)."""
    lib().o_adam_f32(p.size, lr, b1, b2, bias1, bias2, ptr(g), ptr(m), ptr(v), ptr(p))


def clamp_grid(delta, g):
    lib().o_clamp_grid_f32(g.size, delta, ptr(g))


def quantize_grid(bits, delta, g):
    lib().o_quantize_grid_f32(g.size, bits, delta, ptr(g))


def quantize_pack(bits, x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(x.size * bits // 32, np.uint32)
    lib().o_quantize_pack_f32(x.size, bits, ptr(x), ptr(out))
    return out


def quantize_weights(w, prescale):
    w = np.ascontiguousarray(w, np.float32)
    rows, cols = w.shape
    wq = np.zeros(rows * cols // 4, np.uint32)
    ws = np.zeros(rows, np.float32)
    lib().o_quantize_weights(ptr(w), rows, cols, prescale, ptr(wq), ptr(ws))
    return wq, ws


def lanczos_weights():
    w = np.zeros(12, np.float32)
    lib().o_lanczos_weights(ptr(w))
    return w


def filter_lanczos(src):
    src = np.ascontiguousarray(src, np.uint8)
    dim, _, c = src.shape
    dst = np.zeros((dim // 2, dim // 2, c), np.uint8)
    lib().o_filter_lanczos(dim, c, ptr(src), ptr(dst))
    return dst


# ---- textures: channel selection, mips, package (numpy restatement of
# prepare_tex / package_tex, ntc_kernel.cu:533-567, 665-686) ------------------------
def prepare_tex(rgba, cmap):
    idx = [c for c in cmap]
    return np.ascontiguousarray(rgba[:, :, idx])


def build_package(sources, L: Layout):
    """sources: list of (rgba8 array, semantic int, channel string). Stable-sorted
    by semantic like libopenntc.cpp:1195-1199. Returns per-mip [M][M][O] arrays."""
    order = sorted(range(len(sources)), key=lambda i: sources[i][1])
    per_source = []
    for i in order:
        rgba, _, chans = sources[i]
        mip = prepare_tex(rgba, ["RGBA".index(ch) for ch in chans])
        mips = [mip]
        for _ in range(1, L.mip_count):
            mips.append(filter_lanczos(mips[-1]))
        per_source.append(mips)
    pkgs = []
    for m in range(L.mip_count):
        pkgs.append(np.ascontiguousarray(np.concatenate([s[m] for s in per_source], axis=2)))
    return pkgs


# ---- full training step / eval ---------------------------------------------------------
class Trainer:
    """Drives o_train_step_f32 over a schedule (the CPU rendition of the
    reference step loop, libopenntc.cpp:568-790)."""

    def __init__(self, L: Layout, P: Params, packages, seed=123, max_batch=8 * 65536):
        self.L = L
        self.P = P
        self.seed = seed
        self.packages = [np.ascontiguousarray(p) for p in packages]
        self.max_batch = max_batch
        self.ws = np.zeros(lib().o_step_workspace_elems(C.byref(L), max_batch), np.float32)
        self._g0 = _ptr_array(P.g0)
        self._g1 = _ptr_array(P.g1)
        self._g0m = _ptr_array(P.g0m)
        self._g0v = _ptr_array(P.g0v)
        self._g1m = _ptr_array(P.g1m)
        self._g1v = _ptr_array(P.g1v)
        self._pk = _ptr_array(self.packages)

    def step(self, st: Step, index: int) -> float:
        return lib().o_train_step_f32(
            C.byref(self.L), C.byref(st), self.seed, index, self._g0, self._g1, self._g0m,
            self._g0v, self._g1m, self._g1v, self.P.dec.ctypes.data, self.P.dec_m.ctypes.data,
            self.P.dec_v.ctypes.data, self._pk, self.ws.ctypes.data)

    def _eval_batch(self):
        # evaluation walks each mip in tiles of min(M, 256)^2, up to 8 per batch
        T = min(self.L.mip_dim[0], 256)
        return min(8, (self.L.mip_dim[0] // T) ** 2) * T * T

    def eval(self, channel_sse=False):
        if self.max_batch < self._eval_batch():
            raise ValueError(f"Trainer workspace holds {self.max_batch} texels; evaluation needs {self._eval_batch()}")
        cs = np.zeros(16, np.float64)
        mse = lib().o_eval_f32(C.byref(self.L), self._g0, self._g1, self.P.dec.ctypes.data,
                               self._pk, self.ws.ctypes.data, cs.ctypes.data)
        return (mse, cs[: self.L.out_dim]) if channel_sse else mse

    def calibrate(self):
        if self.max_batch < self._eval_batch():
            raise ValueError(f"Trainer workspace holds {self.max_batch} texels; calibration needs {self._eval_batch()}")
        out = np.zeros(2, np.float32)
        lib().o_calibrate_f32(C.byref(self.L), self._g0, self._g1, self.P.dec.ctypes.data,
                              self.ws.ctypes.data, out.ctypes.data)
        return out



# ---- evaluation decode at reduced precision ---------------------------------------------
PREC_FP32, PREC_BF16, PREC_FP8, PREC_INT8 = 0, 1, 2, 3


def bf16_round(x: float) -> float:
    return lib().o_bf16_round(float(x))


def e4m3_encode(x: float) -> int:
    return int(lib().o_e4m3_encode(float(x)))


def e4m3_decode(c: int) -> float:
    return lib().o_e4m3_decode(int(c))


def clamp_s8(x: float) -> int:
    return int(lib().o_clamp_s8(float(x)))


def fp8_scales(F, H, O, dec, max_a0, max_a1):
    dec = np.ascontiguousarray(dec, np.float32)
    s = np.zeros(6, np.float32)
    lib().o_fp8_scales(F, H, O, ptr(dec), float(max_a0), float(max_a1), ptr(s))
    return s


def decode(precision, act, F, Fp, H, O, x, dec=None, s8=None, blob=None, rcp=None):
    """Outputs y [O][B] of the decode of FP32 features x [F][B] at the given
    precision (dec: FP32 W0[H][F] | W1[H][H] | Wout[16][H]; s8: fp8_scales;
    blob: exported INT8 decoder; rcp: (rcp_s_a1, rcp_s_a2))."""
    x = np.ascontiguousarray(x, np.float32)
    B = x.shape[1]
    y = np.zeros((O, B), np.float32)
    dec = None if dec is None else np.ascontiguousarray(dec, np.float32)
    s8 = None if s8 is None else np.ascontiguousarray(s8, np.float32)
    blob = None if blob is None else np.ascontiguousarray(blob, np.uint8)
    rc = None if rcp is None else np.ascontiguousarray(rcp, np.float32)
    lib().o_decode_f32(precision, act, F, Fp, H, O, B, None if dec is None else ptr(dec),
                       None if s8 is None else ptr(s8), None if blob is None else ptr(blob),
                       None if rc is None else ptr(rc), ptr(x), ptr(y))
    return y


def decode_f64(act, F, H, O, x, dec):
    x = np.ascontiguousarray(x, np.float64)
    dec = np.ascontiguousarray(dec, np.float64)
    y = np.zeros((O, x.shape[1]), np.float64)
    lib().o_decode_f64(act, F, H, O, x.shape[1], ptr(dec), ptr(x), ptr(y))
    return y
