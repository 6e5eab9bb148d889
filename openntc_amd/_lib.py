"""ctypes binding of the C ABI in include/ntc_hip.h (libopenntc_amd.so).

The library is built in-tree by ``python -m openntc_amd.build`` (or
``__graft_entry__.build()``). Loading fails loudly when it is missing: there
is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libopenntc_amd.so")

NTC_OK = 0
NTC_E_INVALID = -1
NTC_E_WORKSPACE = -2
NTC_E_NODEVICE = -3

G0_CORNERS, G0_BILINEAR = 0, 1
ACT_HARDGELU, ACT_RELU = 0, 1
GEMM_AB, GEMM_ATB, GEMM_ABT = 0, 1, 2
GEMM_KCHUNK = 1024


class FeatureDesc(C.Structure):
    _fields_ = [("g0_dim", C.c_int32), ("g1_dim", C.c_int32), ("g0_channels", C.c_int32),
                ("g1_channels", C.c_int32), ("g0_mode", C.c_int32), ("pe_first_period", C.c_int32),
                ("pe_count", C.c_int32), ("lod_slot", C.c_int32), ("feature_dim", C.c_int32)]


class Tiles(C.Structure):
    _fields_ = [("origins", C.c_void_p), ("n_tiles", C.c_int32), ("tile_dim", C.c_int32),
                ("mip_dim", C.c_int32), ("norm_lod", C.c_float)]


class Noise(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("step", C.c_uint32), ("level", C.c_int32),
                ("enabled", C.c_int32), ("delta0", C.c_float), ("delta1", C.c_float)]


class GridAdam(C.Structure):
    _fields_ = [("m0", C.c_void_p), ("v0", C.c_void_p), ("m1", C.c_void_p), ("v1", C.c_void_p),
                ("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("bias1", C.c_float),
                ("bias2", C.c_float), ("lo0", C.c_float), ("hi0", C.c_float), ("lo1", C.c_float),
                ("hi1", C.c_float)]


class FusedArgs(C.Structure):
    _fields_ = [("feat", FeatureDesc), ("tiles", Tiles), ("noise", Noise),
                ("hidden_dim", C.c_int32), ("out_dim", C.c_int32), ("activation", C.c_int32),
                ("g0", C.c_void_p), ("g1", C.c_void_p), ("weights", C.c_void_p),
                ("targets", C.c_void_p), ("loss_scale", C.c_float), ("scatter", C.c_int32),
                ("dg0", C.c_void_p), ("dg1", C.c_void_p), ("dw_partials", C.c_void_p),
                ("sse_partials", C.c_void_p), ("blocks", C.c_int32)]


class SplitTailArgs(C.Structure):
    """ntc_split_tail_args: the next step's catch-up (catch_tiles non-null)
    and the decoder update (dec_p non-null) of ntc_split_tail."""
    _fields_ = [("catch_feat", C.c_void_p), ("catch_tiles", C.c_void_p), ("catch_skip", C.c_void_p),
                ("catch_g0", C.c_void_p), ("catch_g1", C.c_void_p), ("catch_adam", C.c_void_p),
                ("catch_steps", C.c_void_p), ("catch_upto", C.c_int32), ("catch_last0", C.c_void_p),
                ("catch_last1", C.c_void_p), ("dec_feat", C.c_void_p), ("dec_lr", C.c_float),
                ("dec_beta1", C.c_float), ("dec_beta2", C.c_float), ("dec_bias1", C.c_float),
                ("dec_bias2", C.c_float), ("dec_grad", C.c_void_p), ("dec_m", C.c_void_p), ("dec_v", C.c_void_p),
                ("dec_p", C.c_void_p), ("dec_table", C.c_void_p), ("sse_out", C.c_void_p)]


class LayoutInfo(C.Structure):
    _fields_ = [("g0_bits", C.c_int32), ("g1_bits", C.c_int32), ("g0_channels", C.c_int32),
                ("g1_channels", C.c_int32), ("g0_scale", C.c_int32), ("hidden_dim", C.c_int32),
                ("feature_layout", C.c_int32), ("activation", C.c_int32)]


class InitEx(C.Structure):
    _fields_ = [("profile", C.c_int32), ("use_layout", C.c_int32), ("layout", LayoutInfo),
                ("min_dimension", C.c_int32), ("mip_min_dim", C.c_int32), ("seed", C.c_uint32)]


class TrainEx(C.Structure):
    _fields_ = [("steps", C.c_int32), ("grids_per_batch", C.c_int32), ("lr_grid", C.c_float),
                ("lr_decoder", C.c_float), ("init_lo", C.c_float), ("init_hi", C.c_float),
                ("path", C.c_int32)]


class Progress(C.Structure):
    _fields_ = [("phase", C.c_int32), ("result", C.c_int32), ("batches_complete", C.c_int32),
                ("total_batches", C.c_int32)]


class CtxLayout(C.Structure):
    _fields_ = [("dim", C.c_int32), ("mip_count", C.c_int32), ("level_count", C.c_int32),
                ("out_dim", C.c_int32), ("mip_dim", C.c_int32 * 12), ("g0_dim", C.c_int32 * 6),
                ("g1_dim", C.c_int32 * 6), ("g0_channels", C.c_int32), ("g1_channels", C.c_int32),
                ("g0_bits", C.c_int32), ("g1_bits", C.c_int32), ("g0_scale", C.c_int32),
                ("feature_dim", C.c_int32), ("feature_dim_padded", C.c_int32),
                ("hidden_dim", C.c_int32), ("out_dim_padded", C.c_int32), ("activation", C.c_int32),
                ("feat", FeatureDesc), ("g0_delta", C.c_float), ("g1_delta", C.c_float),
                ("lock_step", C.c_int32), ("total_steps", C.c_int32),
                ("grids_per_batch", C.c_int32), ("batch_index", C.c_int32), ("seed", C.c_uint32)]


class StepInfo(C.Structure):
    _fields_ = [("lod", C.c_int32), ("level", C.c_int32), ("tile_dim", C.c_int32),
                ("n_tiles", C.c_int32), ("origins", C.c_int32 * 16), ("lr_grid", C.c_float),
                ("lr_decoder", C.c_float), ("bias1", C.c_float), ("bias2", C.c_float),
                ("grid_bias1", C.c_float), ("grid_bias2", C.c_float), ("grid_update", C.c_int32),
                ("lock", C.c_int32)]


class ManifestSource(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("path", C.c_char * 256), ("channels", C.c_char * 8),
                ("semantic", C.c_int32), ("is_srgb", C.c_int32), ("vertical_flip", C.c_int32)]


class Compressed(C.Structure):
    """ntc_compressed: the fields of openntc::CompressedData as plain C types."""
    _fields_ = [("g0", C.c_void_p * 6), ("g1", C.c_void_p * 6), ("decoder", C.c_void_p),
                ("g0_size", C.c_uint64 * 6), ("g1_size", C.c_uint64 * 6),
                ("g0_offset", C.c_uint64 * 6), ("g1_offset", C.c_uint64 * 6),
                ("g0_total_size", C.c_uint64), ("g1_total_size", C.c_uint64),
                ("decoder_size", C.c_uint64), ("g0_grid_dim", C.c_int32 * 6),
                ("g1_grid_dim", C.c_int32 * 6), ("g0_bits", C.c_int32), ("g1_bits", C.c_int32),
                ("g0_channels", C.c_int32), ("g1_channels", C.c_int32), ("dim", C.c_int32),
                ("mip_count", C.c_int32), ("level_count", C.c_int32), ("channel_count", C.c_int32),
                ("channel_semantics", C.c_int32 * 16), ("profile", C.c_int32),
                ("max_abs_a1", C.c_float), ("max_abs_a2", C.c_float), ("s_a1", C.c_float),
                ("s_a2", C.c_float)]


class BatchStats(C.Structure):
    _fields_ = [("n_sets", C.c_int32), ("steps", C.c_int32), ("steps_done", C.c_int32),
                ("pass_texels", C.c_int32), ("fused_grid", C.c_int32), ("grid_update_grid", C.c_int32),
                ("tuned", C.c_int32), ("graph", C.c_int32), ("texels_done", C.c_int64),
                ("tune_ms", C.c_float * 4), ("tune_pass", C.c_int32 * 4), ("tune_grid", C.c_int32 * 4)]


_P = C.c_void_p
_I = C.c_int32
_I64 = C.c_int64
_U = C.c_uint32
_F = C.c_float
_D = C.c_double
_SZ = C.c_size_t

# name -> (restype, argtypes)
SIGNATURES = {
    "ntc_version": (C.c_char_p, []),
    "ntc_device_check": (_I, [_P]),
    "ntc_status_string": (C.c_char_p, [_I]),
    "ntc_draw_features_f32": (_I, [_P, _P, _P, _P, _P, _P, _P]),
    "ntc_draw_features_f64": (_I, [_P, _P, _P, _P, _P, _P, _P]),
    "ntc_draw_targets_f32": (_I, [_P, _I, _P, _P, _P]),
    "ntc_draw_targets_f64": (_I, [_P, _I, _P, _P, _P]),
    "ntc_gemm_workspace_bytes": (_SZ, [_I, _I, _I, _I, _I]),
    "ntc_gemm_f32": (_I, [_I, _I, _I, _I, _P, _P, _P, _P, _SZ, _P]),
    "ntc_gemm_f64": (_I, [_I, _I, _I, _I, _P, _P, _P, _P, _SZ, _P]),
    "ntc_act_forward_f32": (_I, [_I, _I64, _P, _P, _P]),
    "ntc_act_backward_f32": (_I, [_I, _I64, _P, _P, _P, _P]),
    "ntc_act_forward_f64": (_I, [_I, _I64, _P, _P, _P]),
    "ntc_act_backward_f64": (_I, [_I, _I64, _P, _P, _P, _P]),
    "ntc_loss_workspace_bytes": (_SZ, [_I64]),
    "ntc_loss_f32": (_I, [_I64, _F, _P, _P, _P, _P, _P, _P]),
    "ntc_loss_f64": (_I, [_I64, _D, _P, _P, _P, _P, _P, _P]),
    "ntc_channel_sse_workspace_bytes": (_SZ, []),
    "ntc_channel_sse_f32": (_I, [_I, _I64, _P, _P, _P, _P, _P]),
    "ntc_accumulate_grid_gradients_f32": (_I, [_P, _P, _P, _P, _P, _P]),
    "ntc_accumulate_grid_gradients_f64": (_I, [_P, _P, _P, _P, _P, _P]),
    "ntc_adam_f32": (_I, [_I64, _F, _F, _F, _F, _F, _P, _P, _P, _P, _P]),
    "ntc_grid_update_f32": (_I, [_I64, _F, _F, _F, _F, _F, _F, _F, _P, _P, _P, _P, _P]),
    "ntc_clamp_grid_f32": (_I, [_I64, _F, _P, _P]),
    "ntc_quantize_grid_f32": (_I, [_I64, _I, _F, _P, _P]),
    "ntc_quantize_pack_f32": (_I, [_I64, _I, _P, _P, _P]),
    "ntc_max_abs_f32": (_I, [_I64, _P, _P, _P]),
    "ntc_fill_zero": (_I, [_P, _SZ, _P]),
    "ntc_prepare_tex": (_I, [_I, _I, _P, _P, _P, _P]),
    "ntc_filter_lanczos": (_I, [_I, _I, _P, _P, _P, _P]),
    "ntc_package_tex": (_I, [_I, _I, _I, _P, _P, _P, _P, _P, _P]),
    "ntc_ctx_create": (_P, []),
    "ntc_ctx_destroy": (None, [_P]),
    "ntc_ctx_init": (_I, [_P, _I]),
    "ntc_ctx_init_ex": (_I, [_P, _P]),
    "ntc_ctx_load_textures": (_I, [_P, _I, _P, _P, _P, _I]),
    "ntc_ctx_load_manifest": (_I, [_P, C.c_char_p]),
    "ntc_ctx_unload_manifest": (None, [_P]),
    "ntc_ctx_begin_training": (_I, [_P, _I, _I]),
    "ntc_ctx_begin_training_ex": (_I, [_P, _P]),
    "ntc_ctx_train": (_I, [_P, _I, _P]),
    "ntc_ctx_train_until_complete": (_I, [_P, _P]),
    "ntc_ctx_eval": (_I, [_P, _I, _P, _P, _P]),
    "ntc_ctx_calibrate": (_I, [_P, _F, _P]),
    "ntc_ctx_get_last_error": (_I, [_P]),
    "ntc_ctx_get_layout": (_I, [_P, _P]),
    "ntc_ctx_step_info": (_I, [_P, _I, _P]),
    "ntc_ctx_feature_level_for_lod": (_I, [_P, _I]),
    "ntc_ctx_read_grid": (_I, [_P, _I, _I, _P]),
    "ntc_ctx_write_grid": (_I, [_P, _I, _I, _P]),
    "ntc_ctx_read_decoder": (_I, [_P, _I, _P]),
    "ntc_ctx_write_decoder": (_I, [_P, _I, _P]),
    "ntc_ctx_read_package": (_I, [_P, _I, _P]),
    "ntc_ctx_get_compressed": (_I, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "ntc_ctx_dump": (_I, [_P, C.c_char_p]),
    "ntc_ctx_stream": (_P, [_P]),
    "ntc_ctx_kernel_times": (_I, [_P, _P, _I]),
    "ntc_ctx_flush_grids": (_I, [_P]),
    "ntc_ctx_set_timing": (_I, [_P, _I]),
    "ntc_ctx_set_timing_stride": (_I, [_P, _I]),
    "ntc_ctx_set_fused_form": (_I, [_P, _I]),
    "ntc_ctx_fused_form": (_I, [_P]),
    "ntc_debug_inject_fault": (_I, [_I, _I]),
    "ntc_synthetic_pbr": (_I, [_I, _U, _P, _P, _P]),
    "ntc_dev_alloc": (_I, [_SZ, _P]),
    "ntc_dev_free": (_I, [_P]),
    "ntc_dev_upload": (_I, [_P, _P, _SZ]),
    "ntc_dev_download": (_I, [_P, _P, _SZ]),
    "ntc_dev_synchronize": (_I, []),
    "ntc_dev_copy": (_I, [_P, _P, _SZ]),
    "ntc_grid_update2_f32": (_I, [_I64, _P, _P, _P, _P, _F, _F, _I64, _P, _P, _P, _P, _F, _F,
                                  _F, _F, _F, _F, _F, _P]),
    "ntc_fused_supported": (_I, [_P, _I, _I]),
    "ntc_fused_table_elems": (_SZ, [_P, _I]),
    "ntc_fused_max_blocks": (_I, [_P]),
    "ntc_fused_blocks_for": (_I, [_P, _I, _I, _I64, _I]),
    "ntc_pack_decoder_bf16": (_I, [_P, _I, _I, _P, _P, _P]),
    "ntc_fused_train_step": (_I, [_P, _P]),
    "ntc_split_supported": (_I, [_P, _I, _I]),
    "ntc_split_blocks_for": (_I, [_P, _I, _I, _I64, _I]),
    "ntc_split_workspace_bytes": (_SZ, [_P, _I, _I, _I64]),
    "ntc_split_train_step": (_I, [_P, _P, _SZ, _P]),
    "ntc_split_gather": (_I, [_P, _P, _SZ, _P]),
    "ntc_split_mlp": (_I, [_P, _P, _SZ, _P]),
    "ntc_split_scatter": (_I, [_P, _P, _SZ, _P]),
    "ntc_split_tail": (_I, [_P, _P, _SZ, _P, _P]),
    "ntc_grid_update_tiles_f32": (_I, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "ntc_grid_update_tiles_part_f32": (_I, [_P, _P, _P, _P, _P, _P, _P, _I, _I, _P]),
    "ntc_grid_lazy_catchup_f32": (_I, [_P, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P]),
    "ntc_grid_lazy_hot_f32": (_I, [_P, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P]),
    "ntc_grid_lazy_hot_catchup_f32": (_I, [_P, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _P, _P,
                                           _P]),
    "ntc_fill_i32": (_I, [C.c_int64, _I, _P, _P]),
    "ntc_decode_table_bytes": (C.c_size_t, [_I, _I, _I, _I]),
    "ntc_build_decode_table_f": (_I, [_I, _I, _I, _I, _P, _F, _F, _P]),
    "ntc_build_decode_table_int8": (_I, [_I, _I, _I, _I, _P, _F, _F, _P]),
    "ntc_decode_workspace_bytes": (C.c_size_t, [C.c_int64]),
    "ntc_decode_tiles_sse_f32": (_I, [_I, _I, _P, _P, _P, _P, _P, _I, _I, _I, _P, _P, _P, _P]),
    "ntc_decode_sse_f32": (_I, [_I, _I, _I, _I, _I, _I, C.c_int64, _P, _P, _P, _P, _P, _P, _P]),
    "ntc_decoder_update": (_I, [_I, _I, _I, _I, _P, _F, _F, _F, _F, _F, _P, _P, _P, _P, _P, _P, _P]),
    "ntc_decoder_update_bf16": (_I, [_P, _I, _I, _I, _P, _F, _F, _F, _F, _F, _P, _P, _P, _P, _P, _P, _P, _P]),
    "ntc_fused_params_bytes": (C.c_size_t, []),
    "ntc_fused_params_pack": (_I, [_P, _P]),
    "ntc_fused_pass_sizes": (_I, [_P, _I, _I, _P]),
    "ntc_fused_train_sets": (_I, [_P, _I, _I, _I, _P, _P, _I, _P, _I, _P]),
    "ntc_decoder_update_sets": (_I, [_I, _I, _I, _P, _P, _I, _P, _P]),
    "ntc_pack_decoder_bf16_sets": (_I, [_P, _I, _I, _P, _I, _P]),
    "ntc_grid_update_tiles_bytes": (C.c_size_t, []),
    "ntc_grid_update_tiles_pack": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "ntc_grid_update_tiles_sets": (_I, [_P, _P, _I, _P, _I, _P]),
    "ntc_grid_lazy_bytes": (C.c_size_t, []),
    "ntc_grid_lazy_pack": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P]),
    "ntc_grid_lazy_catchup_sets": (_I, [_P, _P, _I, _P, _I, _I, _P]),
    "ntc_grid_lazy_hot_sets": (_I, [_P, _P, _I, _P, _I, _I, _P]),
    "ntc_step_advance": (_I, [_P, _P]),
    "ntc_batch_create": (_I, [_P, _I, _I, _I, _P]),
    "ntc_batch_train": (_I, [_P, _I, _I]),
    "ntc_batch_get_stats": (_I, [_P, _P]),
    "ntc_batch_destroy": (None, [_P]),
    "ntc_manifest_parse": (_I, [C.c_char_p, _I, _P, _P, _P]),
    "ntc_image_decode_rgba8": (_I, [C.c_char_p, _P, _P, _P, _SZ]),
    "ntc_step_forward": (_I, [_P, _P, _P]),
    "ntc_step_backward": (_I, [_P, _P, _P]),
    "ntc_reduce_decoder_partials": (_I, [_I, _I, _I, _P, _P, _P, _P, _P]),
    "ntc_fused_train_step_noalloc_max": (C.c_size_t, []),
    "ntc_fused_train_step_noalloc": (_I, [_P, _P, _P, _P]),
    "ntc_container_dump": (_I, [C.c_char_p, _P]),
    "ntc_container_load": (_I, [C.c_char_p, _P]),
    "ntc_file_data": (_I, [_P, _P]),
    "ntc_file_free": (None, [_P]),
    "ntc_fill_constants": (_I, [_P, _P]),
}

_lib = None


def load():
    """Loads libopenntc_amd.so and declares every C-ABI signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"openntc_amd native library not built: {LIB_PATH} is missing "
            "(run `python -m openntc_amd.build`)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int, what: str = "call"):
    if status != NTC_OK:
        msg = load().ntc_status_string(status).decode()
        raise RuntimeError(f"{what} failed: status {status} ({msg})")
    return status
