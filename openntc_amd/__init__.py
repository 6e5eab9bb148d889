"""openntc_amd: MI355X (gfx950) implementation of the openntc training hot path.

Native library: libopenntc_amd.so (HIP kernels + C++ Context + C ABI, see
include/ntc_hip.h and include/libopenntc/libopenntc.h). This package only
binds it; it contains no compute of its own.
"""
from ._lib import load, check, LIB_PATH  # noqa: F401
from .context import (  # noqa: F401
    Activation, Batch, Context, EvalPrecision, FeatureLayout, LayoutInfo, Profile, Quality, Result,
    Semantic, TrainPath, TrainPhase, synthetic_pbr,
)
