"""gadgetron_amd -- MI355X-native non-Cartesian NFFT with Gadgetron's cuNFFT_plan API.

The compute path is libgtnufft.so (hand-written HIP for gfx950 + rocFFT)
behind the C ABI of include/gt_nufft.h; this package is its Python front end.
"""
from ._lib import EXPORTED_SYMBOLS, GTError, LIB_PATH, lib
from .nfft import (
    ConvolutionType,
    DeviceArray,
    NFFT_comp_mode,
    NFFT_conv_mode,
    NFFT_fft_mode,
    NFFT_prep_mode,
    NFFTOperator,
    NFFTPlan,
    device_count,
)
from .sense import CgBatch, SenseOperator, cg_sense
from . import b1map, buffer, trajectories

__all__ = [
    "CgBatch", "ConvolutionType", "DeviceArray", "EXPORTED_SYMBOLS", "GTError", "LIB_PATH", "NFFT_comp_mode",
    "NFFT_conv_mode", "NFFT_fft_mode", "NFFT_prep_mode", "NFFTOperator", "NFFTPlan", "SenseOperator", "cg_sense",
    "b1map", "buffer", "device_count", "lib", "trajectories",
]
