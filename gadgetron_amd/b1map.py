"""Coil sensitivity (B1) map estimation (include/gt_b1map.h).

Mirrors estimate_b1_map<float, D> (toolboxes/mri/pmri/gpu/b1_map.h:16-24,
b1_map.cu:32-108): RSS-normalised coil images, 7^D box-smoothed coil
correlation, two power iterations, phase of coil 0 removed.  Arrays are torch
tensors or DeviceArrays on the GPU in numpy order: coil images (C, *N), maps
(T, *N), complex64.
"""
import numpy as np

from ._lib import check, lib, sizes
from .nfft import DeviceArray, _ptr, _stream

__all__ = ["estimate_b1_map"]


def estimate_b1_map(data, target_coils=-1, out=None, stream=None):
    """data: (C, *N) coil images, D = len(N) in {2, 3}; returns (T, *N) with T
    = target_coils if 0 < target_coils <= C, else C."""
    shape = [int(s) for s in data.shape]
    if len(shape) < 3:
        raise ValueError("estimate_b1_map:: dimensionality mismatch.")
    C, img = shape[0], shape[1:]
    D = len(img)
    T = C if (target_coils <= 0 or target_coils > C) else int(target_coils)
    out = DeviceArray.zeros((T,) + tuple(img), np.complex64) if out is None else out
    check(lib().gt_estimate_b1_map(_ptr(data), D, sizes(list(reversed(img))), C, int(target_coils), _ptr(out),
                                   _stream(stream)))
    return out
