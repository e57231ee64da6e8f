"""On-device radial trajectories and density compensation (include/gt_radial.h).

Mirrors core/gpu/radial_utilities.h:13-38 (trajectory and radial DCW
generators) and mri/sdc/SDC.h:25-52 (estimate_dcw) of the reference.  Arrays
are torch tensors or DeviceArrays on the GPU, numpy order: a 2D trajectory
(F, P * S, 2), a 3D kooshball (1, n * S, 3), weights (P * S,).
"""
import ctypes
import enum

import numpy as np

from ._lib import check, lib, sizes
from .nfft import DeviceArray, _numel, _ptr, _stream

__all__ = ["GoldenRatioStep", "trajectory_golden_ratio_2d", "trajectory_fixed_angle_2d",
           "trajectory_variable_angle_2d", "dcw_golden_ratio_2d", "dcw_fixed_angle_2d", "kooshball_3d",
           "estimate_dcw"]


class GoldenRatioStep(enum.IntEnum):
    """GOLDEN_RATIO_ANGULAR_STEP_SIZE (radial_utilities.h:10-13)."""
    GR_SMALLEST = 0
    GR_ORIGINAL = 1


def trajectory_golden_ratio_2d(samples_per_profile, profiles_per_frame, frames=1, profile_offset=0,
                               mode=GoldenRatioStep.GR_ORIGINAL, out=None, stream=None):
    """compute_radial_trajectory_golden_ratio_2d: (F, P * S, 2) float32."""
    S, P, F = int(samples_per_profile), int(profiles_per_frame), int(frames)
    out = DeviceArray.zeros((F, P * S, 2), np.float32) if out is None else out
    check(lib().gt_radial_trajectory_golden_ratio_2d(_ptr(out), S, P, F, int(profile_offset), int(mode),
                                                      _stream(stream)))
    return out


def trajectory_fixed_angle_2d(samples_per_profile, profiles_per_frame, frames=1, angular_offset=0.0, out=None,
                              stream=None):
    """compute_radial_trajectory_fixed_angle_2d: (F, P * S, 2) float32."""
    S, P, F = int(samples_per_profile), int(profiles_per_frame), int(frames)
    out = DeviceArray.zeros((F, P * S, 2), np.float32) if out is None else out
    check(lib().gt_radial_trajectory_fixed_angle_2d(_ptr(out), S, P, F, float(angular_offset), _stream(stream)))
    return out


def trajectory_variable_angle_2d(angles, samples_per_profile, profiles_per_frame, out=None, stream=None):
    """compute_radial_trajectory_variable_angle_2d: angles is a device array of
    one angle per frame (every profile of frame f lies at angles[f], as in the
    reference kernel)."""
    S, P, F = int(samples_per_profile), int(profiles_per_frame), _numel(angles)
    out = DeviceArray.zeros((F, P * S, 2), np.float32) if out is None else out
    check(lib().gt_radial_trajectory_variable_angle_2d(_ptr(out), _ptr(angles), S, P, F, _stream(stream)))
    return out


def dcw_golden_ratio_2d(samples_per_profile, num_profiles, alpha, one_over_radial_oversampling_factor,
                        profile_offset=0, mode=GoldenRatioStep.GR_ORIGINAL, out=None, stream=None):
    """compute_radial_dcw_golden_ratio_2d: (P * S,) float32."""
    S, P = int(samples_per_profile), int(num_profiles)
    out = DeviceArray.zeros((P * S,), np.float32) if out is None else out
    check(lib().gt_radial_dcw_golden_ratio_2d(_ptr(out), S, P, float(alpha), float(one_over_radial_oversampling_factor),
                                               int(profile_offset), int(mode), _stream(stream)))
    return out


def dcw_fixed_angle_2d(samples_per_profile, num_profiles, alpha, one_over_radial_oversampling_factor, out=None,
                       stream=None):
    """compute_radial_dcw_fixed_angle_2d: (P * S,) float32."""
    S, P = int(samples_per_profile), int(num_profiles)
    out = DeviceArray.zeros((P * S,), np.float32) if out is None else out
    check(lib().gt_radial_dcw_fixed_angle_2d(_ptr(out), S, P, float(alpha), float(one_over_radial_oversampling_factor),
                                              _stream(stream)))
    return out


def kooshball_3d(samples_per_spoke, spokes, rotation=None, out=None, stream=None):
    """3D kooshball (no reference counterpart): (1, spokes * S, 3) float32;
    rotation: 3 x 3 numpy array (row-major) or None."""
    S, n = int(samples_per_spoke), int(spokes)
    out = DeviceArray.zeros((1, n * S, 3), np.float32) if out is None else out
    rot = None
    if rotation is not None:
        r = np.ascontiguousarray(rotation, dtype=np.float32).reshape(9)
        rot = (ctypes.c_float * 9)(*r.tolist())
    check(lib().gt_kooshball_trajectory_3d(_ptr(out), S, n, rot, _stream(stream)))
    return out


def estimate_dcw(traj, matrix_size, os_factor=2.1, num_iterations=10, kernel_width=5.5, initial_dcw=None, out=None,
                 stream=None):
    """estimate_dcw (SDC.h:25-52): iterative density compensation with the Jinc
    kernel.  traj: device (..., D) float32 in [-1/2, 1/2], taken as one flattened
    frame; returns (n,) float32 weights."""
    D = int(traj.shape[-1])
    n = _numel(traj) // D
    out = DeviceArray.zeros((n,), np.float32) if out is None else out
    check(lib().gt_sdc_estimate_dcw(_ptr(traj), n, D, sizes(list(matrix_size)), float(os_factor),
                                    int(num_iterations), float(kernel_width), _ptr(initial_dcw), _ptr(out),
                                    _stream(stream)))
    return out
