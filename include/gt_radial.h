/* gt_radial.h -- on-device radial trajectories and density compensation for
 * libgtnufft: the step before the NUFFT path.
 *
 * Replaces (reference files, toolboxes/):
 *   core/gpu/radial_utilities.h:13-38   trajectory and radial DCW generators
 *   mri/sdc/SDC.h:25-52, gpu/cuSDC.cu   estimate_dcw (iterative, Jinc kernel)
 * Device pointers are caller-owned; trajectories are float D-vectors per
 * sample, samples of a profile consecutive, profiles of a frame consecutive,
 * then frames (the reference's [samples per frame, frames] array).  Every call
 * is asynchronous on `stream` unless stated.  The reference's requirement that
 * samples per profile be a multiple of the GPU's warp size does not apply.
 */
#ifndef GT_RADIAL_H
#define GT_RADIAL_H

#include "gt_nufft.h"

#ifdef __cplusplus
extern "C" {
#endif

/* GOLDEN_RATIO_ANGULAR_STEP_SIZE (radial_utilities.h:10-13) */
typedef enum {
    GT_GR_SMALLEST = 0, /* pi (3 - sqrt 5) / 2: 68.75 degrees */
    GT_GR_ORIGINAL = 1  /* pi / ((sqrt 5 + 1) / 2): 111.25 degrees */
} gt_golden_ratio_step;

/* compute_radial_trajectory_golden_ratio_2d (radial_utilities.h:24-28): profile
 * p (counted over all frames) at angle (p + profile_offset) step + pi; sample i
 * at (i - S/2) / S along it.  traj: S * P * F float2 (written). */
gt_status gt_radial_trajectory_golden_ratio_2d(float* traj, unsigned samples_per_profile, unsigned profiles_per_frame,
                                               unsigned frames, unsigned profile_offset, int mode, gt_stream_t stream);

/* compute_radial_trajectory_fixed_angle_2d (radial_utilities.h:19-22): profile p
 * of frame f at angle (p + f / F) pi / P + angular_offset + pi. */
gt_status gt_radial_trajectory_fixed_angle_2d(float* traj, unsigned samples_per_profile, unsigned profiles_per_frame,
                                              unsigned frames, float angular_offset, gt_stream_t stream);

/* compute_radial_trajectory_variable_angle_2d (radial_utilities.h:15-17): as the
 * reference kernel does, every profile of frame f lies at angles[f] (device,
 * F floats). */
gt_status gt_radial_trajectory_variable_angle_2d(float* traj, const float* angles, unsigned samples_per_profile,
                                                 unsigned profiles_per_frame, unsigned frames, gt_stream_t stream);

/* compute_radial_dcw_golden_ratio_2d / _fixed_angle_2d (radial_utilities.h:30-38):
 * the area of each sample's Voronoi-like cell among its neighbours on the same
 * and on the nearest other profiles; the centre sample gets pi (alpha /
 * (2 os))^2 / P.  dcw: S * P floats (written).  S even, P >= 4. */
gt_status gt_radial_dcw_golden_ratio_2d(float* dcw, unsigned samples_per_profile, unsigned num_profiles, float alpha,
                                        float one_over_radial_oversampling_factor, unsigned profile_offset, int mode,
                                        gt_stream_t stream);
gt_status gt_radial_dcw_fixed_angle_2d(float* dcw, unsigned samples_per_profile, unsigned num_profiles, float alpha,
                                       float one_over_radial_oversampling_factor, gt_stream_t stream);

/* 3D kooshball trajectory.  No counterpart in the reference (it has no 3D
 * radial generator): spoke k along a generalised spiral direction on the
 * sphere (cos theta = -1 + (2k + 1) / n, longitude steps 3.6 / sqrt(n (1 -
 * cos^2 theta))), rotated by `rotation` (row-major 3 x 3, or NULL for none);
 * sample i at (i - S/2) / S along it, clipped to [-1/2, 1/2].  traj: S * n
 * float3 (written).  Synchronises `stream`. */
gt_status gt_kooshball_trajectory_3d(float* traj, unsigned samples_per_spoke, unsigned spokes, const float* rotation,
                                     gt_stream_t stream);

/* estimate_dcw (SDC.h:25-52, gpu/cuSDC.cu:39-85): starting from initial_dcw
 * (device, n floats, or NULL for ones), `iterations` updates dcw <- dcw /
 * C2NC(NC2C(dcw)) (0 where the sum is 0) with the Jinc kernel of width
 * kernel_width, on a grid of matrix_size * os_factor (truncated) rounded up to
 * multiples of 32 per dimension (the reference rounds to its GPUs' warp size).
 * traj: device n * D floats in [-1/2, 1/2] (one flattened frame); dcw: device n
 * floats (written).  The reference's defaults: os_factor 2.1 (SDC.h:22-27;
 * 1.5 in the cuNDArray overload, cuSDC.h:20-24), 10 iterations, width 5.5.
 * Synchronises `stream`. */
gt_status gt_sdc_estimate_dcw(const float* traj, size_t n, int D, const size_t* matrix_size, float os_factor,
                              unsigned iterations, float kernel_width, const float* initial_dcw, float* dcw,
                              gt_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* GT_RADIAL_H */
