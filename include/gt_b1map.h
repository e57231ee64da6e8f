/* gt_b1map.h -- coil sensitivity (B1) map estimation for libgtnufft: the
 * coil maps CG-SENSE takes, from accumulated coil images (gt_buffer.h).
 *
 * Replaces (reference files, toolboxes/mri/pmri/gpu/):
 *   b1_map.h:16-24, b1_map.cu:32-108    estimate_b1_map<float, D>, D = 2, 3
 */
#ifndef GT_B1MAP_H
#define GT_B1MAP_H

#include "gt_nufft.h"

#ifdef __cplusplus
extern "C" {
#endif

/* estimate_b1_map (b1_map.cu:32-108): coil images normalised by their
 * root-sum-of-squares, the coil correlation matrices box-smoothed over 7^D
 * pixels (zero outside the image), two power iterations from a vector of ones
 * for the dominant eigenvector, and the phase of coil 0 removed.
 *   data: device [N..., num_coils] complex64 (x fastest, coil slowest);
 *   D: 2 or 3; image_dims: D sizes;
 *   target_coils: use the first target_coils coils (<= 0 or > num_coils: all);
 *   csm: device [N..., T] complex64 (written), T the number of coils used.
 * A pixel whose coils are all zero makes its 7^D neighbourhood NaN, as in the
 * reference.  Scratch is allocated per call; synchronises `stream`. */
gt_status gt_estimate_b1_map(const void* data, int D, const size_t* image_dims, unsigned num_coils, int target_coils,
                             void* csm, gt_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* GT_B1MAP_H */
