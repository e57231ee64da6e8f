/* gt_buffer.h -- accumulation of k-space frames into coil images for
 * libgtnufft: the step before CG-SENSE (its right-hand side and coil maps).
 *
 * Replaces (reference files, toolboxes/mri/pmri/gpu/):
 *   cuBuffer.h:13-57, cuBuffer.cpp:8-190       cuBuffer<float, D>
 *   cuSenseBuffer.h, cuSenseBuffer.cpp:6-38      cuSenseBuffer<float, D>::get_combined_coil_image
 * A buffer holds `num_cycles` + 1 oversampled coil k-space slots.  Frames are
 * gridded (NC2C, accumulating) into the current slot; after `num_sub_cycles`
 * frames the slot joins the running sum, the oldest slot leaves it and is
 * cleared, and the next slot becomes current.  Device pointers are caller
 * owned; image arrays are [N..., C] complex64, x fastest.
 */
#ifndef GT_BUFFER_H
#define GT_BUFFER_H

#include "gt_nufft.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gt_buffer gt_buffer;

/* cuBuffer::cuBuffer (cuBuffer.cpp:8-20) */
gt_status gt_buffer_create(gt_buffer** buffer);
gt_status gt_buffer_destroy(gt_buffer* buffer);

/* cuBuffer::setup (cuBuffer.cpp:32-77): D 1..3, matrix_size / matrix_size_os
 * of D entries, kernel width W, coils, full cycles and frames per cycle.  The
 * plan is made again only when the sizes or W change, and the slots only when
 * the oversampled size or coil count change (their contents are kept
 * otherwise). */
gt_status gt_buffer_setup(gt_buffer* buffer, int D, const size_t* matrix_size, const size_t* matrix_size_os,
                          float W, unsigned num_coils, unsigned num_cycles, unsigned num_sub_cycles);

/* cuBuffer::set_dcw (cuBuffer.h:23-25): device pointer kept by reference (NULL
 * for none); broadcast over coils by modulus. */
gt_status gt_buffer_set_dcw(gt_buffer* buffer, const float* dcw, size_t dcw_len);

/* cuBuffer::clear (cuBuffer.cpp:22-30) */
gt_status gt_buffer_clear(gt_buffer* buffer, gt_stream_t stream);

/* cuBuffer::add_frame_data (cuBuffer.cpp:79-135): samples [M, C] complex64 and
 * their trajectory [M] of D floats; *cycle_completed (may be NULL) is set to 1
 * when the frame completed a slot (the accumulated images changed). */
gt_status gt_buffer_add_frame_data(gt_buffer* buffer, const void* samples, size_t num_samples, const float* traj,
                                   int* cycle_completed, gt_stream_t stream);

/* cuBuffer::get_accumulated_coil_images (cuBuffer.cpp:137-171): backwards
 * FFT, deapodization and crop of the running sum into out [N..., C]; ones
 * while no slot has completed. */
gt_status gt_buffer_get_accumulated_coil_images(gt_buffer* buffer, void* out, gt_stream_t stream);

/* cuBuffer::get_normalization_factor (cuBuffer.h:27-29): 1 / (num_cycles * num_sub_cycles). */
gt_status gt_buffer_get_normalization_factor(const gt_buffer* buffer, float* factor);

/* cuSenseBuffer::get_combined_coil_image (cuSenseBuffer.cpp:18-38): the
 * accumulated coil images combined as sum_c conj(csm_c) image_c; csm [N..., C]
 * complex64, out [N...] complex64. */
gt_status gt_buffer_get_combined_coil_image(gt_buffer* buffer, const void* csm, void* out, gt_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* GT_BUFFER_H */
