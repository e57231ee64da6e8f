/*
 * gt_nufft.h -- C ABI of the MI355X non-Cartesian NFFT library (libgtnufft.so).
 *
 * It replaces Gadgetron's cuNFFT_plan<REAL,D> C++ template API
 * (toolboxes/nfft/NFFT.h:54-175, toolboxes/nfft/gpu/cuNFFT.h:32-131) with plain
 * C entry points: opaque plan handles, raw device pointers, explicit dims
 * (x fastest, as cuNDArray::get_dimensions() reports them), status codes
 * instead of exceptions, and an explicit HIP stream (NULL = default stream).
 *
 * Layouts (Gadgetron's, column-major, x fastest):
 *   image       [N_0 .. N_{D-1}, F, C]  or oversampled [N_os..., F, C], complex64
 *   samples     [M, F, C]                complex64
 *   trajectory  [M, F] of D floats (x, y[, z] interleaved), values in [-1/2, 1/2]
 *   dcw         real, broadcast over samples by modulus of its length
 * F = frames (set by preprocess), C = batches (coils); any extra image dims
 * beyond D fold into F*C.
 *
 * Every call validates its arguments on the host before any device work, so
 * the error paths run without a GPU.  A failing call returns a non-zero
 * gt_status and leaves a thread-local message for gt_last_error().
 */
#ifndef GT_NUFFT_H
#define GT_NUFFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gt_status {
    GT_OK = 0,
    GT_ERR_INVALID_ARGUMENT = 1, /* std::runtime_error validation paths of the reference */
    GT_ERR_NOT_PREPROCESSED = 2,
    GT_ERR_SHAPE = 3,            /* dims mismatch (cuGriddingConvolution.cu:213-268)    */
    GT_ERR_RANGE = 4,            /* trajectory outside [-1/2,1/2] (cuGriddingConvolution.cu:145-152) */
    GT_ERR_UNSUPPORTED = 5,
    GT_ERR_DEVICE = 6,           /* HIP runtime failure (the reference's cuda_error)   */
    GT_ERR_FFT = 7,              /* rocFFT failure                                     */
    GT_ERR_OUT_OF_MEMORY = 8,
    GT_ERR_INTERNAL = 9
} gt_status;

/* NFFT_fft_mode (NFFT.h:14-17) */
enum { GT_FFT_FORWARDS = 0, GT_FFT_BACKWARDS = 1 };
/* NFFT_conv_mode (NFFT.h:22-25) */
enum { GT_CONV_C2NC = 0, GT_CONV_NC2C = 1 };
/* NFFT_comp_mode (NFFT.h:31-36) */
enum { GT_FORWARDS_C2NC = 0, GT_FORWARDS_NC2C = 1, GT_BACKWARDS_C2NC = 2, GT_BACKWARDS_NC2C = 3 };
/* NFFT_prep_mode (NFFT.h:41-45) */
enum { GT_PREP_C2NC = 0, GT_PREP_NC2C = 1, GT_PREP_ALL = 2 };
/* ConvolutionType (gpu/cuGriddingConvolution.h:13-18).  STANDARD is the
 * deterministic tile gather; ATOMIC is the float-atomic scatter kept for
 * comparison; SPARSE_MATRIX grids through the transpose of the row-normalised
 * convolution matrix (ConvolverNC2C_sparse.cuh:84-111): each sample is scaled
 * by one over the sum of its window weights, then gridded as STANDARD (its
 * C2NC is STANDARD's, as in the reference). */
enum { GT_CONV_STANDARD = 0, GT_CONV_ATOMIC = 1, GT_CONV_SPARSE_MATRIX = 2 };
/* precision */
enum { GT_FLOAT = 0, GT_DOUBLE = 1 };

typedef void* gt_stream_t; /* a hipStream_t; NULL selects the default stream */
typedef struct gt_nufft_plan gt_nufft_plan;

/* Plan descriptor.  One of N_os[] or os_factor selects the ctor:
 *   os_factor == 0  -> NFFT_plan(N, N_os, W)              (NFFT.hpp:6-20, cuNFFT.cu:28-47)
 *   os_factor >= 1  -> NFFT_plan(N, os_factor, W)         (NFFT.hpp:23-37)
 * Both use the cuNFFT deapodization convention (unscaled filter FFT,
 * cuNFFT.cu:170-202), which equals hoNFFT_plan(N, alpha, W) (hoNFFT.cpp:175-197). */
typedef struct gt_nufft_desc {
    int D;               /* 1 to 4 (4: no trajectory gradient, no SENSE) */
    int precision;       /* GT_FLOAT, or GT_DOUBLE (NFFT<cuNDArray, double, D>, cuNFFT.cu:219-227):
                            STANDARD convolution only, the generic kernels in double;
                            trajectories are double D-vectors, samples and images
                            complex double, and a dcw argument points to doubles
                            (passed through the `const float*` parameter) */
    size_t N[4];
    size_t N_os[4];
    double os_factor;
    double W;            /* kernel width, >= 1.8 (cuGriddingConvolution.cu:67-69) */
    int conv_type;       /* GT_CONV_STANDARD, GT_CONV_ATOMIC or GT_CONV_SPARSE_MATRIX */
    int device;          /* -1: current device */
} gt_nufft_desc;

/* Plan life cycle.  Replaces NFFT<cuNDArray,REAL,D>::make_plan (cuNFFT.cu:206-227). */
gt_status gt_nufft_create(gt_nufft_plan** plan, const gt_nufft_desc* desc);
gt_status gt_nufft_destroy(gt_nufft_plan* plan);

/* NFFT_plan::preprocess (NFFT.hpp:40-62, cuGriddingConvolution.cu:128-188).
 * traj: device pointer to num_samples*num_frames D-vectors of float. */
gt_status gt_nufft_preprocess(gt_nufft_plan* plan, const void* traj, size_t num_samples,
                              size_t num_frames, int prep_mode, gt_stream_t stream);

/* NFFT_plan::compute (NFFT.hpp:104-202).  dcw may be NULL. */
gt_status gt_nufft_compute(gt_nufft_plan* plan, const void* in, const size_t* in_dims, int in_ndim,
                           void* out, const size_t* out_dims, int out_ndim, const float* dcw,
                           size_t dcw_len, int comp_mode, gt_stream_t stream);

/* gt_nufft_compute that allocates no device memory.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as gt_nufft_compute, but the oversampled grid
 * and the split tiles' partial sums live in a fixed device array of the
 * library, 64 MiB, of which each GPU holds its own copy from the library's
 * first use on it until the process ends.  The call needs nb * prod(N_os) * 8
 * bytes for the grid (nb = frames * coils) plus, for the NC2C modes, the
 * partial tiles, together at most 64 MiB (the 256^2, 32-coil workload: 36 MiB
 * of grid); a larger call returns GT_ERR_UNSUPPORTED before any device work.
 * Taken: float 2D STANDARD plans whose oversampled sizes are fused FFT lengths
 * (96, 128, 144, 160, 192, 240, 256, 288, 320, 384), image side of the matrix
 * size; anything else returns GT_ERR_UNSUPPORTED.  Nothing in the array needs
 * to start at zero (every call writes what it reads).  Every call on a GPU
 * shares that GPU's array: calls that overlap in time on one GPU (several
 * streams or threads) race, and the caller must order them.  Preprocess, which
 * allocates, must have run before. */
gt_status gt_nufft_compute_noalloc(gt_nufft_plan* plan, const void* in, const size_t* in_dims, int in_ndim,
                                   void* out, const size_t* out_dims, int out_ndim, const float* dcw,
                                   size_t dcw_len, int comp_mode, gt_stream_t stream);

/* NFFT_plan::mult_MH_M (NFFT.hpp:89-101): FORWARDS_C2NC then BACKWARDS_NC2C. */
gt_status gt_nufft_mult_MH_M(gt_nufft_plan* plan, const void* in, void* out, const size_t* dims,
                             int ndim, const float* dcw, size_t dcw_len, gt_stream_t stream);

/* NFFT_plan::convolve (NFFT.hpp:65-86); the image side is oversampled. */
gt_status gt_nufft_convolve(gt_nufft_plan* plan, const void* in, const size_t* in_dims, int in_ndim,
                            void* out, const size_t* out_dims, int out_ndim, int conv_mode,
                            int accumulate, gt_stream_t stream);

/* cuNFFT_impl::fft (cuNFFT.cu:57-70 -> cuNDFFT): centred FFT over the first D
 * dims of an oversampled array, scaled by 1/sqrt(prod N_os) when do_scale. */
gt_status gt_nufft_fft(gt_nufft_plan* plan, void* data, const size_t* dims, int ndim, int fft_mode,
                       int do_scale, gt_stream_t stream);

/* cuNFFT_impl::deapodize (cuNFFT.cu:73-92); image must be oversampled. */
gt_status gt_nufft_deapodize(gt_nufft_plan* plan, void* image, const size_t* dims, int ndim,
                             int fourier_domain, gt_stream_t stream);

/* Trajectory gradient of the differentiable forward NUFFT (no counterpart in
 * the reference: the derivative-of-Kaiser-Bessel kernel behind the torch
 * autograd function).  For y = compute(image, FORWARDS_C2NC, dcw) and a real
 * loss L with grad_samples = 2 dL/d conj(y) ([M, F, C], torch's convention),
 * writes grad_traj = dL/dk, [M, F] of D floats (the trajectory's layout).  The
 * oversampled grid of the forward pass is recomputed from image. */
gt_status gt_nufft_grad_traj(gt_nufft_plan* plan, const void* image, const size_t* dims, int ndim,
                             const void* grad_samples, const size_t* sdims, int sndim, const float* dcw,
                             size_t dcw_len, float* grad_traj, gt_stream_t stream);

/* NFFT_plan getters (NFFT.h:135-151) and preprocess results. */
gt_status gt_nufft_get_matrix_size(const gt_nufft_plan* plan, size_t* N);
gt_status gt_nufft_get_matrix_size_os(const gt_nufft_plan* plan, size_t* N_os);
gt_status gt_nufft_get_W(const gt_nufft_plan* plan, double* W);
gt_status gt_nufft_get_num_samples_frames(const gt_nufft_plan* plan, size_t* M, size_t* F);

/* Tuning knobs of the STANDARD gridding (tile edge of dims 0..2, samples per
 * work item).  Values <= 0 keep the defaults; a 4-D plan's 4th dimension keeps
 * tiles as wide as the window.  Takes effect at the next preprocess. */
gt_status gt_nufft_set_tiling(gt_nufft_plan* plan, const int* tile, int max_samples_per_item);

/* Stage timing with HIP events recorded on the caller's stream around each
 * stage of compute/convolve/fft (the role of GPUTimer, core/gpu/GPUTimer.h:16-67).
 * Stages: 0 NC2C gridding (incl. the split-tile reduction), 1 C2NC gridding,
 * 2 FFT, 3 pad+deapodize, 4 crop+deapodize.  enable=0 stops recording;
 * enabling creates a pool of 1024 events up front, so recording creates none
 * until that many are outstanding.  gt_nufft_profile_read synchronizes,
 * returns total milliseconds and launch counts per stage (n entries) and
 * clears the record. */
enum { GT_STAGE_NC2C = 0, GT_STAGE_C2NC = 1, GT_STAGE_FFT = 2, GT_STAGE_PAD = 3, GT_STAGE_CROP = 4,
       GT_NUM_STAGES = 5 };
gt_status gt_nufft_profile(gt_nufft_plan* plan, int enable);
/* As gt_nufft_profile, recording only the stages whose bit (1 << stage) is
 * set in stage_mask (0 stops recording).  Every recorded event waits for the
 * work before it, so a timed loop that records only the stages it needs keeps
 * the others at their unrecorded speed. */
gt_status gt_nufft_profile_stages(gt_nufft_plan* plan, unsigned stage_mask);
gt_status gt_nufft_profile_read(gt_nufft_plan* plan, double* total_ms, int* counts, int n);

/* Thread-local message of the last failing call on this thread. */
const char* gt_last_error(void);

/* Device-memory plumbing for callers without their own allocator (tests,
 * ctypes).  Thin wrappers over the HIP runtime. */
gt_status gt_device_count(int* count);
gt_status gt_set_device(int device);
gt_status gt_malloc(void** ptr, size_t bytes);
gt_status gt_free(void* ptr);
gt_status gt_memcpy_htod(void* dst, const void* src, size_t bytes, gt_stream_t stream);
gt_status gt_memcpy_dtoh(void* dst, const void* src, size_t bytes, gt_stream_t stream);
gt_status gt_memcpy_dtod(void* dst, const void* src, size_t bytes, gt_stream_t stream);
gt_status gt_memset(void* ptr, int value, size_t bytes, gt_stream_t stream);
/* y += x over n real components (GT_FLOAT: float, GT_DOUBLE: double; complex
 * arrays pass twice their element count): cuNDArray operator+=
 * (core/gpu/cuNDArray_operators.cu:85-88), which NFFTOperator's accumulating
 * applications use (NFFTOperator.hpp:22-33). */
gt_status gt_array_add(void* y, const void* x, size_t n, int precision, gt_stream_t stream);
/* y = sqrt(y) over n floats: sqrt_inplace (core/gpu/cuNDArray_elemwise.h:64),
 * which gpuCgSenseGadget applies to the density weights of a job
 * (gadgets/pmri/gpuCgSenseGadget.cpp:129). */
gt_status gt_array_sqrt(float* y, size_t n, gt_stream_t stream);
/* y *= a over n floats: cuNDArray operator*= with a scalar
 * (core/gpu/cuNDArray_operators.h:112), as on the regulariser image
 * (gpuCgSenseGadget.cpp:170). */
gt_status gt_array_scale(float* y, size_t n, float a, gt_stream_t stream);
/* x[b][i] *= w[i] for complex64 x [batches][n] and float w [n]: cuNDArray
 * operator*= of a complex array by a real array of its leading dimensions,
 * batched over the rest (core/gpu/cuNDArray_operators.h:114-123), as on a
 * job's samples (gpuCgSenseGadget.cpp:179). */
gt_status gt_array_mul_weights(void* x, const float* w, size_t n, size_t batches, gt_stream_t stream);
gt_status gt_stream_synchronize(gt_stream_t stream);
gt_status gt_device_synchronize(void);

#ifdef __cplusplus
}
#endif

#endif /* GT_NUFFT_H */
