/*
 * gt_cgsense.h -- C ABI of the non-Cartesian SENSE operator and its CG solver
 * (libgtnufft.so), the callers of the NFFT on the reconstruction path.
 *
 * It replaces cuNonCartesianSenseOperator<float,D>
 * (toolboxes/mri/pmri/gpu/cuNonCartesianSenseOperator.cu:14-90, with the coil
 * kernels of sense_utilities.cu:7-123) and cuCgSolver / cgSolver
 * (toolboxes/solvers/cgSolver.h:105-390, relativeResidualTCB
 * cgCallback.h:56-104, cgPreconditioner.h:24-39).
 *
 *   E x    = NFFT_C2NC(dcw; csm .* x)              x: [N..., F], y: [M, F, C]
 *   E^H y  = sum_c conj(csm_c) .* NFFT_NC2C(dcw; y)
 *   CG     : x_0 = 0 (or a given x0), r = p = w_E E^H d, then the reference
 *            recurrence with alpha = rq / <p, q> (complex), <a, b> = sum conj(a) b,
 *            the preconditioner applied twice, identity (kappa) and image
 *            (imageOperator) regularisers and termination after the iteration
 *            where rq / rq_0 < tolerance.
 *
 * Arithmetic is complex FP32; the dot products and norms that set the step
 * sizes are reduced in FP64 on the device, in a fixed order (results repeat
 * bit for bit).  As in the reference, the caller passes sqrt(DCW) as dcw
 * (NFFT applies it in both directions).  Arrays are caller-owned device memory,
 * x fastest.
 */
#ifndef GT_CGSENSE_H
#define GT_CGSENSE_H

#include "gt_nufft.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gt_sense_op gt_sense_op;

/* cuNonCartesianSenseOperator::setup + set_csm + set_dcw
 * (cuNonCartesianSenseOperator.cu:64-90, cuSenseOperator.h).  plan: a
 * preprocessed 2D or 3D plan (not owned; it must outlive the operator);
 * csm: [N..., C] complex64; dcw: NULL or real, length M*F (or a divisor, as
 * NFFT_plan::compute broadcasts it). */
gt_status gt_sense_create(gt_sense_op** op, gt_nufft_plan* plan, const void* csm, size_t ncoils,
                          const float* dcw, size_t dcw_len);
gt_status gt_sense_destroy(gt_sense_op* op);

/* mult_M: [N..., F] -> [M, F, C]; mult_MH: [M, F, C] -> [N..., F];
 * mult_MH_M: [N..., F] -> [N..., F] (linearOperator::mult_MH_M,
 * operators/linearOperator.h:75-85).  accumulate adds to out. */
gt_status gt_sense_mult_M(gt_sense_op* op, const void* in, void* out, int accumulate, gt_stream_t stream);
gt_status gt_sense_mult_MH(gt_sense_op* op, const void* in, void* out, int accumulate, gt_stream_t stream);
gt_status gt_sense_mult_MH_M(gt_sense_op* op, const void* in, void* out, int accumulate, gt_stream_t stream);

/* An image regularisation operator in the CG sum (cgSolver::add_regularization_operator,
 * cgSolver.h:376-384): weight * R^H R x with R = imageOperator's real image w
 * (imageOperator::mult_MH_M, operators/imageOperator.h:49-66: x .* w .* w).
 * image: [N..., F] floats, broadcast by modulus; NULL is the identity. */
typedef struct gt_cg_reg {
    double weight;        /* linearOperator::set_weight of the regulariser */
    const float* image;
    size_t image_len;
} gt_cg_reg;

typedef struct gt_cg_desc {
    int iterations;        /* cgSolver::set_max_iterations */
    double tc_tolerance;   /* relativeResidualTCB: stop after rq / rq_0 < tc_tolerance */
    double encoding_weight;/* linearOperator weight of E (1.0 in the reference default) */
    double kappa;          /* weight of an identity regularisation operator (0: none) */
    const void* precond;   /* NULL or complex64 diagonal weights, broadcast by modulus
                              (cgPreconditioner::apply) */
    size_t precond_len;
    const gt_cg_reg* regs; /* NULL or n_regs image regularisers, summed after E^H E in
                              list order (cgSolver::mult_MH_M, cgSolver.h:355-390) */
    int n_regs;
    const void* x0;        /* NULL or the initial guess [N..., F] complex64 (solver::set_x0;
                              cgSolver::initialize, cgSolver.h:240-280): x = x0,
                              r = rhs - A x0, p = P^2 r; rq_0 stays that of the
                              right-hand side, as in the reference */
} gt_cg_desc;

/* cgSolver::solve(d): x = CG(A, w_E E^H d) with A = w_E E^H E + kappa I +
 * sum_k weight_k diag(image_k)^2.  data: [M, F, C]; x: [N..., F] (written).
 * iterations == 0 returns the right-hand side w_E E^H d, as
 * cgSolver::solve_from_rhs does (cgSolver.h:124-129).  rq_history: NULL or
 * iterations + 1 doubles receiving rq_0 and rq after each iteration (0 after
 * termination).  The iterations run without host synchronisation. */
gt_status gt_cgsense_solve(gt_sense_op* op, const void* data, void* x, const gt_cg_desc* desc,
                           double* rq_history, gt_stream_t stream);

/* The same solver with the right-hand side given as an image: x = CG(w_E E^H E
 * + kappa I, rhs), rhs and x: [N..., F].  No counterpart in the reference: the
 * backward pass of the differentiable solve needs it, since the gradient of
 * x = A^-1 w_E E^H d with respect to d is w_E E A^-1 g (A is Hermitian), one
 * more solve with the image-space gradient g as right-hand side. */
gt_status gt_cgsense_solve_image(gt_sense_op* op, const void* rhs, void* x, const gt_cg_desc* desc,
                                 double* rq_history, gt_stream_t stream);

/* Many independent reconstructions (uneven sizes, one operator each), set up
 * once and run any number of times (the operators, data and x buffers are
 * bound at creation; every problem needs its own operator, plan and x buffer).
 *   GT_CG_PLAIN:   one problem at a time, each iteration's dot products read
 *                  back to the host, as the reference's cgSolver does;
 *   GT_CG_BATCHED: every problem's whole solve (all iterations) captured into
 *                  one hipGraph with the problems spread over parallel
 *                  branches (longest first), launched as one graph with no
 *                  host round trips;
 *   GT_CG_FUSED:   every step of every problem in one launch, from
 *                  per-problem descriptor tables: one gridding launch per
 *                  direction over the work items of all problems, the FFT
 *                  passes once per group of problems with equal matrix sizes,
 *                  the CG updates with a problem index; all iterations
 *                  captured into one hipGraph.  2D float plans with fused FFT
 *                  lengths and the matrix-core gridding, an identity
 *                  regulariser (kappa) only.  It holds one oversampled grid
 *                  per size group for all its problems' coils.
 * PLAIN and BATCHED compute the same arithmetic (results agree bit for bit);
 * FUSED runs every problem's gridding on the matrix cores (the others do so
 * from 9 / 17 coils), so its results agree with theirs to rounding. */
enum { GT_CG_PLAIN = 0, GT_CG_BATCHED = 1, GT_CG_FUSED = 2 };
typedef struct gt_cg_batch gt_cg_batch;
/* gt_cgsense_batch_create (GT_CG_BATCHED) times candidate graphs by running
 * them: it writes every problem's x (whole solves), reads the data and runs
 * one E^H E per operator to size its workspaces.  After creation the caller
 * may change the contents of data (and x, which each run overwrites) but not
 * the bound plans (no preprocess or set_tiling) nor the operators' sizes:
 * gt_cgsense_batch_run then returns GT_ERR_INVALID_ARGUMENT without running
 * anything, and the batch must be created again.  desc->x0 is not taken by a
 * batch (GT_ERR_UNSUPPORTED). */
gt_status gt_cgsense_batch_create(gt_cg_batch** batch, gt_sense_op* const* ops, const void* const* data,
                                  void* const* x, int nprob, const gt_cg_desc* desc, int mode, gt_stream_t stream);
gt_status gt_cgsense_batch_run(gt_cg_batch* batch, gt_stream_t stream);
gt_status gt_cgsense_batch_destroy(gt_cg_batch* batch);
/* Parallel branches of a batched graph, picked at creation by timing one
 * replay per candidate (2, 4, ..., 64, and one per problem); 0 for GT_CG_PLAIN. */
gt_status gt_cgsense_batch_branches(const gt_cg_batch* batch, int* branches);
/* GT_CG_FUSED: NC2C table entries per work item its autotuner chose for every
 * bound plan (0: each plan's own rule; -1 for the other modes).  Creation tries
 * 0, 256, 512, 1024 and 2048 (re-tiling the bound plans' NC2C work items,
 * which changes only the summation order of split tiles) and keeps the
 * fastest; the choice is cached per problem set for the process and, where
 * GT_NUFFT_TUNE_CACHE names a file, across processes. */
gt_status gt_cgsense_batch_item_length(const gt_cg_batch* batch, int* entries);

/* cuImageOperator<T>::compute (operators/imageOperator.h:28-47 with the
 * histogram offset of operators/gpu/cuImageOperator.h:30-63): the real
 * regularisation image 1 / (|x| s + offset) of a complex image x (n elements),
 * s = n / asum(x) (asum = sum |re| + |im|, the BLAS definition).  With
 * offset_estimation the offset is (k + 1) max / 50000, k the number of
 * 50000-bin histogram bins of |x| s / max up to and including the bin that
 * brings the count to floor(0.01 n); otherwise *offset is used as given.  The
 * offset used is written back to *offset.  out: n floats. */
gt_status gt_image_operator_compute(const void* image, size_t n, int offset_estimation, float* offset, float* out,
                                    gt_stream_t stream);

/* The CG-SENSE preconditioner weights of gpuCgSenseGadget::process
 * (gadgets/pmri/gpuCgSenseGadget.cpp:165-176): w = 1 / sqrt(sum_c |csm_c|^2 +
 * kappa reg) over the nimg image elements (reg NULL: 0), written as complex64
 * (imaginary part 0).  csm: [N..., C] complex64; reg: nimg floats. */
gt_status gt_cgsense_precond_weights(const void* csm, size_t nimg, size_t ncoils, const float* reg, double kappa,
                                     void* weights, gt_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* GT_CGSENSE_H */
