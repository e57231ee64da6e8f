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
 *   CG     : x_0 = 0, r = p = w_E E^H d, then the reference recurrence with
 *            alpha = rq / <p, q> (complex), <a, b> = sum conj(a) b, the
 *            preconditioner applied twice, identity regularisation kappa and
 *            termination after the iteration where rq / rq_0 < tolerance.
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

typedef struct gt_cg_desc {
    int iterations;        /* cgSolver::set_max_iterations */
    double tc_tolerance;   /* relativeResidualTCB: stop after rq / rq_0 < tc_tolerance */
    double encoding_weight;/* linearOperator weight of E (1.0 in the reference default) */
    double kappa;          /* weight of an identity regularisation operator (0: none) */
    const void* precond;   /* NULL or complex64 diagonal weights, broadcast by modulus
                              (cgPreconditioner::apply) */
    size_t precond_len;
} gt_cg_desc;

/* cgSolver::solve(d): x = CG(E^H E + kappa I, w_E E^H d).  data: [M, F, C];
 * x: [N..., F] (written).  rq_history: NULL or iterations + 1 doubles receiving
 * rq_0 and rq after each iteration (0 after termination).  The iterations run
 * without host synchronisation. */
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
 * bound at creation).
 *   GT_CG_PLAIN:   one problem at a time, each iteration's dot products read
 *                  back to the host, as the reference's cgSolver does;
 *   GT_CG_BATCHED: every problem's whole solve (all iterations) captured into
 *                  one hipGraph with the problems spread over parallel
 *                  branches (longest first), launched as one graph with no
 *                  host round trips.
 * Both modes compute the same arithmetic (results agree bit for bit). */
enum { GT_CG_PLAIN = 0, GT_CG_BATCHED = 1 };
typedef struct gt_cg_batch gt_cg_batch;
gt_status gt_cgsense_batch_create(gt_cg_batch** batch, gt_sense_op* const* ops, const void* const* data,
                                  void* const* x, int nprob, const gt_cg_desc* desc, int mode, gt_stream_t stream);
gt_status gt_cgsense_batch_run(gt_cg_batch* batch, gt_stream_t stream);
gt_status gt_cgsense_batch_destroy(gt_cg_batch* batch);
/* Parallel branches of a batched graph, picked at creation by timing one
 * replay per candidate (2, 4, ..., 64, and one per problem); 0 for GT_CG_PLAIN. */
gt_status gt_cgsense_batch_branches(const gt_cg_batch* batch, int* branches);

#ifdef __cplusplus
}
#endif

#endif /* GT_CGSENSE_H */
