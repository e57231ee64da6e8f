"""Differentiable forward NUFFT for torch: y = A(x; k).

A is NFFT_plan::compute(FORWARDS_C2NC) (toolboxes/nfft/NFFT.hpp:125-140) of
libgtnufft.so.  Its backward pass uses the library's own kernels:
  * image gradient: the adjoint A^H (BACKWARDS_NC2C, NFFT.hpp:163-184);
  * trajectory gradient: the derivative-of-Kaiser-Bessel gather
    (gt_nufft_grad_traj), with the forward pass's oversampled grid recomputed
    from x rather than stored.
Gradients follow torch's convention for complex tensors (a real loss L gives
grad = 2 dL/d conj(z) for complex z and dL/dk for the real trajectory).

Shapes (numpy / torch order): x (C, F, [Nz,] Ny, Nx) complex64, k (F, M, D)
float32 in [-1/2, 1/2], y (C, F, M) complex64, dcw None or (F, M) float32.
All tensors live on the GPU and are contiguous.

cg_sense_solve makes the CG-SENSE reconstruction x = A^-1 w_E E^H d,
A = w_E E^H E + kappa I (cgSolver::solve, toolboxes/solvers/cgSolver.h:105-390),
differentiable in the data d.  A is Hermitian, so the gradient is
w_E E A^-1 g for the image gradient g: the backward pass runs one more CG
solve (right-hand side g) and one E, recomputing rather than storing anything
of the forward iterations -- memory does not grow with the iteration count.
It is the gradient of the exact solution, which the forward iterate
approaches as the solve converges.
"""
import torch

from .nfft import NFFT_comp_mode, NFFTPlan
from .sense import SenseOperator, cg_sense, cg_sense_image


def _stream():
    return torch.cuda.current_stream().cuda_stream


class NufftForward(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, k, N, W, os_factor, dcw):
        x = x.contiguous()
        k = k.detach().contiguous()
        plan = NFFTPlan(list(N), None, W=W, os_factor=os_factor)
        s = _stream()
        plan.preprocess(k, stream=s)
        C, F = x.shape[0], x.shape[1]
        y = torch.empty((C, F, k.shape[-2]), dtype=torch.complex64, device=x.device)
        plan.compute(x, y, dcw, NFFT_comp_mode.FORWARDS_C2NC, stream=s)
        ctx.plan = plan
        ctx.dcw = dcw
        ctx.save_for_backward(x, k)
        return y

    @staticmethod
    def backward(ctx, gy):
        x, k = ctx.saved_tensors
        plan, dcw = ctx.plan, ctx.dcw
        gy = gy.contiguous()
        s = _stream()
        gx = gk = None
        if ctx.needs_input_grad[0]:
            gx = torch.empty_like(x)
            plan.compute(gy, gx, dcw, NFFT_comp_mode.BACKWARDS_NC2C, stream=s)
        if ctx.needs_input_grad[1]:
            gk = torch.empty_like(k)
            plan.grad_traj(x, gy, gk, dcw, stream=s)
        return gx, gk, None, None, None, None


def nufft(x, k, N=None, W=5.5, os_factor=1.5, dcw=None):
    """y = A(x; k), differentiable in x and k."""
    if N is None:
        N = list(reversed(x.shape[2:]))
    return NufftForward.apply(x, k, tuple(int(n) for n in N), float(W), float(os_factor), dcw)


class CgSenseSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, data, op, image_shape, iterations, tc_tolerance, encoding_weight, kappa):
        data = data.contiguous()
        s = _stream()
        x = torch.zeros(image_shape, dtype=torch.complex64, device=data.device)
        cg_sense(op, data, x, iterations, tc_tolerance, encoding_weight, kappa, stream=s)
        ctx.op = op
        ctx.args = (iterations, tc_tolerance, encoding_weight, kappa)
        ctx.data_shape = data.shape
        return x

    @staticmethod
    def backward(ctx, gx):
        op = ctx.op
        iterations, tc_tolerance, encoding_weight, kappa = ctx.args
        gd = None
        if ctx.needs_input_grad[0]:
            s = _stream()
            gx = gx.contiguous()
            z = torch.zeros_like(gx)
            cg_sense_image(op, gx, z, iterations, tc_tolerance, encoding_weight, kappa, stream=s)
            gd = torch.empty(ctx.data_shape, dtype=torch.complex64, device=gx.device)
            op.mult_M(z, gd, stream=s)
            if encoding_weight != 1.0:
                gd = gd * encoding_weight
        return gd, None, None, None, None, None, None


def cg_sense_solve(op: SenseOperator, data, iterations=10, tc_tolerance=1e-6, encoding_weight=1.0, kappa=0.0):
    """x = CG(w_E E^H E + kappa I, w_E E^H d) for data d (C, F, M), differentiable in d.
    Returns x with shape (F, [Nz,] Ny, Nx)."""
    N = op.plan.get_matrix_size()
    shape = (int(data.shape[1]),) + tuple(int(n) for n in reversed(N))
    return CgSenseSolve.apply(data, op, shape, int(iterations), float(tc_tolerance), float(encoding_weight),
                              float(kappa))
