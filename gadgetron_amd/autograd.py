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
"""
import torch

from .nfft import NFFT_comp_mode, NFFTPlan


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
