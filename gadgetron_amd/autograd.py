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
That is the gradient of the exact solution, which the forward iterate
approaches as the solve converges (grad="exact").  grad="unrolled" gives the
gradient of the truncated iterate the forward pass returns: the backward pass
recomputes the CG iterations (nothing of them is kept between the passes) with
the library's E^H E and E^H as linear autograd functions and differentiates
through the recurrence.

The forward NUFFT keeps a few plans, keyed on the trajectory tensor itself
(its identity, storage address, shape and version counter): repeated calls on
an unchanged trajectory -- image-only gradient loops -- skip preprocess, and a
trajectory changed in place gets a new plan.
"""
import collections
import weakref

import torch

from .nfft import NFFT_comp_mode, NFFTPlan
from .sense import SenseOperator, cg_sense, cg_sense_image


def _stream():
    return torch.cuda.current_stream().cuda_stream


# (id of the trajectory tensor, N, W, os_factor, device) ->
# (weak reference, version, data_ptr, shape, plan); least recently used first
_PLANS = collections.OrderedDict()
PLAN_CACHE_SIZE = 4


def _plan_for(k, N, W, os_factor, stream):
    """A preprocessed plan for trajectory k, reused while k is unchanged."""
    key = (id(k), tuple(N), float(W), float(os_factor), k.device.index)
    ent = _PLANS.get(key)
    if ent is not None:
        ref, ver, ptr, shape, plan = ent
        if ref() is k and ver == k._version and ptr == k.data_ptr() and shape == tuple(k.shape):
            _PLANS.move_to_end(key)
            return plan
        del _PLANS[key]
    plan = NFFTPlan(list(N), None, W=W, os_factor=os_factor)
    plan.preprocess(k.detach().contiguous(), stream=stream)
    _PLANS[key] = (weakref.ref(k), k._version, k.data_ptr(), tuple(k.shape), plan)
    while len(_PLANS) > PLAN_CACHE_SIZE:
        _PLANS.popitem(last=False)
    return plan


def clear_plan_cache():
    _PLANS.clear()


class NufftForward(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, k, N, W, os_factor, dcw):
        x = x.contiguous()
        s = _stream()
        plan = _plan_for(k, N, W, os_factor, s)
        k = k.detach()
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


def _normal(op, x, we, kappa):
    """A x = w_E E^H E x + kappa x (cgSolver::mult_MH_M with an identity regulariser)."""
    y = torch.empty_like(x)
    op.mult_MH_M(x, y, stream=_stream())
    if we != 1.0:
        y = y * we
    if kappa != 0.0:
        y = y + kappa * x
    return y


class _Normal(torch.autograd.Function):
    """y = A x with A = w_E E^H E + kappa I.  A is Hermitian, so the
    vector-Jacobian product of a gradient g is A g."""

    @staticmethod
    def forward(ctx, x, op, we, kappa):
        ctx.op, ctx.we, ctx.kappa = op, we, kappa
        return _normal(op, x.contiguous(), we, kappa)

    @staticmethod
    def backward(ctx, g):
        return _normal(ctx.op, g.contiguous(), ctx.we, ctx.kappa), None, None, None


class _Rhs(torch.autograd.Function):
    """b = w_E E^H d; the vector-Jacobian product of g is w_E E g."""

    @staticmethod
    def forward(ctx, d, op, we, image_shape):
        ctx.op, ctx.we, ctx.data_shape = op, we, d.shape
        b = torch.empty(image_shape, dtype=torch.complex64, device=d.device)
        op.mult_MH(d.contiguous(), b, stream=_stream())
        return b * we if we != 1.0 else b

    @staticmethod
    def backward(ctx, g):
        gd = torch.empty(ctx.data_shape, dtype=torch.complex64, device=g.device)
        ctx.op.mult_M(g.contiguous(), gd, stream=_stream())
        return (gd * ctx.we if ctx.we != 1.0 else gd), None, None, None


def _dot(a, b):
    """<a, b> = sum conj(a) b in complex128 (the solver's FP64 dot products)."""
    return torch.vdot(a.reshape(-1).to(torch.complex128), b.reshape(-1).to(torch.complex128))


def _cg_unrolled(b, apply_A, iterations, tc_tolerance):
    """The CG recurrence of cgSolver (cgSolver.h:219-390, no preconditioner,
    x_0 = 0, relativeResidualTCB) written in torch operations, so that autograd
    can differentiate the truncated iterate."""
    x = torch.zeros_like(b)
    r = b
    p = r
    rq0 = _dot(r, r).real
    rq = rq0
    if not float(rq0.detach()) > 0.0:
        return x
    for _ in range(int(iterations)):
        q = apply_A(p)
        alpha = (rq / _dot(p, q)).to(torch.complex64)
        x = x + alpha * p
        r = r - alpha * q
        tmp = _dot(r, r).real
        p = p * (tmp / rq).to(torch.float32) + r
        rq = tmp
        if float((rq / rq0).detach()) < tc_tolerance:
            break
    return x


class CgSenseSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, data, op, image_shape, iterations, tc_tolerance, encoding_weight, kappa, grad):
        data = data.contiguous()
        s = _stream()
        x = torch.zeros(image_shape, dtype=torch.complex64, device=data.device)
        cg_sense(op, data, x, iterations, tc_tolerance, encoding_weight, kappa, stream=s)
        ctx.op = op
        ctx.args = (iterations, tc_tolerance, encoding_weight, kappa)
        ctx.image_shape = image_shape
        ctx.data_shape = data.shape
        ctx.grad = grad
        if grad == "unrolled":
            ctx.save_for_backward(data)
        return x

    @staticmethod
    def backward(ctx, gx):
        op = ctx.op
        iterations, tc_tolerance, encoding_weight, kappa = ctx.args
        gd = None
        if not ctx.needs_input_grad[0]:
            return None, None, None, None, None, None, None, None
        s = _stream()
        gx = gx.contiguous()
        if ctx.grad == "unrolled":
            (data,) = ctx.saved_tensors
            with torch.enable_grad():
                d = data.detach().requires_grad_()
                b = _Rhs.apply(d, op, encoding_weight, ctx.image_shape)
                x = _cg_unrolled(b, lambda v: _Normal.apply(v, op, encoding_weight, kappa), iterations,
                                 tc_tolerance)
                (gd,) = torch.autograd.grad(x, d, gx)
            return gd, None, None, None, None, None, None, None
        z = torch.zeros_like(gx)
        cg_sense_image(op, gx, z, iterations, tc_tolerance, encoding_weight, kappa, stream=s)
        gd = torch.empty(ctx.data_shape, dtype=torch.complex64, device=gx.device)
        op.mult_M(z, gd, stream=s)
        if encoding_weight != 1.0:
            gd = gd * encoding_weight
        return gd, None, None, None, None, None, None, None


def cg_sense_solve(op: SenseOperator, data, iterations=10, tc_tolerance=1e-6, encoding_weight=1.0, kappa=0.0,
                   grad="exact"):
    """x = CG(w_E E^H E + kappa I, w_E E^H d) for data d (C, F, M), differentiable in d.
    Returns x with shape (F, [Nz,] Ny, Nx).

    grad="exact": the gradient of the exact solution map (one more CG solve in
    the backward pass); grad="unrolled": the gradient of the truncated iterate,
    with the iterations recomputed in the backward pass."""
    if grad not in ("exact", "unrolled"):
        raise ValueError("grad must be 'exact' or 'unrolled'")
    N = op.plan.get_matrix_size()
    shape = (int(data.shape[1]),) + tuple(int(n) for n in reversed(N))
    return CgSenseSolve.apply(data, op, shape, int(iterations), float(tc_tolerance), float(encoding_weight),
                              float(kappa), grad)
