"""Non-Cartesian SENSE operator and CG solver over libgtnufft.so.

Mirrors cuNonCartesianSenseOperator<float, D>
(toolboxes/mri/pmri/gpu/cuNonCartesianSenseOperator.cu:14-90) and
cuCgSolver / cgSolver (toolboxes/solvers/cgSolver.h:105-390) through the C ABI
of include/gt_cgsense.h.  Shapes are numpy order (reversed Gadgetron dims):
image x (F, [Nz,] Ny, Nx), coil maps (C, [Nz,] Ny, Nx), data (C, F, M).
"""
import ctypes

import numpy as np

from ._lib import GTError, check, gt_cg_desc, gt_cg_reg, lib
from .nfft import NFFTPlan, _numel, _ptr, _stream

__all__ = ["SenseOperator", "cg_sense", "cg_sense_image", "CgBatch", "ImageOperator", "precond_weights"]


class SenseOperator:
    """E x = NFFT_C2NC(dcw; csm * x),  E^H y = sum_c conj(csm_c) NFFT_NC2C(dcw; y).

    As in the reference, pass sqrt(DCW) as dcw: the NFFT applies it both ways.
    The plan must be preprocessed; plan, csm and dcw must stay alive as long as
    the operator (this object keeps references to them)."""

    def __init__(self, plan: NFFTPlan, csm, dcw=None):
        D = plan.D
        shp = [int(s) for s in csm.shape]
        if len(shp) != D + 1:
            raise GTError(3, "coil maps must have shape (C, *image)")
        self.C = shp[0]
        self._keep = (plan, csm, dcw)
        h = ctypes.c_void_p()
        dl = 0 if dcw is None else _numel(dcw)
        check(lib().gt_sense_create(ctypes.byref(h), plan._h, _ptr(csm), self.C, _ptr(dcw), dl))
        self._h = h.value
        self.plan = plan

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().gt_sense_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def mult_M(self, x, y, accumulate=False, stream=None):
        check(lib().gt_sense_mult_M(self._h, _ptr(x), _ptr(y), int(bool(accumulate)), _stream(stream)))

    def mult_MH(self, y, x, accumulate=False, stream=None):
        check(lib().gt_sense_mult_MH(self._h, _ptr(y), _ptr(x), int(bool(accumulate)), _stream(stream)))

    def mult_MH_M(self, x, out, accumulate=False, stream=None):
        check(lib().gt_sense_mult_MH_M(self._h, _ptr(x), _ptr(out), int(bool(accumulate)), _stream(stream)))


def _desc(iterations, tc_tolerance, encoding_weight, kappa, precond, regs=None, x0=None):
    """regs: None or a list of (weight, image) image regularisers (image None:
    the identity; otherwise a float32 device array).  The returned descriptor
    keeps the regulariser list alive as d._regs."""
    d = gt_cg_desc()
    d.iterations = int(iterations)
    d.tc_tolerance = float(tc_tolerance)
    d.encoding_weight = float(encoding_weight)
    d.kappa = float(kappa)
    d.precond = None if precond is None else int(precond.data_ptr())
    d.precond_len = 0 if precond is None else _numel(precond)
    regs = list(regs or [])
    arr = (gt_cg_reg * max(1, len(regs)))()
    for k, (w, img) in enumerate(regs):
        arr[k].weight = float(w)
        arr[k].image = None if img is None else int(img.data_ptr())
        arr[k].image_len = 0 if img is None else _numel(img)
    d._regs = arr
    d.regs = ctypes.cast(arr, ctypes.POINTER(gt_cg_reg)) if regs else None
    d.n_regs = len(regs)
    d.x0 = None if x0 is None else int(x0.data_ptr())
    return d


def cg_sense(op: SenseOperator, data, x, iterations=10, tc_tolerance=1e-6, encoding_weight=1.0, kappa=0.0,
             precond=None, regs=None, x0=None, stream=None):
    """cgSolver::solve(data) into x; returns rq_0 and rq after each iteration
    (zeros after termination).  precond: complex64 diagonal weights (applied
    twice, as cgPreconditioner::apply in the reference recurrence); regs: image
    regularisers (weight, real image) added as weight * image^2 * x
    (imageOperator::mult_MH_M); x0: initial guess (may be x itself).
    iterations 0 returns the right-hand side, as cgSolver::solve_from_rhs."""
    d = _desc(iterations, tc_tolerance, encoding_weight, kappa, precond, regs, x0)
    hist = (ctypes.c_double * (int(iterations) + 1))()
    check(lib().gt_cgsense_solve(op._h, _ptr(data), _ptr(x), ctypes.byref(d), hist, _stream(stream)))
    return np.array(hist[:], dtype=np.float64)


def cg_sense_image(op: SenseOperator, rhs, x, iterations=10, tc_tolerance=1e-6, encoding_weight=1.0, kappa=0.0,
                   precond=None, regs=None, x0=None, stream=None):
    """The solver of cg_sense with an image-space right-hand side:
    x = CG(w_E E^H E + kappa I + regularisers, rhs); returns the residual history."""
    d = _desc(iterations, tc_tolerance, encoding_weight, kappa, precond, regs, x0)
    hist = (ctypes.c_double * (int(iterations) + 1))()
    check(lib().gt_cgsense_solve_image(op._h, _ptr(rhs), _ptr(x), ctypes.byref(d), hist, _stream(stream)))
    return np.array(hist[:], dtype=np.float64)


class CgBatch:
    """Many independent CG-SENSE reconstructions, set up once, run repeatedly.

    mode "plain": one problem at a time with the dot products read back to the
    host each time (the reference's method); "batched": every solve captured
    into one hipGraph with the problems on parallel branches; "fused": one
    launch per step for every problem (gridding over all problems' work items,
    FFT passes per size group, CG updates with a problem index), captured into
    one hipGraph (2D plans, identity regulariser only)."""

    def __init__(self, ops, data, xs, iterations=10, tc_tolerance=1e-6, encoding_weight=1.0, kappa=0.0,
                 mode="batched", stream=None):
        n = len(ops)
        if not (n == len(data) == len(xs)) or n == 0:
            raise GTError(1, "ops, data and xs must be non-empty lists of one length")
        self._keep = (list(ops), list(data), list(xs))
        PA = ctypes.c_void_p * n
        hops = PA(*[o._h for o in ops])
        hd = PA(*[int(a.data_ptr()) for a in data])
        hx = PA(*[int(a.data_ptr()) for a in xs])
        d = _desc(iterations, tc_tolerance, encoding_weight, kappa, None)
        m = {"plain": 0, "batched": 1, "fused": 2}[mode]
        h = ctypes.c_void_p()
        check(lib().gt_cgsense_batch_create(ctypes.byref(h), hops, hd, hx, n, ctypes.byref(d), m, _stream(stream)))
        self._h = h.value
        self.mode = mode

    def run(self, stream=None):
        check(lib().gt_cgsense_batch_run(self._h, _stream(stream)))

    @property
    def branches(self):
        """Parallel graph branches the autotuner chose (0 for the plain mode)."""
        n = ctypes.c_int()
        check(lib().gt_cgsense_batch_branches(self._h, ctypes.byref(n)))
        return n.value

    @property
    def item_length(self):
        """NC2C entries per work item the fused mode's autotuner chose (0: each
        plan's own rule; -1 for the other modes)."""
        n = ctypes.c_int()
        check(lib().gt_cgsense_batch_item_length(self._h, ctypes.byref(n)))
        return n.value

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().gt_cgsense_batch_destroy(self._h)
                self._h = None
        except Exception:
            pass


class ImageOperator:
    """cuImageOperator<float_complext> (operators/imageOperator.h,
    operators/gpu/cuImageOperator.h): a regulariser whose real image is
    1 / (|x| n / asum(x) + offset), computed from a complex image x; the
    offset comes from the reference's 50000-bin histogram (the 1 % quantile)
    or is given.  `image` is a float32 device array of x's size; `offset` the
    offset used.  As a CG regulariser it contributes weight * image^2 * x
    (pass (weight, op.image) in cg_sense's regs)."""

    def __init__(self, weight=1.0):
        self.weight = float(weight)
        self.image = None
        self.offset = 0.0

    def compute(self, x, offset_estimation=True, stream=None):
        from .nfft import DeviceArray
        n = _numel(x)
        if self.image is None or _numel(self.image) != n:
            self.image = DeviceArray.zeros(tuple(int(v) for v in x.shape), np.float32)
        off = ctypes.c_float(self.offset)
        check(lib().gt_image_operator_compute(_ptr(x), n, int(bool(offset_estimation)), ctypes.byref(off),
                                              _ptr(self.image), _stream(stream)))
        self.offset = float(off.value)
        return self.image


def precond_weights(csm, reg=None, kappa=0.0, out=None, stream=None):
    """The CG-SENSE gadget's preconditioner (gpuCgSenseGadget.cpp:165-176):
    1 / sqrt(sum_c |csm_c|^2 + kappa reg) as complex64 of the image shape.
    csm: (C, *image) complex64 device array; reg: None or float32 image."""
    from .nfft import DeviceArray
    shp = tuple(int(v) for v in csm.shape)
    C, img = shp[0], shp[1:]
    nimg = int(np.prod(img))
    if out is None:
        out = DeviceArray.zeros(img, np.complex64)
    check(lib().gt_cgsense_precond_weights(_ptr(csm), nimg, C, _ptr(reg), float(kappa), _ptr(out), _stream(stream)))
    return out
