"""Non-Cartesian SENSE operator and CG solver over libgtnufft.so.

Mirrors cuNonCartesianSenseOperator<float, D>
(toolboxes/mri/pmri/gpu/cuNonCartesianSenseOperator.cu:14-90) and
cuCgSolver / cgSolver (toolboxes/solvers/cgSolver.h:105-390) through the C ABI
of include/gt_cgsense.h.  Shapes are numpy order (reversed Gadgetron dims):
image x (F, [Nz,] Ny, Nx), coil maps (C, [Nz,] Ny, Nx), data (C, F, M).
"""
import ctypes

import numpy as np

from ._lib import GTError, check, gt_cg_desc, lib
from .nfft import NFFTPlan, _numel, _ptr, _stream

__all__ = ["SenseOperator", "cg_sense", "cg_sense_image", "CgBatch"]


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


def _desc(iterations, tc_tolerance, encoding_weight, kappa, precond):
    d = gt_cg_desc()
    d.iterations = int(iterations)
    d.tc_tolerance = float(tc_tolerance)
    d.encoding_weight = float(encoding_weight)
    d.kappa = float(kappa)
    d.precond = None if precond is None else int(precond.data_ptr())
    d.precond_len = 0 if precond is None else _numel(precond)
    return d


def cg_sense(op: SenseOperator, data, x, iterations=10, tc_tolerance=1e-6, encoding_weight=1.0, kappa=0.0,
             precond=None, stream=None):
    """cgSolver::solve(data) into x; returns rq_0 and rq after each iteration
    (zeros after termination).  precond: complex64 diagonal weights (applied
    twice, as cgPreconditioner::apply in the reference recurrence)."""
    d = _desc(iterations, tc_tolerance, encoding_weight, kappa, precond)
    hist = (ctypes.c_double * (int(iterations) + 1))()
    check(lib().gt_cgsense_solve(op._h, _ptr(data), _ptr(x), ctypes.byref(d), hist, _stream(stream)))
    return np.array(hist[:], dtype=np.float64)


def cg_sense_image(op: SenseOperator, rhs, x, iterations=10, tc_tolerance=1e-6, encoding_weight=1.0, kappa=0.0,
                   precond=None, stream=None):
    """The solver of cg_sense with an image-space right-hand side:
    x = CG(w_E E^H E + kappa I, rhs); returns the residual history."""
    d = _desc(iterations, tc_tolerance, encoding_weight, kappa, precond)
    hist = (ctypes.c_double * (int(iterations) + 1))()
    check(lib().gt_cgsense_solve_image(op._h, _ptr(rhs), _ptr(x), ctypes.byref(d), hist, _stream(stream)))
    return np.array(hist[:], dtype=np.float64)


class CgBatch:
    """Many independent CG-SENSE reconstructions, set up once, run repeatedly.

    mode "plain": one problem at a time with the dot products read back to the
    host each time (the reference's method); "batched": every solve captured
    into one hipGraph with the problems on parallel branches."""

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
        m = {"plain": 0, "batched": 1}[mode]
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

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().gt_cgsense_batch_destroy(self._h)
                self._h = None
        except Exception:
            pass
