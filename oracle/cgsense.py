"""TEST INFRASTRUCTURE ONLY (the oracle): CPU restatement of the non-Cartesian
SENSE operator and the CG solver that calls the NFFT.

Only tests/ and bench.py's cpu_baseline leg use it; the product never does.
It runs on the hoNFFT restatement (pyoracle.Plan) and follows:
  mri/pmri/gpu/cuNonCartesianSenseOperator.cu:14-62   mult_M / mult_MH
  mri/pmri/gpu/sense_utilities.cu:7-123                csm multiply, conj(csm) coil sum
  operators/linearOperator.h:75-85                     mult_MH_M = mult_MH(mult_M(x))
  solvers/cgSolver.h:219-390                           compute_rhs, initialize, iterate, mult_MH_M
  solvers/cgCallback.h:56-104                          relativeResidualTCB (stop once rq/rq0 < tol)
  solvers/cgPreconditioner.h:24-39                     apply = elementwise weights (used twice)
Dot products are FP64 (sum conj(a) b); vectors are complex64 like the
reference's cuNDArray<complext<float>>.  Pinned in tests/test_oracle_cg.py
against a dense operator: adjoint identities, convergence to the direct
solution of the normal equations, Krylov residual orthogonality.
"""
import numpy as np

from . import pyoracle as po

C64 = np.complex64


def dotc(a, b):
    """<a, b> = sum conj(a) b in FP64 (cuNDArray_blas.cu:49-77 cdotc)."""
    return complex(np.vdot(np.asarray(a, np.complex128).ravel(), np.asarray(b, np.complex128).ravel()))


class SenseOp:
    """E x = NFFT_C2NC(dcw; csm x), E^H y = sum_c conj(csm_c) NFFT_NC2C(dcw; y).

    plan: preprocessed pyoracle.Plan; csm: (C, *img) complex; dcw: (F, M) or None.
    x: (F, *img); y: (C, F, M)."""

    def __init__(self, plan, csm, dcw=None):
        self.plan = plan
        self.csm = np.asarray(csm, C64)
        self.C = self.csm.shape[0]
        self.img = tuple(self.csm.shape[1:])
        self.dcw = dcw
        self.F = plan.F
        self.M = plan.M

    def mult_M(self, x):
        tmp = (self.csm[:, None] * np.asarray(x, C64)[None]).astype(C64)
        return self.plan.compute(tmp, (self.C, self.F, self.M), po.FORWARDS_C2NC, dcw=self.dcw)

    def mult_MH(self, y):
        tmp = self.plan.compute(np.asarray(y, C64), (self.C, self.F) + self.img, po.BACKWARDS_NC2C, dcw=self.dcw)
        return np.sum(np.conj(self.csm)[:, None] * tmp, axis=0).astype(C64)

    def mult_MH_M(self, x):
        return self.mult_MH(self.mult_M(x))


def cg_solve(op, data, iterations, tc_tolerance=1e-6, encoding_weight=1.0, kappa=0.0, precond=None,
             record=False):
    """cgSolver::solve(data) with an identity regulariser of weight kappa.

    Returns (x, history of rq) and, with record=True, the list of residuals r_k
    (for the orthogonality check)."""
    we = np.float32(encoding_weight)
    w = None if precond is None else np.asarray(precond, C64).ravel()

    def apply2(v):  # precond_->apply twice
        flat = v.ravel()
        ww = np.resize(w, flat.shape)
        return (flat * ww * ww).astype(C64).reshape(v.shape)

    rhs = (we * op.mult_MH(data)).astype(C64)
    x = np.zeros_like(rhs)
    r = rhs.copy()
    p = r.copy() if w is None else apply2(r)
    rq0 = dotc(r, p).real
    rq = rq0
    hist = [rq0]
    rs = [r.copy()]
    if not rq0 > 0:
        return (x, hist, rs) if record else (x, hist)
    for _ in range(int(iterations)):
        q = (we * op.mult_MH_M(p) + np.float32(kappa) * p).astype(C64)
        alpha = rq / dotc(p, q)
        x = (x + alpha * p).astype(C64)
        r = (r - alpha * q).astype(C64)
        z = r if w is None else apply2(r)
        tmp = dotc(r, z).real
        p = (p * (tmp / rq) + z).astype(C64)
        rq = tmp
        hist.append(rq)
        rs.append(r.copy())
        if rq / rq0 < tc_tolerance:
            break
    return (x, hist, rs) if record else (x, hist)
