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
  operators/imageOperator.h:28-96                      compute (asum normalisation, abs, offset,
                                                       reciprocal), mult_MH_M = x * image^2
  operators/gpu/cuImageOperator.h:30-63                estimate_offset (50000-bin histogram, 1 %)
  gadgets/pmri/gpuCgSenseGadget.cpp:165-176            preconditioner 1/sqrt(sum |csm|^2 + kappa R)
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


HIST_BINS = 50000


def image_operator_compute(image, offset_estimation=True, offset=0.0):
    """imageOperator::compute (imageOperator.h:28-47) with
    cuImageOperator::estimate_offset (cuImageOperator.h:30-63), REAL = float:
    tmp = image * (n / asum(image)), asum = sum |re| + |im|; R = |tmp|; the
    offset is (k + 1) * max / 50000 where k bins of the histogram of R / max
    were consumed to reach floor(0.01 n) elements; R += offset (if > 0);
    returns (1 / R, offset)."""
    x = np.asarray(image, C64).ravel()
    n = x.size
    s = np.float32(np.sum(np.abs(x.real).astype(np.float64)) + np.sum(np.abs(x.imag).astype(np.float64)))
    scale = np.float32(n) / s
    re = (x.real * scale).astype(np.float32)
    im = (x.imag * scale).astype(np.float32)
    R = np.sqrt(re * re + im * im).astype(np.float32)
    if offset_estimation:
        mx = R.max()
        bins = np.minimum(np.floor((R / mx) * np.float32(HIST_BINS)).astype(np.int64), HIST_BINS - 1)
        hist = np.bincount(bins, minlength=HIST_BINS)
        target = int(np.float32(0.01) * np.float32(n))
        cumsum = counter = 0
        while cumsum < target:
            cumsum += int(hist[counter])
            counter += 1
        offset = float(np.float32(counter + 1) * mx / np.float32(HIST_BINS))
    if offset > 0:
        R = (R + np.float32(offset)).astype(np.float32)
    return (np.float32(1.0) / R).astype(np.float32).reshape(np.shape(image)), float(offset)


def precond_weights(csm, reg=None, kappa=0.0):
    """gpuCgSenseGadget::process (gpuCgSenseGadget.cpp:165-176): 1 / sqrt(sum_c
    |csm_c|^2 + kappa reg), complex64 of the image shape (csm: (C, *img))."""
    a = np.sum(np.abs(np.asarray(csm, C64)) ** 2, axis=0).astype(np.float32)
    if reg is not None:
        a = (a + np.float32(kappa) * np.asarray(reg, np.float32).reshape(a.shape)).astype(np.float32)
    return (np.float32(1.0) / np.sqrt(a)).astype(C64)


def cg_solve(op, data, iterations, tc_tolerance=1e-6, encoding_weight=1.0, kappa=0.0, precond=None,
             record=False, regs=None, x0=None, rhs=None):
    """cgSolver::solve(data) with an identity regulariser of weight kappa and
    image regularisers regs = [(weight, image)] (cgSolver::mult_MH_M adds
    weight * image^2 * x for each, cgSolver.h:376-384); x0: initial guess
    (cgSolver::initialize, cgSolver.h:240-280: rq_0 is taken before x0 is
    applied); rhs: an image right-hand side instead of w_E E^H data.

    Returns (x, history of rq) and, with record=True, the list of residuals r_k
    (for the orthogonality check).  iterations 0 returns the right-hand side
    (cgSolver::solve_from_rhs, cgSolver.h:124-129)."""
    we = np.float32(encoding_weight)
    w = None if precond is None else np.asarray(precond, C64).ravel()
    regs = list(regs or [])

    def apply2(v):  # precond_->apply twice
        flat = v.ravel()
        ww = np.resize(w, flat.shape)
        return (flat * ww * ww).astype(C64).reshape(v.shape)

    def mhm(v):  # cgSolver::mult_MH_M: encoding, then the regularisers in list order
        out = (we * op.mult_MH_M(v) + np.float32(kappa) * v).astype(C64)
        for wt, img in regs:
            g = np.ones(1, np.float32) if img is None else np.asarray(img, np.float32).ravel()
            g2 = np.resize(g, v.size).reshape(v.shape)
            out = (out + np.float32(wt) * (v * g2 * g2)).astype(C64)
        return out

    if rhs is None:
        rhs = (we * op.mult_MH(data)).astype(C64)
    rhs = np.asarray(rhs, C64)
    if int(iterations) == 0:
        return (rhs.copy(), [], [rhs.copy()]) if record else (rhs.copy(), [])
    x = np.zeros_like(rhs)
    r = rhs.copy()
    p = r.copy() if w is None else apply2(r)
    rq0 = dotc(r, p).real
    if x0 is not None:
        x = np.asarray(x0, C64).copy()
        r = (r - mhm(x)).astype(C64)
        p = r.copy() if w is None else apply2(r)
    rq = dotc(r, p).real
    hist = [rq0]
    rs = [r.copy()]
    if not rq > 0:
        return (x, hist, rs) if record else (x, hist)
    for _ in range(int(iterations)):
        q = mhm(p)
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
