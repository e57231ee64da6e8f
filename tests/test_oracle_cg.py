"""Pins of the CG-SENSE oracle (oracle/cgsense.py) on the CPU.

The SENSE operator and the CG recurrence are checked against what the
mathematics fixes, on problems small enough to write the operator as a dense
matrix E (columns = mult_M of unit images):
  * mult_MH is the adjoint of mult_M (dense E^H);
  * CG converges to the direct solution of (E^H E + kappa I) x = E^H d, with and
    without a preconditioner (a preconditioner changes the path, not the fixed point);
  * successive CG residuals are mutually orthogonal (the Krylov property), which
    a dropped term or a wrong sign in alpha, beta or the updates breaks;
  * relativeResidualTCB stops after the first iteration whose rq / rq0 < tol;
  * with a preconditioner, the iterates equal those of textbook preconditioned
    CG in FP64 with M^-1 = diag(w)^2 (cgSolver applies the weights twice,
    cgSolver.h:248-249,328-329), and the residuals are diag(w)^2-orthogonal.
    The weights vary over the image (a constant preconditioner leaves CG's
    iterates unchanged, so it could not tell once from twice), and the test
    checks that a once-applied preconditioner would have failed it.
"""
import numpy as np
import pytest

from oracle import cgsense as cg
from oracle import pyoracle as po
from gadgetron_amd import trajectories as tj


def smooth_csm(rng, C, N):
    """Normalised complex Gaussian bumps (sigma = N/3) at random centres."""
    y, x = np.mgrid[0:N, 0:N]
    maps = []
    for _ in range(C):
        cy, cx = rng.uniform(0, N, 2)
        ph = rng.uniform(0, 2 * np.pi)
        maps.append(np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * (N / 3) ** 2)) * np.exp(1j * ph))
    maps = np.array(maps)
    maps /= np.sqrt(np.sum(np.abs(maps) ** 2, axis=0, keepdims=True))
    return maps.astype(np.complex64)


def small_problem(seed=0, N=12, spokes=8, C=2):
    rng = np.random.default_rng(seed)
    traj = tj.golden_radial_2d(2 * N, spokes)
    plan = po.Plan([N, N], 5.5, alpha=1.5, precision="float")
    plan.preprocess(traj)
    csm = smooth_csm(rng, C, N)
    dcw = np.sqrt(tj.ramp_dcf(traj)).astype(np.float32)
    op = cg.SenseOp(plan, csm, dcw)
    return rng, op


def dense_E(op):
    N2 = int(np.prod(op.img))
    cols = []
    for i in range(N2):
        e = np.zeros((1,) + op.img, np.complex64)
        e.flat[i] = 1.0
        cols.append(op.mult_M(e).ravel().astype(np.complex128))
    return np.array(cols).T  # (C*M, N2)


def test_sense_adjoint_is_dense_conjugate_transpose():
    rng, op = small_problem()
    E = dense_E(op)
    y = (rng.standard_normal((op.C, 1, op.M)) + 1j * rng.standard_normal((op.C, 1, op.M))).astype(np.complex64)
    mh = op.mult_MH(y).ravel()
    ref = E.conj().T @ y.ravel().astype(np.complex128)
    assert np.linalg.norm(mh - ref) / np.linalg.norm(ref) < 2e-5


# kappa > 0: with 8 spokes the unregularised normal matrix is singular (k-space
# corners are not sampled), so only the regularised system has a unique solution.
@pytest.mark.parametrize("kappa,precond", [(0.05, False), (0.05, True), (0.5, False)])
def test_cg_converges_to_direct_solution(kappa, precond):
    rng, op = small_problem(seed=1)
    E = dense_E(op)
    N2 = E.shape[1]
    A = E.conj().T @ E + kappa * np.eye(N2)
    d = (rng.standard_normal((op.C, 1, op.M)) + 1j * rng.standard_normal((op.C, 1, op.M))).astype(np.complex64)
    b = E.conj().T @ d.ravel().astype(np.complex128)
    x_ref = np.linalg.solve(A, b)
    w = None
    if precond:
        w = (1.0 / np.sqrt(np.sum(np.abs(op.csm) ** 2, axis=0) + 0.5)).astype(np.complex64)
    x, hist = cg.cg_solve(op, d, iterations=3 * N2, tc_tolerance=1e-14, kappa=kappa, precond=w)
    err = np.linalg.norm(x.ravel() - x_ref) / np.linalg.norm(x_ref)
    assert err < 1e-3, err
    assert hist[-1] < 1e-6 * hist[0]


def test_cg_residuals_are_orthogonal():
    rng, op = small_problem(seed=2)
    d = (rng.standard_normal((op.C, 1, op.M)) + 1j * rng.standard_normal((op.C, 1, op.M))).astype(np.complex64)
    x, hist, rs = cg.cg_solve(op, d, iterations=6, tc_tolerance=0.0, kappa=0.01, record=True)
    for i in range(len(rs)):
        for j in range(i):
            c = abs(cg.dotc(rs[i], rs[j])) / (np.linalg.norm(rs[i]) * np.linalg.norm(rs[j]))
            assert c < 2e-3, (i, j, c)


def test_cg_termination_after_first_iteration_below_tolerance():
    rng, op = small_problem(seed=3)
    d = (rng.standard_normal((op.C, 1, op.M)) + 1j * rng.standard_normal((op.C, 1, op.M))).astype(np.complex64)
    _, full = cg.cg_solve(op, d, iterations=8, tc_tolerance=0.0)
    ratios = np.array(full[1:]) / full[0]
    k = 3
    tol = 0.5 * (ratios[k - 1] + ratios[k]) if ratios[k] < ratios[k - 1] else ratios[k - 1] * 0.999
    _, part = cg.cg_solve(op, d, iterations=8, tc_tolerance=tol)
    first = int(np.argmax(ratios < tol)) + 1
    assert len(part) == first + 1


def dense_pcg(A, b, minv, iterations):
    """Textbook preconditioned CG in FP64 (Saad, Iterative Methods for Sparse
    Linear Systems, Alg. 9.1) with M^-1 = diag(minv), x_0 = 0; returns the
    iterates x_1 .. x_iterations."""
    x = np.zeros_like(b)
    r = b.copy()
    z = minv * r
    p = z.copy()
    rz = np.vdot(r, z)
    xs = []
    for _ in range(iterations):
        Ap = A @ p
        alpha = rz / np.vdot(p, Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        z = minv * r
        rz_new = np.vdot(r, z)
        p = z + (rz_new / rz) * p
        rz = rz_new
        xs.append(x.copy())
    return xs


def varying_weights(rng, img):
    """Real positive preconditioner weights that vary over the image."""
    return rng.uniform(0.4, 2.5, img).astype(np.complex64)


@pytest.mark.parametrize("iterations", [1, 2, 3])
def test_preconditioned_cg_equals_dense_pcg_with_squared_weights(iterations):
    rng, op = small_problem(seed=4)
    E = dense_E(op)
    kappa = 0.05
    A = E.conj().T @ E + kappa * np.eye(E.shape[1])
    d = (rng.standard_normal((op.C, 1, op.M)) + 1j * rng.standard_normal((op.C, 1, op.M))).astype(np.complex64)
    b = E.conj().T @ d.ravel().astype(np.complex128)
    w = varying_weights(rng, op.img)
    wd = w.ravel().astype(np.complex128)
    x, _ = cg.cg_solve(op, d, iterations=iterations, tc_tolerance=0.0, kappa=kappa, precond=w)
    twice = dense_pcg(A, b, wd * wd, iterations)[-1]
    once = dense_pcg(A, b, wd, iterations)[-1]
    err = np.linalg.norm(x.ravel() - twice) / np.linalg.norm(twice)
    assert err < 2e-5, err
    # the same check would reject a preconditioner applied once
    err_once = np.linalg.norm(x.ravel() - once) / np.linalg.norm(once)
    assert err_once > 100 * max(err, 1e-7), (err, err_once)


def test_preconditioned_residuals_are_weighted_orthogonal():
    rng, op = small_problem(seed=5)
    d = (rng.standard_normal((op.C, 1, op.M)) + 1j * rng.standard_normal((op.C, 1, op.M))).astype(np.complex64)
    w = varying_weights(rng, op.img)
    w2 = (np.abs(w.ravel()) ** 2).astype(np.float64)
    x, hist, rs = cg.cg_solve(op, d, iterations=6, tc_tolerance=0.0, kappa=0.01, precond=w, record=True)

    def wdot(a, b):
        return np.vdot(a.ravel().astype(np.complex128), w2 * b.ravel().astype(np.complex128))

    for i in range(len(rs)):
        for j in range(i):
            c = abs(wdot(rs[i], rs[j])) / np.sqrt(abs(wdot(rs[i], rs[i])) * abs(wdot(rs[j], rs[j])))
            assert c < 2e-3, (i, j, c)
    # rq is the weighted residual norm <r, diag(w)^2 r> (the reference's real(dot(r, P^2 r)))
    for k, r in enumerate(rs):
        assert abs(hist[k] - wdot(r, r).real) <= 1e-4 * hist[k]


# ---- image regularisers, initial guess, zero iterations -------------------

def test_image_operator_offset_equals_sorted_quantile():
    """estimate_offset's histogram walk, restated by sorting: the element at
    position floor(0.01 n) of the sorted |x| s lies in bin b, and the walk
    stops one bin past it, so offset = (b + 2) max / 50000."""
    rng = np.random.default_rng(11)
    for n in (4096, 10000, 65536):
        x = (rng.gamma(2.0, 1.0, n) * np.exp(2j * np.pi * rng.uniform(0, 1, n))).astype(np.complex64)
        recip, off = cg.image_operator_compute(x)
        xs = x.astype(np.complex128)
        s = np.float32(np.sum(np.abs(xs.real)) + np.sum(np.abs(xs.imag)))
        scale = np.float32(n) / s
        R = np.sqrt((x.real * scale).astype(np.float32) ** 2 + (x.imag * scale).astype(np.float32) ** 2).astype(np.float32)
        mx = R.max()
        target = int(np.float32(0.01) * np.float32(n))
        v = np.sort(R)[target - 1]
        b = min(int(np.floor((v / mx) * np.float32(50000))), 49999)
        assert off == pytest.approx(float(np.float32(b + 2) * mx / np.float32(50000)), rel=1e-6)
        np.testing.assert_allclose(recip, 1.0 / (R.astype(np.float64) + off), rtol=1e-6)
        # the normalisation: mean of |re| + |im| of the scaled image is one
        assert np.mean(np.abs(x.real * scale) + np.abs(x.imag * scale)) == pytest.approx(1.0, rel=1e-5)


def test_image_operator_constant_image_closed_form():
    """x = a > 0 everywhere: |x| s = 1, every element in the last bin, so the
    walk consumes all 50000 bins: offset = 50001 / 50000, image = 1 / (1 + offset)."""
    x = np.full((32, 32), 3.7, np.complex64)
    recip, off = cg.image_operator_compute(x)
    assert off == pytest.approx(50001.0 / 50000.0, rel=1e-6)
    np.testing.assert_allclose(recip, 1.0 / (1.0 + 50001.0 / 50000.0), rtol=1e-6)
    # no estimation: the given offset (here 0: no offset) is used
    recip0, off0 = cg.image_operator_compute(x, offset_estimation=False, offset=0.0)
    assert off0 == 0.0
    np.testing.assert_allclose(recip0, 1.0, rtol=1e-6)


def test_precond_weights_closed_form():
    """Normalised coil maps (sum |csm|^2 = 1): w = 1 / sqrt(1 + kappa R)."""
    rng = np.random.default_rng(12)
    csm = smooth_csm(rng, 4, 16)
    reg = rng.uniform(0.1, 3.0, (16, 16)).astype(np.float32)
    w = cg.precond_weights(csm, reg, 0.3)
    np.testing.assert_allclose(w.real, 1.0 / np.sqrt(1.0 + 0.3 * reg), rtol=2e-6)
    assert np.all(w.imag == 0)


def test_cg_with_image_regularisers_converges_to_direct_solution():
    rng, op = small_problem(seed=13)
    E = dense_E(op)
    N2 = E.shape[1]
    g1 = rng.uniform(0.2, 1.5, op.img).astype(np.float32)
    g2 = rng.uniform(0.5, 1.0, op.img).astype(np.float32)
    regs = [(0.3, g1), (0.1, g2), (0.02, None)]
    diag = 0.05 + 0.3 * g1.ravel() ** 2 + 0.1 * g2.ravel() ** 2 + 0.02
    A = E.conj().T @ E + np.diag(diag.astype(np.float64))
    d = (rng.standard_normal((op.C, 1, op.M)) + 1j * rng.standard_normal((op.C, 1, op.M))).astype(np.complex64)
    b = E.conj().T @ d.ravel().astype(np.complex128)
    x_ref = np.linalg.solve(A, b)
    x, hist = cg.cg_solve(op, d, iterations=3 * N2, tc_tolerance=1e-14, kappa=0.05, regs=regs)
    assert np.linalg.norm(x.ravel() - x_ref) / np.linalg.norm(x_ref) < 1e-3


def dense_pcg_x0(A, b, minv, iterations, x0):
    """Textbook preconditioned CG from x_0 (Saad Alg. 9.1)."""
    x = x0.copy()
    r = b - A @ x
    z = minv * r
    p = z.copy()
    rz = np.vdot(r, z)
    for _ in range(iterations):
        Ap = A @ p
        alpha = rz / np.vdot(p, Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        z = minv * r
        rz_new = np.vdot(r, z)
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x


@pytest.mark.parametrize("precond", [False, True])
def test_cg_initial_guess_equals_dense_pcg_from_x0(precond):
    rng, op = small_problem(seed=14)
    E = dense_E(op)
    kappa = 0.05
    g = rng.uniform(0.2, 1.5, op.img).astype(np.float32)
    regs = [(0.2, g)]
    A = E.conj().T @ E + np.diag((kappa + 0.2 * g.ravel() ** 2).astype(np.float64))
    d = (rng.standard_normal((op.C, 1, op.M)) + 1j * rng.standard_normal((op.C, 1, op.M))).astype(np.complex64)
    b = E.conj().T @ d.ravel().astype(np.complex128)
    x0 = (rng.standard_normal((1,) + op.img) + 1j * rng.standard_normal((1,) + op.img)).astype(np.complex64)
    w = varying_weights(rng, op.img) if precond else None
    minv = np.ones(b.size) if w is None else np.abs(w.ravel().astype(np.complex128)) ** 2
    for it in (1, 3):
        x, hist = cg.cg_solve(op, d, iterations=it, tc_tolerance=0.0, kappa=kappa, precond=w, regs=regs, x0=x0)
        ref = dense_pcg_x0(A, b, minv, it, x0.ravel().astype(np.complex128))
        assert np.linalg.norm(x.ravel() - ref) / np.linalg.norm(ref) < 2e-5
        # rq_0 is the right-hand side's <b, P^2 b>, taken before x0 is applied
        assert hist[0] == pytest.approx(np.vdot(b, minv * b).real, rel=1e-5)
    # the exact solution as x0: no residual left, x stays
    xs = np.linalg.solve(A, b).astype(np.complex64).reshape(x0.shape)
    x, _ = cg.cg_solve(op, d, iterations=3, tc_tolerance=0.0, kappa=kappa, precond=w, regs=regs, x0=xs)
    assert np.linalg.norm((x - xs).ravel()) / np.linalg.norm(xs.ravel()) < 1e-4


def test_cg_zero_iterations_returns_rhs():
    rng, op = small_problem(seed=15)
    d = (rng.standard_normal((op.C, 1, op.M)) + 1j * rng.standard_normal((op.C, 1, op.M))).astype(np.complex64)
    x, _ = cg.cg_solve(op, d, iterations=0, encoding_weight=0.5)
    np.testing.assert_allclose(x, 0.5 * op.mult_MH(d), rtol=1e-6, atol=1e-7)
