"""Pins of the CG-SENSE oracle (oracle/cgsense.py) on the CPU.

The SENSE operator and the CG recurrence are checked against what the
mathematics fixes, on problems small enough to write the operator as a dense
matrix E (columns = mult_M of unit images):
  * mult_MH is the adjoint of mult_M (dense E^H);
  * CG converges to the direct solution of (E^H E + kappa I) x = E^H d, with and
    without a preconditioner (a preconditioner changes the path, not the fixed point);
  * successive CG residuals are mutually orthogonal (the Krylov property), which
    a dropped term or a wrong sign in alpha, beta or the updates breaks;
  * relativeResidualTCB stops after the first iteration whose rq / rq0 < tol.
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
