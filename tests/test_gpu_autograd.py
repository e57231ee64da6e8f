"""Differentiable NUFFT (gadgetron_amd.autograd): torch autograd through
libgtnufft.so, loss = ||A(x; k) - y||^2 with y from a perturbed trajectory.

Checks (bar: 1e-3, the differentiable workload's tolerance):
  * image gradient against the analytic adjoint, 2 A^H (A x - y), computed by
    the FP64 oracle;
  * trajectory gradient against FP64 central finite differences (h = 1e-6)
    of the loss on 64 random coordinates.  A coordinate moves one sample, so
    each difference is evaluated on that sample alone (a one-sample
    trajectory through the FP64 oracle).  Coordinates whose +-h step moves a
    cell across a window edge (where the truncated Kaiser-Bessel weight
    jumps) are skipped.
"""
import sys

import numpy as np
import pytest
import torch

from gadgetron_amd import trajectories as tj
from gadgetron_amd.autograd import nufft
from oracle import pyoracle as po
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu
W, ALPHA, H = 5.5, 1.5, 1e-6


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def crosses_edge(p, nos, h):
    r = W / 2
    lo = [np.ceil(q - r) for q in (p - h * nos, p + h * nos)]
    hi = [np.floor(q + r) for q in (p - h * nos, p + h * nos)]
    return lo[0] != lo[1] or hi[0] != hi[1]


def check(N, C, spokes, S, seed, ncoord=64, with_dcw=False):
    rng = np.random.default_rng(seed)
    traj = tj.golden_radial_2d(S, spokes)                       # (1, M, 2)
    M = traj.shape[1]
    pert = np.clip(traj + rng.normal(0, 1e-3, traj.shape), -0.5, 0.5).astype(np.float32)
    x = tj.complex_normal(rng, (C, 1, N, N))
    dcw = tj.ramp_dcf(traj) if with_dcw else None               # (1, M), a constant of k
    d64 = None if dcw is None else dcw.astype(np.float64)
    ref = po.Plan([N, N], W, alpha=ALPHA, precision="double")
    ref.preprocess(pert.astype(np.float64))
    y = ref.compute(x.astype(np.complex128), (C, 1, M), po.FORWARDS_C2NC, dcw=d64).astype(np.complex64)

    dev = torch.device("cuda", 0)
    xt = torch.tensor(x, device=dev, requires_grad=True)
    kt = torch.tensor(traj, device=dev, requires_grad=True)
    yt = torch.tensor(y, device=dev)
    dt = None if dcw is None else torch.tensor(dcw, device=dev)
    loss = (nufft(xt, kt, [N, N], dcw=dt) - yt).abs().pow(2).sum()
    loss.backward()
    torch.cuda.synchronize()
    gx = xt.grad.cpu().numpy()
    gk = kt.grad.cpu().numpy()

    # image gradient: 2 A^H (A x - y) in FP64
    ref.preprocess(traj.astype(np.float64))
    Ax = ref.compute(x.astype(np.complex128), (C, 1, M), po.FORWARDS_C2NC, dcw=d64)
    resid = Ax - y.astype(np.complex128)
    gx_ref = 2.0 * ref.compute(resid, (C, 1, N, N), po.BACKWARDS_NC2C, dcw=d64)
    ex = np.linalg.norm(gx - gx_ref) / np.linalg.norm(gx_ref)

    # trajectory gradient: central differences, one sample at a time
    nos = N * ALPHA
    one = po.Plan([N, N], W, alpha=ALPHA, precision="double")
    gmax = float(np.max(np.abs(gk)))
    errs = []
    tried = 0
    while len(errs) < ncoord and tried < 20 * ncoord:
        tried += 1
        j = int(rng.integers(0, M))
        d = int(rng.integers(0, 2))
        k0 = traj[0, j].astype(np.float64)
        p = (k0[d] + 0.5) * nos
        if crosses_edge(p, nos, H) or abs(k0[d]) > 0.5 - 2 * H:
            continue
        vals = []
        for sgn in (1.0, -1.0):
            kk = k0.copy()
            kk[d] += sgn * H
            one.preprocess(kk.reshape(1, 1, 2))
            yj = one.compute(x.astype(np.complex128), (C, 1, 1), po.FORWARDS_C2NC,
                             dcw=None if d64 is None else d64[:, j:j + 1])[:, 0, 0]
            vals.append(float(np.sum(np.abs(yj - y[:, 0, j].astype(np.complex128)) ** 2)))
        fd = (vals[0] - vals[1]) / (2 * H)
        errs.append(abs(gk[0, j, d] - fd) / (abs(fd) + 1e-3 * gmax))
    return ex, np.array(errs)


def test_autograd_small(gpu):
    say("autograd: 32^2, 2 coils, 16 spokes x 64")
    ex, ek = check(32, 2, 16, 64, seed=4)
    print(f"image grad rel err {ex:.2e}, trajectory grad max rel err {ek.max():.2e} over {len(ek)} coords")
    assert len(ek) >= 48
    assert ex <= 1e-4
    assert ek.max() <= 1e-3


def test_autograd_small_with_dcw(gpu):
    """The same checks with a ramp DCW applied in the forward NUFFT (a constant
    of the trajectory, so dy/dk carries the sample's DCW)."""
    say("autograd: 32^2, 2 coils, 16 spokes x 64, ramp dcw")
    ex, ek = check(32, 2, 16, 64, seed=5, with_dcw=True)
    print(f"image grad rel err {ex:.2e}, trajectory grad max rel err {ek.max():.2e} over {len(ek)} coords")
    assert len(ek) >= 48
    assert ex <= 1e-4
    assert ek.max() <= 1e-3


def test_autograd_plan_reuse(gpu):
    """Repeated calls on an unchanged trajectory reuse one preprocessed plan;
    changing the trajectory in place gives a new plan and the new result."""
    from gadgetron_amd import autograd as ag
    say("autograd: plan reuse, 32^2, 1 coil")
    rng = np.random.default_rng(6)
    traj = tj.golden_radial_2d(64, 16)
    M = traj.shape[1]
    x = tj.complex_normal(rng, (1, 1, 32, 32))
    dev = torch.device("cuda", 0)
    ag.clear_plan_cache()
    kt = torch.tensor(traj, device=dev)
    xt = torch.tensor(x, device=dev)
    y1 = ag.nufft(xt, kt, [32, 32])
    p1 = ag._plan_for(kt, (32, 32), W, ALPHA, ag._stream())
    y2 = ag.nufft(xt, kt, [32, 32])
    assert ag._plan_for(kt, (32, 32), W, ALPHA, ag._stream()) is p1
    assert torch.equal(y1, y2)
    kt.mul_(0.5)  # in place: version counter moves
    y3 = ag.nufft(xt, kt, [32, 32])
    assert ag._plan_for(kt, (32, 32), W, ALPHA, ag._stream()) is not p1
    ref = po.Plan([32, 32], W, alpha=ALPHA, precision="float")
    ref.preprocess((traj * 0.5).astype(np.float32))
    e = rel_l2(y3.cpu().numpy(), ref.compute(x, (1, 1, M), po.FORWARDS_C2NC))
    say(f"  after the in-place change vs oracle: {e:.2e}")
    assert e <= 1e-5


def test_autograd_differentiable_workload(gpu):
    say("autograd: 256^2, 16 coils, 128 golden-angle spokes x 512, seed 4")
    ex, ek = check(256, 16, 128, 512, seed=4)
    print(f"image grad rel err {ex:.2e}, trajectory grad max rel err {ek.max():.2e} over {len(ek)} coords")
    assert len(ek) >= 48
    assert ex <= 1e-3
    assert ek.max() <= 1e-3


def test_cg_sense_solve_gradient(gpu):
    """Differentiable CG-SENSE: x = A^-1 w_E E^H d with A = w_E E^H E + kappa I.
    The forward solve is checked against the direct FP64 solution, the data
    gradient against the analytic adjoint w_E E A^-1 g and against FP64 central
    differences of the exact solution map, on a problem small enough for a
    dense E (12^2, 8 spokes, 2 coils)."""
    import gadgetron_amd as ga
    from gadgetron_amd.autograd import cg_sense_solve
    from oracle import cgsense as cg
    from tests.test_oracle_cg import dense_E, smooth_csm

    say("cg-sense gradient, 12^2 x 2 coils")
    N, spokes, C, kappa, we = 12, 8, 2, 0.5, 0.7
    rng = np.random.default_rng(21)
    traj = tj.golden_radial_2d(2 * N, spokes)
    M = traj.shape[1]
    csm = smooth_csm(rng, C, N)
    dcw = np.sqrt(tj.ramp_dcf(traj)).astype(np.float32)
    oplan = po.Plan([N, N], 5.5, alpha=1.5, precision="float")
    oplan.preprocess(traj)
    E = dense_E(cg.SenseOp(oplan, csm, dcw))  # (C*M, N*N), used in FP64
    A = we * (E.conj().T @ E) + kappa * np.eye(N * N)

    dev = torch.device("cuda")
    plan = ga.NFFTPlan([N, N], None, W=5.5, os_factor=1.5)
    t_traj = torch.as_tensor(traj, device=dev)
    plan.preprocess(t_traj)
    t_csm = torch.as_tensor(csm, device=dev)
    t_dcw = torch.as_tensor(dcw, device=dev)
    op = ga.SenseOperator(plan, t_csm, t_dcw)

    d = (rng.standard_normal((C, 1, M)) + 1j * rng.standard_normal((C, 1, M))).astype(np.complex64)
    t = (rng.standard_normal((1, N, N)) + 1j * rng.standard_normal((1, N, N))).astype(np.complex64)
    td = torch.as_tensor(d, device=dev).requires_grad_()
    x = cg_sense_solve(op, td, iterations=40, tc_tolerance=1e-12, encoding_weight=we, kappa=kappa)
    loss = (x - torch.as_tensor(t, device=dev)).abs().pow(2).sum()
    loss.backward()
    g = td.grad.detach().cpu().numpy().ravel().astype(np.complex128)

    d64 = d.ravel().astype(np.complex128)
    t64 = t.ravel().astype(np.complex128)
    solve = lambda dd: np.linalg.solve(A, we * (E.conj().T @ dd))
    x64 = solve(d64)
    ex = np.linalg.norm(x.detach().cpu().numpy().ravel() - x64) / np.linalg.norm(x64)
    say(f"  forward vs direct FP64 solve: {ex:.2e}")
    assert ex <= 1e-4
    # torch's gradient of a real loss is 2 dL/d conj(z): g_x = 2 (x - t)
    g_ref = we * (E @ np.linalg.solve(A, 2.0 * (x64 - t64)))
    eg = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    say(f"  gradient vs analytic adjoint: {eg:.2e}")
    assert eg <= 1e-3
    L = lambda dd: float(np.sum(np.abs(solve(dd) - t64) ** 2))
    worst = 0.0
    for _ in range(6):
        v = rng.standard_normal(d64.size) + 1j * rng.standard_normal(d64.size)
        h = 1e-6
        fd = (L(d64 + h * v) - L(d64 - h * v)) / (2 * h)
        an = float(np.real(np.vdot(g, v)))
        worst = max(worst, abs(fd - an) / max(abs(fd), 1e-12))
    say(f"  directional derivatives vs FP64 central differences: worst {worst:.2e}")
    assert worst <= 1e-3


def _truncated_cg64(A, b, iterations):
    """The CG recurrence of cgSolver (x_0 = 0, no preconditioner) in FP64 on a
    dense Hermitian A: the truncated iterate the unrolled gradient differentiates."""
    x = np.zeros_like(b)
    r = b.copy()
    p = r.copy()
    rq = np.vdot(r, r).real
    for _ in range(iterations):
        q = A @ p
        alpha = rq / np.vdot(p, q)
        x = x + alpha * p
        r = r - alpha * q
        tmp = np.vdot(r, r).real
        p = p * (tmp / rq) + r
        rq = tmp
    return x


def test_cg_sense_solve_unrolled_gradient(gpu):
    """grad="unrolled": the gradient of the truncated 3-iteration CG iterate
    (iterations recomputed in the backward pass), against FP64 central
    differences of the same truncated map on a dense E (12^2, 2 coils); it is
    not the exact-solution gradient, which the test also shows."""
    import gadgetron_amd as ga
    from gadgetron_amd.autograd import cg_sense_solve
    from oracle import cgsense as cg
    from tests.test_oracle_cg import dense_E, smooth_csm

    say("cg-sense unrolled gradient, 12^2 x 2 coils, 3 iterations")
    N, spokes, C, kappa, we, its = 12, 8, 2, 0.5, 0.7, 3
    rng = np.random.default_rng(22)
    traj = tj.golden_radial_2d(2 * N, spokes)
    M = traj.shape[1]
    csm = smooth_csm(rng, C, N)
    dcw = np.sqrt(tj.ramp_dcf(traj)).astype(np.float32)
    oplan = po.Plan([N, N], 5.5, alpha=1.5, precision="float")
    oplan.preprocess(traj)
    E = dense_E(cg.SenseOp(oplan, csm, dcw))
    A = we * (E.conj().T @ E) + kappa * np.eye(N * N)

    dev = torch.device("cuda")
    plan = ga.NFFTPlan([N, N], None, W=5.5, os_factor=1.5)
    plan.preprocess(torch.as_tensor(traj, device=dev))
    op = ga.SenseOperator(plan, torch.as_tensor(csm, device=dev), torch.as_tensor(dcw, device=dev))
    d = (rng.standard_normal((C, 1, M)) + 1j * rng.standard_normal((C, 1, M))).astype(np.complex64)
    t = (rng.standard_normal((1, N, N)) + 1j * rng.standard_normal((1, N, N))).astype(np.complex64)
    td = torch.as_tensor(d, device=dev).requires_grad_()
    x = cg_sense_solve(op, td, iterations=its, tc_tolerance=0.0, encoding_weight=we, kappa=kappa, grad="unrolled")
    loss = (x - torch.as_tensor(t, device=dev)).abs().pow(2).sum()
    loss.backward()
    g = td.grad.detach().cpu().numpy().ravel().astype(np.complex128)

    d64 = d.ravel().astype(np.complex128)
    t64 = t.ravel().astype(np.complex128)
    xn = lambda dd: _truncated_cg64(A, we * (E.conj().T @ dd), its)
    ex = np.linalg.norm(x.detach().cpu().numpy().ravel() - xn(d64)) / np.linalg.norm(xn(d64))
    say(f"  forward vs FP64 truncated CG: {ex:.2e}")
    assert ex <= 1e-4
    L = lambda dd: float(np.sum(np.abs(xn(dd) - t64) ** 2))
    worst = 0.0
    for _ in range(6):
        v = rng.standard_normal(d64.size) + 1j * rng.standard_normal(d64.size)
        h = 1e-6
        fd = (L(d64 + h * v) - L(d64 - h * v)) / (2 * h)
        an = float(np.real(np.vdot(g, v)))
        worst = max(worst, abs(fd - an) / max(abs(fd), 1e-12))
    say(f"  directional derivatives vs FP64 central differences: worst {worst:.2e}")
    assert worst <= 1e-3
    # the exact-solution gradient is a different vector at 3 iterations
    x_exact = np.linalg.solve(A, we * (E.conj().T @ d64))
    g_exact = we * (E @ np.linalg.solve(A, 2.0 * (xn(d64) - t64)))
    gap = np.linalg.norm(g - g_exact) / np.linalg.norm(g_exact)
    say(f"  unrolled vs exact-solution gradient: {gap:.2e}")
    # the two gradients differ by far more than the unrolled one's distance from the differences
    assert gap > max(100 * worst, 1e-3) and np.linalg.norm(xn(d64) - x_exact) > 1e-3 * np.linalg.norm(x_exact)
