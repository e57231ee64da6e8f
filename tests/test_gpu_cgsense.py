"""GPU parity of the SENSE operator and the CG solver (include/gt_cgsense.h)
against the oracle (oracle/cgsense.py over the hoNFFT restatement).

Bars: operator applications within relative L2 1e-5 of the oracle; CG
solutions after 10 iterations within 1e-4 (the recurrence amplifies the FP32
noise of the NFFT); the batched hipGraph solver equal bit for bit to the plain
per-problem solver and to single solves.
"""
import sys

import numpy as np
import pytest

import gadgetron_amd as ga
from gadgetron_amd import DeviceArray, NFFTPlan
from gadgetron_amd import trajectories as tj
from oracle import cgsense as cg
from oracle import pyoracle as po
from tests.helpers import rel_l2
from tests.test_oracle_cg import smooth_csm

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def dev(a, dtype):
    return DeviceArray.from_numpy(np.ascontiguousarray(a, dtype=dtype))


class Problem:
    """One 2D radial SENSE problem on the GPU and on the oracle."""

    def __init__(self, seed, N, spokes, C):
        rng = np.random.default_rng(seed)
        self.N, self.C = N, C
        self.traj = tj.golden_radial_2d(2 * N, spokes)
        self.M = self.traj.shape[1]
        self.csm = smooth_csm(rng, C, N)
        self.dcw = np.sqrt(tj.ramp_dcf(self.traj)).astype(np.float32)
        self.x_true = (rng.standard_normal((1, N, N)) + 1j * rng.standard_normal((1, N, N))).astype(np.complex64)
        self.plan = NFFTPlan([N, N], None, W=5.5, os_factor=1.5)
        self.d_traj = dev(self.traj, np.float32)
        self.plan.preprocess(self.d_traj)
        self.d_csm = dev(self.csm, np.complex64)
        self.d_dcw = dev(self.dcw, np.float32)
        self.op = ga.SenseOperator(self.plan, self.d_csm, self.d_dcw)
        oplan = po.Plan([N, N], 5.5, alpha=1.5, precision="float")
        oplan.preprocess(self.traj)
        self.oop = cg.SenseOp(oplan, self.csm, self.dcw)
        y = self.oop.mult_M(self.x_true)
        noise = 1e-3 * (rng.standard_normal(y.shape) + 1j * rng.standard_normal(y.shape)) / np.sqrt(2)
        self.data = (y + noise).astype(np.complex64)
        self.d_data = dev(self.data, np.complex64)


def test_sense_operator_matches_oracle(gpu):
    say("SENSE mult_M / mult_MH / mult_MH_M, 64^2, 32 spokes, 4 coils")
    P = Problem(0, 64, 32, 4)
    rng = np.random.default_rng(5)
    x = (rng.standard_normal((1, 64, 64)) + 1j * rng.standard_normal((1, 64, 64))).astype(np.complex64)
    y = (rng.standard_normal((4, 1, P.M)) + 1j * rng.standard_normal((4, 1, P.M))).astype(np.complex64)
    dx, dy = dev(x, np.complex64), dev(y, np.complex64)
    out_y = DeviceArray.zeros((4, 1, P.M), np.complex64)
    out_x = DeviceArray.zeros((1, 64, 64), np.complex64)
    P.op.mult_M(dx, out_y)
    P.plan.synchronize()
    assert rel_l2(out_y.numpy(), P.oop.mult_M(x)) <= 1e-5
    P.op.mult_MH(dy, out_x)
    P.plan.synchronize()
    assert rel_l2(out_x.numpy(), P.oop.mult_MH(y)) <= 1e-5
    P.op.mult_MH_M(dx, out_x)
    P.plan.synchronize()
    assert rel_l2(out_x.numpy(), P.oop.mult_MH_M(x)) <= 1e-5
    # accumulate adds to the output
    base = out_x.numpy()
    P.op.mult_MH(dy, out_x, accumulate=True)
    P.plan.synchronize()
    assert rel_l2(out_x.numpy(), base + P.oop.mult_MH(y)) <= 1e-5
    # adjoint identity on the device results
    P.op.mult_M(dx, out_y)
    P.op.mult_MH(dy, out_x)
    P.plan.synchronize()
    lhs = np.vdot(y.astype(np.complex128), out_y.numpy().astype(np.complex128))
    rhs = np.vdot(out_x.numpy().astype(np.complex128), x.astype(np.complex128))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


@pytest.mark.parametrize("kappa,use_precond", [(0.0, False), (0.01, True)])
def test_cgsense_matches_oracle(gpu, kappa, use_precond):
    say(f"CG-SENSE 10 iterations, 96^2, 48 spokes, 6 coils, kappa {kappa}, precond {use_precond}")
    P = Problem(1, 96, 48, 6)
    w = None
    d_w = None
    if use_precond:
        # weights that vary over the image (a constant one leaves the iterates unchanged)
        w = np.random.default_rng(7).uniform(0.4, 2.5, (1, 96, 96)).astype(np.complex64)
        d_w = dev(w, np.complex64)
    x = DeviceArray.zeros((1, 96, 96), np.complex64)
    hist = ga.cg_sense(P.op, P.d_data, x, iterations=10, tc_tolerance=1e-12, kappa=kappa, precond=d_w)
    xo, ho = cg.cg_solve(P.oop, P.data, 10, tc_tolerance=1e-12, kappa=kappa, precond=w)
    assert rel_l2(x.numpy(), xo) <= 1e-4
    assert np.allclose(hist[: len(ho)], ho, rtol=1e-3)


def test_cgsense_terminates_like_relative_residual_callback(gpu):
    say("CG-SENSE termination, 64^2, 24 spokes, 3 coils")
    P = Problem(2, 64, 24, 3)
    _, ho = cg.cg_solve(P.oop, P.data, 10, tc_tolerance=0.0)
    ratios = np.array(ho[1:]) / ho[0]
    tol = float(np.sqrt(ratios[1] * ratios[2]))  # between the 2nd and 3rd iteration
    x = DeviceArray.zeros((1, 64, 64), np.complex64)
    hist = ga.cg_sense(P.op, P.d_data, x, iterations=10, tc_tolerance=tol)
    xo, ho2 = cg.cg_solve(P.oop, P.data, 10, tc_tolerance=tol)
    n_it = len(ho2) - 1
    assert n_it == 3
    assert np.all(hist[n_it + 1:] == 0.0) and hist[n_it] > 0.0
    assert rel_l2(x.numpy(), xo) <= 1e-4


def test_batched_equals_plain_and_single(gpu):
    say("batched vs plain CG-SENSE, 5 uneven problems")
    specs = [(10, 64, 32, 4), (11, 96, 40, 6), (12, 128, 24, 3), (13, 64, 64, 8), (14, 96, 16, 2)]
    probs = [Problem(*s) for s in specs]
    xs_p = [DeviceArray.zeros((1, p.N, p.N), np.complex64) for p in probs]
    xs_b = [DeviceArray.zeros((1, p.N, p.N), np.complex64) for p in probs]
    plain = ga.CgBatch([p.op for p in probs], [p.d_data for p in probs], xs_p, iterations=10, mode="plain")
    batched = ga.CgBatch([p.op for p in probs], [p.d_data for p in probs], xs_b, iterations=10, mode="batched")
    plain.run()
    batched.run()
    batched.run()  # a second replay gives the same result (state is re-initialised on device)
    probs[0].plan.synchronize()
    for p, a, b in zip(probs, xs_p, xs_b):
        xa, xb = a.numpy(), b.numpy()
        assert np.array_equal(xa, xb)
        xs = DeviceArray.zeros((1, p.N, p.N), np.complex64)
        ga.cg_sense(p.op, p.d_data, xs, iterations=10)
        assert np.array_equal(xs.numpy(), xa)
        xo, _ = cg.cg_solve(p.oop, p.data, 10)
        assert rel_l2(xa, xo) <= 1e-4


def test_batched_refuses_replay_after_plan_changes(gpu):
    say("batched CG-SENSE after a bound plan is preprocessed again: run must refuse")
    probs = [Problem(20, 64, 16, 3), Problem(21, 96, 24, 4)]
    xs = [DeviceArray.zeros((1, p.N, p.N), np.complex64) for p in probs]
    batched = ga.CgBatch([p.op for p in probs], [p.d_data for p in probs], xs, iterations=4, mode="batched")
    batched.run()
    probs[0].plan.synchronize()
    # the same trajectory again: the tables are rebuilt in place, the graph holds stale launch sizes
    probs[1].plan.preprocess(probs[1].d_traj)
    with pytest.raises(ga.GTError) as e:
        batched.run()
    assert e.value.status == 1 and "create the batch again" in str(e.value)
    # a new batch over the same operators works and agrees with single solves
    again = ga.CgBatch([p.op for p in probs], [p.d_data for p in probs], xs, iterations=4, mode="batched")
    again.run()
    probs[0].plan.synchronize()
    for p, x in zip(probs, xs):
        xs1 = DeviceArray.zeros((1, p.N, p.N), np.complex64)
        ga.cg_sense(p.op, p.d_data, xs1, iterations=4)
        assert np.array_equal(xs1.numpy(), x.numpy())


def test_image_operator_and_precond_weights_match_oracle(gpu):
    say("cuImageOperator::compute and the gadget's preconditioner weights, 96^2, 6 coils")
    rng = np.random.default_rng(30)
    img = (rng.gamma(2.0, 1.0, (1, 96, 96)) * np.exp(2j * np.pi * rng.uniform(0, 1, (1, 96, 96)))).astype(np.complex64)
    R = ga.sense.ImageOperator()
    R.compute(dev(img, np.complex64))
    ref, off = cg.image_operator_compute(img)
    assert R.offset == pytest.approx(off, rel=1e-6)
    assert rel_l2(R.image.numpy(), ref) <= 1e-6
    # a given offset is used as is
    R2 = ga.sense.ImageOperator()
    R2.offset = 0.25
    R2.compute(dev(img, np.complex64), offset_estimation=False)
    ref2, _ = cg.image_operator_compute(img, offset_estimation=False, offset=0.25)
    assert rel_l2(R2.image.numpy(), ref2) <= 1e-6
    csm = smooth_csm(rng, 6, 96)
    w = ga.sense.precond_weights(dev(csm, np.complex64), R.image, kappa=0.3)
    assert rel_l2(w.numpy(), cg.precond_weights(csm, ref, 0.3)) <= 1e-6


@pytest.mark.parametrize("with_x0", [False, True])
def test_cgsense_regularised_preconditioned_matches_oracle(gpu, with_x0):
    say(f"CG-SENSE with an image regulariser, gadget preconditioner, x0 {with_x0}: 96^2, 48 spokes, 6 coils")
    P = Problem(3, 96, 48, 6)
    rng = np.random.default_rng(31)
    # the gadget's order: R from a reference image, then precond 1 / sqrt(sum |csm|^2 + kappa R)
    ref_img = (P.x_true + 0.1 * (rng.standard_normal(P.x_true.shape) + 1j * rng.standard_normal(P.x_true.shape))
               ).astype(np.complex64)
    R = ga.sense.ImageOperator()
    R.compute(dev(ref_img, np.complex64))
    rimg, _ = cg.image_operator_compute(ref_img)
    kappa = 0.1
    w = ga.sense.precond_weights(P.d_csm, R.image, kappa=kappa)
    w_ref = cg.precond_weights(P.csm, rimg, kappa)
    x0 = (rng.standard_normal((1, 96, 96)) + 1j * rng.standard_normal((1, 96, 96))).astype(np.complex64)
    x = dev(x0, np.complex64) if with_x0 else DeviceArray.zeros((1, 96, 96), np.complex64)
    hist = ga.cg_sense(P.op, P.d_data, x, iterations=10, tc_tolerance=0.0, precond=w, regs=[(kappa, R.image)],
                       x0=x if with_x0 else None)  # x is its own initial guess
    xo, ho = cg.cg_solve(P.oop, P.data, 10, tc_tolerance=0.0, precond=w_ref, regs=[(kappa, rimg)],
                         x0=x0 if with_x0 else None)
    assert rel_l2(x.numpy(), xo) <= 1e-4
    assert np.allclose(hist[: len(ho)], ho, rtol=1e-3)


def test_cgsense_zero_iterations_returns_rhs(gpu):
    say("CG-SENSE with zero iterations returns w_E E^H d, 64^2, 24 spokes, 3 coils")
    P = Problem(4, 64, 24, 3)
    x = DeviceArray.zeros((1, 64, 64), np.complex64)
    ga.cg_sense(P.op, P.d_data, x, iterations=0, encoding_weight=0.5)
    xo, _ = cg.cg_solve(P.oop, P.data, 0, encoding_weight=0.5)
    assert rel_l2(x.numpy(), xo) <= 1e-5
