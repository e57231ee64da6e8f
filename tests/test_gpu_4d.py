"""GPU parity of 4-D plans (cuNFFT_impl<float, 4> / <double, 4>, cuNFFT.cu:270-283)
against the oracle on the same inputs: every compute mode, the standalone
convolve / fft / deapodize and mult_MH_M, for STANDARD, ATOMIC and
SPARSE_MATRIX float plans (relative L2 <= 1e-5) and STANDARD double plans
(<= 1e-10); NC2C bitwise repeatable; the features kept to D <= 3 (trajectory
gradient, SENSE operator) refuse a 4-D plan.  Every case prints its name
before touching the GPU.
"""
import sys

import numpy as np
import pytest

import gadgetron_amd as ga
from gadgetron_amd import DeviceArray, GTError, NFFT_comp_mode as CM, NFFT_conv_mode, NFFT_fft_mode, NFFTPlan
from oracle import pyoracle as po
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu
N4 = [8, 8, 4, 8]  # N_os = 12 x 12 x 6 x 12: even in every dimension


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def dev(a, dtype):
    return DeviceArray.from_numpy(np.ascontiguousarray(a, dtype=dtype))


def cnormal(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2)


@pytest.mark.parametrize("precision,conv", [("float", ga.ConvolutionType.STANDARD),
                                            ("float", ga.ConvolutionType.ATOMIC),
                                            ("double", ga.ConvolutionType.STANDARD)])
def test_4d_plan_matches_oracle(gpu, precision, conv):
    say(f"4-D plan {N4} {precision} {conv.name}")
    dbl = precision == "double"
    cdt, rdt = (np.complex128, np.float64) if dbl else (np.complex64, np.float32)
    tol = 1e-10 if dbl else 1e-5
    rng = np.random.default_rng(4)
    M, F, C = 700, 2, 2
    traj = rng.uniform(-0.5, 0.5, (F, M, 4)).astype(rdt)
    plan = NFFTPlan(N4, None, W=5.5, os_factor=1.5, conv=conv, precision=precision)
    plan.preprocess(dev(traj, rdt))
    assert plan.get_matrix_size() == N4 and plan.get_matrix_size_os() == [12, 12, 6, 12]
    op = po.Plan(N4, 5.5, alpha=1.5, precision=precision)
    op.preprocess(traj)
    Nos = [int(n * 1.5) for n in N4]
    img_shape = (C, F) + tuple(reversed(N4))
    os_shape = (C, F) + tuple(reversed(Nos))
    smp_shape = (C, F, M)
    smp = cnormal(rng, smp_shape).astype(cdt)
    img = cnormal(rng, img_shape).astype(cdt)
    w_s = rng.uniform(0.2, 1.0, (F, M)).astype(rdt)
    w_c = rng.uniform(0.2, 1.0, (C, F, M)).astype(rdt)

    def run(x, out_shape, mode, dcw=None):
        dout = DeviceArray.zeros(out_shape, cdt)
        plan.compute(dev(x, cdt), dout, None if dcw is None else dev(dcw, rdt), mode)
        plan.synchronize()
        return dout.numpy()

    # On the oversampled grid the deapodization reaches a product of four edge
    # factors at the Nyquist corners and multiplies the float rounding of the
    # FFT there by as much: float plans are compared before that factor (both
    # results divided by the filter), double plans as they are.
    filt = op.deapodize(np.ones(os_shape, cdt), C * F)
    for mode, x, shape, dcw in ((CM.FORWARDS_C2NC, img, smp_shape, w_s), (CM.BACKWARDS_NC2C, smp, img_shape, w_c),
                                (CM.FORWARDS_NC2C, smp, img_shape, None), (CM.BACKWARDS_C2NC, img, smp_shape, None),
                                (CM.BACKWARDS_NC2C, smp, os_shape, w_s)):
        got = run(x, shape, mode, dcw)
        want = op.compute(x, shape, int(mode), dcw)
        if shape == os_shape and not dbl:
            got, want = got / filt, want / filt
        assert rel_l2(got, want) <= tol, (mode, rel_l2(got, want))
    d_in, d_out = dev(img, cdt), DeviceArray.zeros(img_shape, cdt)
    plan.mult_MH_M(d_in, d_out, dev(w_s, rdt))
    plan.synchronize()
    assert rel_l2(d_out.numpy(), op.mult_MH_M(img, w_s)) <= tol
    grid = cnormal(rng, os_shape).astype(cdt)
    d_g = DeviceArray.zeros(os_shape, cdt)
    plan.convolve(dev(smp, cdt), d_g, NFFT_conv_mode.NC2C)
    plan.synchronize()
    assert rel_l2(d_g.numpy().reshape(-1), op.convolve(smp.reshape(-1), 1, C * F)) <= tol
    d_s = DeviceArray.zeros(smp_shape, cdt)
    plan.convolve(dev(grid, cdt), d_s, NFFT_conv_mode.C2NC)
    plan.synchronize()
    assert rel_l2(d_s.numpy().reshape(-1), op.convolve(grid.reshape(-1), 0, C * F)) <= tol
    for fm in (NFFT_fft_mode.FORWARDS, NFFT_fft_mode.BACKWARDS):
        d = dev(grid, cdt)
        plan.fft(d, fm, do_scale=True)
        plan.synchronize()
        assert rel_l2(d.numpy(), op.fft(grid, C * F, fm == NFFT_fft_mode.FORWARDS, True)) <= tol
    # The library builds the separable filter in double from the 1-D kernel
    # transforms; the float oracle takes it from a float 4-D FFT, whose
    # cancellation at the Nyquist corners leaves ~8e-4 relative error: the
    # filter is checked against the FP64 oracle.
    d = dev(grid, cdt)
    plan.deapodize(d)
    plan.synchronize()
    op64 = op if dbl else po.Plan(N4, 5.5, alpha=1.5, precision="double")
    assert rel_l2(d.numpy(), op64.deapodize(grid.astype(np.complex128), C * F)) <= tol
    if conv == ga.ConvolutionType.STANDARD:
        a = run(smp, img_shape, CM.BACKWARDS_NC2C, w_c)
        assert np.array_equal(a, run(smp, img_shape, CM.BACKWARDS_NC2C, w_c))


def test_4d_sparse_matrix_row_normalised(gpu):
    """SPARSE_MATRIX in 4-D: NC2C = the oracle's gridding of samples divided by
    the sum of each sample's window weights (from the oracle's convolution rows)."""
    say(f"4-D SPARSE_MATRIX {N4}")
    rng = np.random.default_rng(44)
    M, C = 600, 2
    traj = rng.uniform(-0.5, 0.5, (1, M, 4)).astype(np.float32)
    plan = NFFTPlan(N4, None, W=5.5, os_factor=1.5, conv=ga.ConvolutionType.SPARSE_MATRIX)
    plan.preprocess(dev(traj, np.float32))
    op = po.Plan(N4, 5.5, alpha=1.5, precision="float")
    op.preprocess(traj)
    wsum = np.array([op.conv_row(j)[1].astype(np.float64).sum() for j in range(M)])
    smp = cnormal(rng, (C, M)).astype(np.complex64)
    img_shape = (C,) + tuple(reversed(N4))
    d_out = DeviceArray.zeros(img_shape, np.complex64)
    plan.compute(dev(smp, np.complex64), d_out, None, CM.BACKWARDS_NC2C)
    plan.synchronize()
    want = op.compute(smp, img_shape, CM.BACKWARDS_NC2C, (1.0 / wsum).astype(np.float32))
    assert rel_l2(d_out.numpy().reshape(-1), want.reshape(-1)) <= 1e-5


def test_4d_refusals(gpu):
    say("4-D refusals")
    plan = NFFTPlan(N4, None, W=5.5, os_factor=1.5)
    traj = np.random.default_rng(0).uniform(-0.5, 0.5, (1, 50, 4)).astype(np.float32)
    plan.preprocess(dev(traj, np.float32))
    img = DeviceArray.zeros((1, 1) + tuple(reversed(N4)), np.complex64)
    gs = DeviceArray.zeros((1, 1, 50), np.complex64)
    gt = DeviceArray.zeros((1, 50, 4), np.float32)
    with pytest.raises(GTError):
        plan.grad_traj(img, gs, gt)
    with pytest.raises(GTError):
        ga.SenseOperator(plan, dev(np.ones((2,) + tuple(reversed(N4)), np.complex64), np.complex64), None)
    with pytest.raises(GTError):  # a window of 9 cells does not fit a double 4-D tile in LDS
        NFFTPlan([8, 8, 8, 8], None, W=8.5, os_factor=1.5, precision="double")
