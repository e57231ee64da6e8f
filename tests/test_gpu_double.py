"""GPU parity of double-precision plans (NFFT<cuNDArray, double, D>, STANDARD
convolution, the generic kernels in double) against the FP64 oracle on the
same inputs: relative L2 <= 1e-10 for every compute mode, the standalone
convolve / fft / deapodize, mult_MH_M, with and without density weights, in
1D, 2D and 3D; NC2C bitwise repeatable; the float-only features refuse double
plans.  Every case prints its name before touching the GPU.
"""
import sys

import numpy as np
import pytest

import gadgetron_amd as ga
from gadgetron_amd import DeviceArray, GTError, NFFT_comp_mode as CM, NFFT_conv_mode, NFFT_fft_mode, NFFTPlan
from oracle import pyoracle as po
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-10


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def dev(a, dtype):
    return DeviceArray.from_numpy(np.ascontiguousarray(a, dtype=dtype))


def cnormal(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2)


CASES = [([64], 300, 1, 2), ([40, 36], 900, 2, 3), ([16, 12, 20], 1500, 1, 2)]


@pytest.mark.parametrize("N,M,F,C", CASES)
def test_double_plan_matches_fp64_oracle(gpu, N, M, F, C):
    say(f"double plan {N}, {M} samples x {F} frames, {C} coils")
    rng = np.random.default_rng(len(N))
    D = len(N)
    traj = rng.uniform(-0.5, 0.5, (F, M, D))
    plan = NFFTPlan(N, None, W=5.5, os_factor=1.5, precision="double")
    plan.preprocess(dev(traj, np.float64))
    op = po.Plan(N, 5.5, alpha=1.5, precision="double")
    op.preprocess(traj)
    Nos = [int(n * 1.5) for n in N]
    img_shape = (C, F) + tuple(reversed(N))
    os_shape = (C, F) + tuple(reversed(Nos))
    smp_shape = (C, F, M)
    smp = cnormal(rng, smp_shape)
    img = cnormal(rng, img_shape)
    w_s = rng.uniform(0.2, 1.0, (F, M))
    w_c = rng.uniform(0.2, 1.0, (C, F, M))

    def run(x, out_shape, mode, dcw=None):
        dx = dev(x, np.complex128)
        dout = DeviceArray.zeros(out_shape, np.complex128)
        plan.compute(dx, dout, None if dcw is None else dev(dcw, np.float64), mode)
        plan.synchronize()
        return dout.numpy()

    for mode, x, shape, dcw in ((CM.FORWARDS_C2NC, img, smp_shape, w_s), (CM.BACKWARDS_NC2C, smp, img_shape, w_c),
                                (CM.FORWARDS_NC2C, smp, img_shape, None), (CM.BACKWARDS_C2NC, img, smp_shape, None),
                                (CM.BACKWARDS_NC2C, smp, os_shape, w_s)):
        got = run(x, shape, mode, dcw)
        want = op.compute(x, shape, int(mode), dcw)
        assert rel_l2(got, want) <= TOL, (mode, rel_l2(got, want))
    # mult_MH_M
    d_in, d_out = dev(img, np.complex128), DeviceArray.zeros(img_shape, np.complex128)
    plan.mult_MH_M(d_in, d_out, dev(w_s, np.float64))
    plan.synchronize()
    assert rel_l2(d_out.numpy(), op.mult_MH_M(img, w_s)) <= TOL
    # standalone convolve, fft, deapodize
    grid = cnormal(rng, os_shape)
    d_g = DeviceArray.zeros(os_shape, np.complex128)
    plan.convolve(dev(smp, np.complex128), d_g, NFFT_conv_mode.NC2C)
    plan.synchronize()
    ref = op.convolve(smp.reshape(-1), 1, C * F)
    assert rel_l2(d_g.numpy().reshape(-1), ref) <= TOL
    d_g2 = dev(grid, np.complex128)
    plan.convolve(dev(smp, np.complex128), d_g2, NFFT_conv_mode.NC2C, accumulate=True)
    plan.synchronize()
    assert rel_l2(d_g2.numpy().reshape(-1), grid.reshape(-1) + ref) <= TOL
    d_s = DeviceArray.zeros(smp_shape, np.complex128)
    plan.convolve(dev(grid, np.complex128), d_s, NFFT_conv_mode.C2NC)
    plan.synchronize()
    assert rel_l2(d_s.numpy().reshape(-1), op.convolve(grid.reshape(-1), 0, C * F)) <= TOL
    for fm in (NFFT_fft_mode.FORWARDS, NFFT_fft_mode.BACKWARDS):
        for scale in (True, False):
            d = dev(grid, np.complex128)
            plan.fft(d, fm, do_scale=scale)
            plan.synchronize()
            assert rel_l2(d.numpy(), op.fft(grid, C * F, fm == NFFT_fft_mode.FORWARDS, scale)) <= TOL
    d = dev(grid, np.complex128)
    plan.deapodize(d)
    plan.synchronize()
    assert rel_l2(d.numpy(), op.deapodize(grid, C * F)) <= TOL
    # NC2C repeatable bit for bit
    a = run(smp, img_shape, CM.BACKWARDS_NC2C, w_c)
    assert np.array_equal(a, run(smp, img_shape, CM.BACKWARDS_NC2C, w_c))


def test_double_plan_refusals(gpu):
    say("double plan refusals")
    with pytest.raises(GTError):
        NFFTPlan([32, 32], None, os_factor=1.5, conv=ga.ConvolutionType.ATOMIC, precision="double")
    plan = NFFTPlan([32, 32], None, os_factor=1.5, precision="double")
    traj = np.random.default_rng(0).uniform(-0.5, 0.5, (1, 100, 2))
    plan.preprocess(dev(traj, np.float64))
    csm = dev(np.ones((2, 32, 32), np.complex64), np.complex64)
    with pytest.raises(GTError):
        ga.SenseOperator(plan, csm, None)
    with pytest.raises(GTError):  # out of range, in double
        plan.preprocess(dev(np.full((1, 4, 2), 0.5000001), np.float64))
