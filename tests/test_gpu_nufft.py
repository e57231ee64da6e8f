"""GPU parity of libgtnufft.so against the oracle (the CPU hoNFFT restatement).

Bar: relative L2 <= 1e-5 against the float oracle on identical inputs and
kernel parameters (complex FP32 arithmetic, FP32 accumulation); bitwise
repeatability of the STANDARD (tile gather) NC2C.  Every case prints its name
before touching the GPU.
"""
import os
import sys

import numpy as np
import pytest

import gadgetron_amd as ga
from gadgetron_amd import DeviceArray, NFFT_comp_mode as CM, NFFT_conv_mode, NFFT_fft_mode, NFFTPlan
from gadgetron_amd import trajectories as tj
from oracle import pyoracle as po
from tests.helpers import c1_inputs, rel_l2

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-5


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def dev(a, dtype=None):
    return DeviceArray.from_numpy(np.ascontiguousarray(a, dtype=dtype))


def gpu_plan(N, W=5.5, conv=ga.ConvolutionType.STANDARD, tile=None, max_items=0):
    p = NFFTPlan(N, None, W=W, os_factor=1.5, conv=conv)
    if tile is not None or max_items:
        p.set_tiling(tile, max_items)
    return p


def oracle_plan(N, traj, W=5.5):
    p = po.Plan(N, W, alpha=1.5, precision="float")
    p.preprocess(traj)
    return p


def run(plan, x, out_shape, mode, dcw=None):
    dx = dev(x, np.complex64)
    dout = DeviceArray.zeros(out_shape, np.complex64)
    ddcw = None if dcw is None else dev(dcw, np.float32)
    plan.compute(dx, dout, ddcw, mode)
    plan.synchronize()
    return dout.numpy()


def test_c1_plain_2d_adjoint_then_forward(gpu):
    """Plain 2D workload: BACKWARDS_NC2C then FORWARDS_C2NC (128^2, 64x256 radial)."""
    say("c1 plain 2D 128^2: preprocess")
    traj, samples, dcw = c1_inputs()
    g = np.load(os.path.join(GOLD, "c1_golden.npz"))
    plan = gpu_plan([128, 128])
    plan.preprocess(dev(traj))
    assert plan.get_matrix_size_os() == [192, 192]
    assert plan.get_num_samples_frames() == (traj.shape[1], 1)
    say("c1: BACKWARDS_NC2C")
    img = run(plan, samples, (1, 1, 128, 128), CM.BACKWARDS_NC2C, dcw)
    say("c1: FORWARDS_C2NC")
    fwd = run(plan, img, (1, 1, traj.shape[1]), CM.FORWARDS_C2NC, dcw)
    op = oracle_plan([128, 128], traj)
    ref_img = op.compute(samples, (1, 1, 128, 128), po.BACKWARDS_NC2C, dcw=dcw)
    ref_fwd = op.compute(img, (1, 1, traj.shape[1]), po.FORWARDS_C2NC, dcw=dcw)
    e1, e2 = rel_l2(img, ref_img), rel_l2(fwd, ref_fwd)
    e_gold = rel_l2(img.reshape(128, 128), g["adjoint"])
    e_ndft = rel_l2(img.reshape(128, 128), g["ndft_adjoint"])
    say(f"c1: adjoint vs oracle {e1:.2e}, forward vs oracle {e2:.2e}, vs golden {e_gold:.2e}, vs FP64 NDFT {e_ndft:.2e}")
    assert e1 <= TOL and e2 <= TOL and e_gold <= TOL
    assert e_ndft < 1e-3


@pytest.mark.parametrize("oversampled", [False, True])
def test_all_compute_modes_multiframe_multicoil(gpu, oversampled):
    say(f"small 2D all modes, oversampled={oversampled}")
    g = np.load(os.path.join(GOLD, "small2d_golden.npz"))
    traj, img, smp, dcw = g["traj"], g["image"], g["samples"], g["dcw"]
    C, F, M = smp.shape
    N = [32, 24]
    plan = gpu_plan(N)
    plan.preprocess(dev(traj))
    op = oracle_plan(N, traj)
    if oversampled:
        img = tj.complex_normal(np.random.default_rng(1), (C, F, 36, 48))
    ishape = img.shape
    cases = [
        (CM.FORWARDS_C2NC, img, (C, F, M), po.FORWARDS_C2NC),
        (CM.BACKWARDS_NC2C, smp, ishape, po.BACKWARDS_NC2C),
        (CM.FORWARDS_NC2C, smp, ishape, po.FORWARDS_NC2C),
        (CM.BACKWARDS_C2NC, img, (C, F, M), po.BACKWARDS_C2NC),
    ]
    for mode, x, oshape, omode in cases:
        say(f"  mode {mode.name}")
        got = run(plan, x, oshape, mode, dcw)
        ref = op.compute(x, oshape, omode, dcw=dcw)
        err = rel_l2(got, ref)
        say(f"  mode {mode.name}: rel err {err:.2e}")
        assert err <= TOL, mode.name
        if not oversampled:
            key = {CM.FORWARDS_C2NC: "fwd_c2nc", CM.BACKWARDS_NC2C: "bwd_nc2c",
                   CM.FORWARDS_NC2C: "fwd_nc2c", CM.BACKWARDS_C2NC: "bwd_c2nc"}[mode]
            assert rel_l2(got, g[key]) <= TOL
    say("  mult_MH_M")
    dx = dev(img, np.complex64)
    dout = DeviceArray.zeros(img.shape, np.complex64)
    plan.mult_MH_M(dx, dout, dev(dcw, np.float32))
    plan.synchronize()
    ref = op.mult_MH_M(img, dcw=dcw)
    assert rel_l2(dout.numpy(), ref) <= TOL


@pytest.mark.parametrize("C,dcw_kind", [(1, "sample"), (6, "element"), (12, "sample"), (16, "element"),
                                         (17, "sample"), (40, "element")])
def test_coil_groups_2d(gpu, C, dcw_kind):
    """2D gridding at coil counts that take 4, 2 and 1 coil groups per wave (and
    a second coil chunk), two frames, per-sample and per-element dcw, and
    convolve with accumulation."""
    say(f"2D 64^2 coil groups: C={C}, dcw {dcw_kind}")
    rng = np.random.default_rng(100 + C)
    N = [64, 64]
    traj = tj.golden_radial_2d(128, 60, frames=2)
    F, M = traj.shape[0], traj.shape[1]
    plan = gpu_plan(N)
    plan.preprocess(dev(traj))
    op = oracle_plan(N, traj)
    smp = tj.complex_normal(rng, (C, F, M))
    img = tj.complex_normal(rng, (C, F, 64, 64))
    if dcw_kind == "sample":
        dcw = tj.ramp_dcf(traj)
    else:
        dcw = rng.uniform(0.2, 1.0, C * F * M).astype(np.float32)
    got = run(plan, smp, img.shape, CM.BACKWARDS_NC2C, dcw)
    assert rel_l2(got, op.compute(smp, img.shape, po.BACKWARDS_NC2C, dcw=dcw)) <= TOL
    got = run(plan, img, smp.shape, CM.FORWARDS_C2NC, dcw)
    assert rel_l2(got, op.compute(img, smp.shape, po.FORWARDS_C2NC, dcw=dcw)) <= TOL
    say(f"  C={C}: convolve with accumulation")
    grid = tj.complex_normal(rng, (C, F, 96, 96))
    acc = tj.complex_normal(rng, (C, F, M))
    d_g, d_a = dev(grid, np.complex64), dev(acc, np.complex64)
    plan.convolve(d_g, d_a, NFFT_conv_mode.C2NC, accumulate=True)
    plan.synchronize()
    ref = op.convolve(grid.reshape(-1), 0, C * F, out=acc.reshape(-1).copy(), accumulate=True)
    assert rel_l2(d_a.numpy().reshape(-1), ref) <= TOL
    d_s, d_g2 = dev(smp, np.complex64), dev(grid, np.complex64)
    plan.convolve(d_s, d_g2, NFFT_conv_mode.NC2C, accumulate=True)
    plan.synchronize()
    ref = op.convolve(smp.reshape(-1), 1, C * F, out=grid.reshape(-1).copy(), accumulate=True)
    assert rel_l2(d_g2.numpy().reshape(-1), ref) <= TOL


@pytest.mark.parametrize("W,C", [(4.5, 8), (4.5, 32), (6.5, 32), (4.0, 32), (3.5, 8)])
def test_window_widths_2d(gpu, W, C):
    """2D gridding at other kernel widths, both directions: a window of KW =
    floor(W) + 1 slots whose radius W / 2 has a fraction below 1/2 (4.5, 6.5)
    reaches one staged cell further past the home tile than one whose
    fraction is 1/2 or more (3.5, 5.5) or 0 (4.0); 8 coils take the FP32
    C2NC, 32 the matrix cores."""
    say(f"2D 64^2 W={W}, C={C}, per-sample dcw")
    rng = np.random.default_rng(300 + C + int(10 * W))
    N = [64, 64]
    traj = tj.golden_radial_2d(128, 90)
    F, M = traj.shape[0], traj.shape[1]
    plan = gpu_plan(N, W=W)
    plan.preprocess(dev(traj))
    op = oracle_plan(N, traj, W=W)
    smp = tj.complex_normal(rng, (C, F, M))
    img = tj.complex_normal(rng, (C, F, 64, 64))
    dcw = tj.ramp_dcf(traj)
    ea = rel_l2(run(plan, smp, img.shape, CM.BACKWARDS_NC2C, dcw), op.compute(smp, img.shape, po.BACKWARDS_NC2C, dcw=dcw))
    ef = rel_l2(run(plan, img, smp.shape, CM.FORWARDS_C2NC, dcw), op.compute(img, smp.shape, po.FORWARDS_C2NC, dcw=dcw))
    say(f"  rel err adjoint {ea:.2e} forward {ef:.2e}")
    assert ea <= TOL and ef <= TOL


def test_standalone_convolve_fft_deapodize(gpu):
    say("standalone convolve / fft / deapodize")
    rng = np.random.default_rng(21)
    N = [32, 40]
    M, C = 700, 2
    traj = rng.uniform(-0.5, 0.5, (1, M, 2)).astype(np.float32)
    plan = gpu_plan(N)
    plan.preprocess(dev(traj))
    op = oracle_plan(N, traj)
    grid = tj.complex_normal(rng, (C, 60, 48))
    smp = tj.complex_normal(rng, (C, M))
    # C2NC with accumulate
    d_out = dev(smp)
    plan.convolve(dev(grid), d_out, NFFT_conv_mode.C2NC, accumulate=True)
    ref = smp.reshape(-1) + op.convolve(grid.reshape(-1), 0, C)
    assert rel_l2(d_out.numpy().reshape(-1), ref) <= TOL
    # NC2C plain and accumulating
    d_g = DeviceArray.zeros(grid.shape, np.complex64)
    plan.convolve(dev(smp), d_g, NFFT_conv_mode.NC2C)
    assert rel_l2(d_g.numpy().reshape(-1), op.convolve(smp.reshape(-1), 1, C)) <= TOL
    d_g = dev(grid)
    plan.convolve(dev(smp), d_g, NFFT_conv_mode.NC2C, accumulate=True)
    assert rel_l2(d_g.numpy().reshape(-1), grid.reshape(-1) + op.convolve(smp.reshape(-1), 1, C)) <= TOL
    # centred FFT, both directions, with and without scaling
    for mode in (NFFT_fft_mode.FORWARDS, NFFT_fft_mode.BACKWARDS):
        for scale in (True, False):
            d = dev(grid)
            plan.fft(d, mode, do_scale=scale)
            ref = op.fft(grid, C, mode == NFFT_fft_mode.FORWARDS, scale)
            assert rel_l2(d.numpy(), ref) <= TOL, (mode, scale)
    for fd in (False, True):
        d = dev(grid)
        plan.deapodize(d, fourier_domain=fd)
        assert rel_l2(d.numpy(), op.deapodize(grid, C, fd)) <= TOL


def test_atomic_variant_matches_oracle(gpu):
    say("ATOMIC NC2C vs oracle")
    rng = np.random.default_rng(31)
    N = [64, 64]
    M, C = 4000, 3
    traj = tj.golden_radial_2d(128, 32)
    smp = tj.complex_normal(rng, (C, 1, traj.shape[1]))
    plan = gpu_plan(N, conv=ga.ConvolutionType.ATOMIC)
    plan.preprocess(dev(traj))
    op = oracle_plan(N, traj)
    got = run(plan, smp, (C, 1, 64, 64), CM.BACKWARDS_NC2C)
    assert rel_l2(got, op.compute(smp, (C, 1, 64, 64), po.BACKWARDS_NC2C)) <= TOL
    _ = M


@pytest.mark.parametrize("tile,max_items", [(None, 0), ((8, 8), 0), ((32, 16), 0), ((16, 16), 64)])
def test_standard_nc2c_is_deterministic_and_tiling_independent(gpu, tile, max_items):
    say(f"STANDARD NC2C determinism tile={tile} max_items={max_items}")
    rng = np.random.default_rng(41)
    N = [96, 96]
    traj = tj.golden_radial_2d(192, 48)
    C = 5
    smp = tj.complex_normal(rng, (C, 1, traj.shape[1]))
    plan = gpu_plan(N, tile=tile, max_items=max_items)
    plan.preprocess(dev(traj))
    d_s = dev(smp)
    outs = []
    for _ in range(3):
        d_o = DeviceArray.zeros((C, 1, 144, 144), np.complex64)
        plan.convolve(d_s, d_o, NFFT_conv_mode.NC2C)
        plan.synchronize()
        outs.append(d_o.numpy())
    assert all(np.array_equal(outs[0].view(np.uint32), o.view(np.uint32)) for o in outs[1:])
    op = oracle_plan(N, traj)
    assert rel_l2(outs[0].reshape(-1), op.convolve(smp.reshape(-1), 1, C)) <= TOL


def test_3d_small(gpu):
    say("3D 24^3 parity")
    rng = np.random.default_rng(51)
    N = [24, 24, 24]
    traj = tj.kooshball_3d(48, 300, seed=5)
    C = 2
    M = traj.shape[1]
    plan = gpu_plan(N)
    plan.preprocess(dev(traj))
    op = oracle_plan(N, traj)
    smp = tj.complex_normal(rng, (C, 1, M))
    img = tj.complex_normal(rng, (C, 1, 24, 24, 24))
    dcw = tj.ramp_dcf(traj, power=2)
    a = run(plan, smp, img.shape, CM.BACKWARDS_NC2C, dcw)
    assert rel_l2(a, op.compute(smp, img.shape, po.BACKWARDS_NC2C, dcw=dcw)) <= TOL
    f = run(plan, img, smp.shape, CM.FORWARDS_C2NC, dcw)
    assert rel_l2(f, op.compute(img, smp.shape, po.FORWARDS_C2NC, dcw=dcw)) <= TOL


def test_1d(gpu):
    say("1D parity")
    rng = np.random.default_rng(61)
    N = [200]
    M = 900
    traj = rng.uniform(-0.5, 0.5, (1, M, 1)).astype(np.float32)
    plan = gpu_plan(N)
    plan.preprocess(dev(traj))
    op = oracle_plan(N, traj)
    smp = tj.complex_normal(rng, (3, 1, M))
    img = tj.complex_normal(rng, (3, 1, 200))
    assert rel_l2(run(plan, smp, img.shape, CM.BACKWARDS_NC2C), op.compute(smp, img.shape, po.BACKWARDS_NC2C)) <= TOL
    assert rel_l2(run(plan, img, smp.shape, CM.FORWARDS_C2NC), op.compute(img, smp.shape, po.FORWARDS_C2NC)) <= TOL


def test_c2_multicoil_full_size(gpu):
    """32-coil 2D radial workload at full size: 256^2, 400 spokes x 512, seed 1."""
    say("c2 32-coil 256^2 full size: preprocess")
    rng = np.random.default_rng(1)
    traj = tj.golden_radial_2d(512, 400)
    M = traj.shape[1]
    C = 32
    smp = tj.complex_normal(rng, (C, 1, M))
    dcw = tj.ramp_dcf(traj)
    plan = gpu_plan([256, 256])
    plan.preprocess(dev(traj))
    say("c2: adjoint")
    img = run(plan, smp, (C, 1, 256, 256), CM.BACKWARDS_NC2C, dcw)
    say("c2: forward")
    fwd = run(plan, img, (C, 1, M), CM.FORWARDS_C2NC, dcw)
    say("c2: oracle")
    op = oracle_plan([256, 256], traj)
    e1 = rel_l2(img, op.compute(smp, (C, 1, 256, 256), po.BACKWARDS_NC2C, dcw=dcw))
    e2 = rel_l2(fwd, op.compute(img, (C, 1, M), po.FORWARDS_C2NC, dcw=dcw))
    say(f"c2: adjoint err {e1:.2e} forward err {e2:.2e}")
    assert e1 <= TOL and e2 <= TOL
    # size-independent property: adjoint identity <A x, y> = <x, A^H y>
    x = tj.complex_normal(rng, (C, 1, 256, 256))
    Ax = run(plan, x, (C, 1, M), CM.FORWARDS_C2NC)
    AHy = run(plan, smp, (C, 1, 256, 256), CM.BACKWARDS_NC2C)
    lhs = np.vdot(smp.astype(np.complex128), Ax.astype(np.complex128))
    rhs = np.vdot(AHy.astype(np.complex128), x.astype(np.complex128))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


def test_errors_on_device(gpu):
    say("error paths")
    plan = gpu_plan([64, 64])
    bad = np.zeros((1, 10, 2), np.float32)
    bad[0, 3, 0] = 0.75
    with pytest.raises(ga.GTError) as e:
        plan.preprocess(dev(bad))
    assert e.value.status == 4
    nan = np.zeros((1, 10, 2), np.float32)
    nan[0, 1, 1] = np.nan
    with pytest.raises(ga.GTError):
        plan.preprocess(dev(nan))
    x = DeviceArray.zeros((1, 64, 64), np.complex64)
    y = DeviceArray.zeros((1, 10), np.complex64)
    with pytest.raises(ga.GTError) as e:
        plan.compute(x, y, None, CM.FORWARDS_C2NC)
    assert e.value.status == 2
    plan.preprocess(dev(np.zeros((1, 10, 2), np.float32)))
    y_bad = DeviceArray.zeros((2, 10), np.complex64)
    with pytest.raises(ga.GTError) as e:
        plan.compute(x, y_bad, None, CM.FORWARDS_C2NC)
    assert e.value.status == 3
    z = DeviceArray.zeros((1, 60, 64), np.complex64)
    with pytest.raises(ga.GTError) as e:
        plan.compute(z, y, None, CM.FORWARDS_C2NC)
    assert e.value.status == 3


@pytest.mark.parametrize("N", [[64, 64], [80], [24, 24, 24]])
def test_sparse_matrix_row_normalised(gpu, N):
    """SPARSE_MATRIX: NC2C through the transpose of the row-normalised
    convolution matrix = the oracle's gridding of samples / (sum of the
    sample's window weights, from the oracle's own convolution rows); C2NC as
    STANDARD."""
    say(f"SPARSE_MATRIX {N}")
    rng = np.random.default_rng(41)
    D, M, C = len(N), 1500, 3
    traj = rng.uniform(-0.5, 0.5, (1, M, D)).astype(np.float32)
    plan = gpu_plan(N, conv=ga.ConvolutionType.SPARSE_MATRIX)
    plan.preprocess(dev(traj))
    op = oracle_plan(N, traj)
    wsum = np.array([op.conv_row(j)[1].astype(np.float64).sum() for j in range(M)])
    smp = tj.complex_normal(rng, (C, M))
    gshape = (C,) + tuple(int(n * 1.5) for n in reversed(N))
    d_g = DeviceArray.zeros(gshape, np.complex64)
    plan.convolve(dev(smp), d_g, NFFT_conv_mode.NC2C)
    plan.synchronize()
    ref = op.convolve((smp / wsum).astype(np.complex64).reshape(-1), 1, C)
    assert rel_l2(d_g.numpy().reshape(-1), ref) <= TOL
    grid = tj.complex_normal(rng, gshape)
    d_g = dev(grid)
    plan.convolve(dev(smp), d_g, NFFT_conv_mode.NC2C, accumulate=True)
    plan.synchronize()
    assert rel_l2(d_g.numpy().reshape(-1), grid.reshape(-1) + ref) <= TOL
    img_shape = (C,) + tuple(reversed(N))
    w_s = rng.uniform(0.2, 1.0, M).astype(np.float32)
    w_c = rng.uniform(0.2, 1.0, C * M).astype(np.float32)
    for name, dcw, odcw in (("none", None, 1.0 / wsum), ("per sample", w_s, w_s / wsum),
                            ("per coil", w_c, w_c / np.tile(wsum, C))):
        got = run(plan, smp, img_shape, CM.BACKWARDS_NC2C, dcw)
        want = op.compute(smp, img_shape, CM.BACKWARDS_NC2C, odcw.astype(np.float32))
        assert rel_l2(got.reshape(-1), want.reshape(-1)) <= TOL, name
    img = tj.complex_normal(rng, img_shape)
    got = run(plan, img, (C, M), CM.FORWARDS_C2NC, w_s)
    want = op.compute(img, (C, M), CM.FORWARDS_C2NC, w_s)
    assert rel_l2(got.reshape(-1), want.reshape(-1)) <= TOL
    # repeatable
    a = run(plan, smp, img_shape, CM.BACKWARDS_NC2C, w_c)
    assert np.array_equal(a, run(plan, smp, img_shape, CM.BACKWARDS_NC2C, w_c))
