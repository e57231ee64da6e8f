"""GPU parity of the 3D gridding path (register-tile NC2C, staged C2NC).

Small and medium cases compare with the oracle (the CPU hoNFFT restatement) at
relative L2 <= 1e-5.  The full 3D kooshball workload (128^3, 192^3 grid,
20000 spokes x 256, 8 coils) is too large for the oracle's whole transform, so
there the 3D kernels are checked against the generic per-sample kernels of the
same library (themselves pinned against the oracle by the small cases), the
oracle's C2NC on a subset of the samples, and the adjoint identity
<C2NC g, s> = <g, NC2C s>.  Every case prints its name before touching the GPU.
"""
import sys

import numpy as np
import pytest

import gadgetron_amd as ga
from gadgetron_amd import DeviceArray, NFFT_comp_mode as CM, NFFT_conv_mode, NFFTPlan
from gadgetron_amd import trajectories as tj
from oracle import pyoracle as po
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-5


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def dev(a, dtype=None):
    return DeviceArray.from_numpy(np.ascontiguousarray(a, dtype=dtype))


def frames_traj(spokes, S, F, seed):
    return np.concatenate([tj.kooshball_3d(S, spokes, seed=seed + f) for f in range(F)], axis=0)


def run(plan, x, out_shape, mode, dcw=None):
    dx = dev(x, np.complex64)
    dout = DeviceArray.zeros(out_shape, np.complex64)
    ddcw = None if dcw is None else dev(dcw, np.float32)
    plan.compute(dx, dout, ddcw, mode)
    plan.synchronize()
    return dout.numpy()


def conv(plan, x, out_shape, mode, acc=None):
    dx = dev(x, np.complex64)
    dout = DeviceArray.zeros(out_shape, np.complex64) if acc is None else dev(acc, np.complex64)
    plan.convolve(dx, dout, mode, accumulate=acc is not None)
    plan.synchronize()
    return dout.numpy()


def check_both_ways(N, traj, C, dcw, seed, max_items=0, W=5.5):
    F, M = traj.shape[0], traj.shape[1]
    rng = np.random.default_rng(seed)
    plan = NFFTPlan(N, None, W=W, os_factor=1.5)
    if max_items:
        plan.set_tiling(None, max_items)
    plan.preprocess(dev(traj))
    op = po.Plan(N, W, alpha=1.5, precision="float")
    op.preprocess(traj)
    smp = tj.complex_normal(rng, (C, F, M))
    img = tj.complex_normal(rng, (C, F) + tuple(N[::-1]))
    a = run(plan, smp, img.shape, CM.BACKWARDS_NC2C, dcw)
    ea = rel_l2(a, op.compute(smp, img.shape, po.BACKWARDS_NC2C, dcw=dcw))
    f = run(plan, img, smp.shape, CM.FORWARDS_C2NC, dcw)
    ef = rel_l2(f, op.compute(img, smp.shape, po.FORWARDS_C2NC, dcw=dcw))
    say(f"  rel err adjoint {ea:.2e} forward {ef:.2e}")
    assert ea <= TOL and ef <= TOL


def test_3d_eight_coils_two_frames_sample_dcw(gpu):
    say("3D 32^3, 8 coils, 2 frames, per-sample dcw")
    traj = frames_traj(180, 64, 2, seed=11)
    check_both_ways([32, 32, 32], traj, 8, tj.ramp_dcf(traj, power=2), seed=71)


def test_3d_eleven_coils_anisotropic_element_dcw(gpu):
    say("3D 24 x 32 x 40, 11 coils (two coil chunks), per-element dcw")
    traj = tj.kooshball_3d(80, 150, seed=12)
    M = traj.shape[1]
    rng = np.random.default_rng(72)
    dcw = rng.uniform(0.2, 1.0, 11 * M).astype(np.float32)
    check_both_ways([24, 32, 40], traj, 11, dcw, seed=73)


def test_3d_no_dcw_one_coil(gpu):
    say("3D 28^3, 1 coil, no dcw")
    traj = tj.kooshball_3d(56, 120, seed=13)
    check_both_ways([28, 28, 28], traj, 1, None, seed=74)


def test_3d_split_tiles(gpu):
    say("3D 32^3, 8 coils: tiles split into 40-entry items (partials + reduce)")
    traj = tj.kooshball_3d(64, 400, seed=14)
    check_both_ways([32, 32, 32], traj, 8, tj.ramp_dcf(traj, power=2), seed=75, max_items=40)


@pytest.mark.parametrize("N,C,F", [([64, 96, 64], 3, 2), ([128, 64, 96], 9, 1), ([64, 192, 64], 2, 1)])
def test_3d_band_fft_all_modes(gpu, N, C, F):
    # grid sizes 96 / 144 / 192, 192 / 96 / 144 and 96 / 288 / 96: the
    # band-limited 3D transforms of the library (x pass with pad or crop, y
    # over the z band, z over every line), with 32-line workgroups along y and
    # z up to length 240 and 16-line ones above
    say(f"3D {N[0]} x {N[1]} x {N[2]}, {C} coils, {F} frames: every compute mode")
    traj = frames_traj(120, 64, F, seed=31)
    M = traj.shape[1]
    rng = np.random.default_rng(91)
    plan = NFFTPlan(N, None, W=5.5, os_factor=1.5)
    plan.preprocess(dev(traj))
    op = po.Plan(N, 5.5, alpha=1.5, precision="float")
    op.preprocess(traj)
    smp = tj.complex_normal(rng, (C, F, M))
    img = tj.complex_normal(rng, (C, F) + tuple(N[::-1]))
    dcw = tj.ramp_dcf(traj, power=2)
    cases = [
        (CM.FORWARDS_C2NC, img, (C, F, M), po.FORWARDS_C2NC),
        (CM.BACKWARDS_NC2C, smp, img.shape, po.BACKWARDS_NC2C),
        (CM.FORWARDS_NC2C, smp, img.shape, po.FORWARDS_NC2C),
        (CM.BACKWARDS_C2NC, img, (C, F, M), po.BACKWARDS_C2NC),
    ]
    for mode, x, oshape, omode in cases:
        err = rel_l2(run(plan, x, oshape, mode, dcw), op.compute(x, oshape, omode, dcw=dcw))
        say(f"  mode {mode.name}: rel err {err:.2e}")
        assert err <= TOL, mode.name


@pytest.mark.parametrize("conv,W", [("ATOMIC", 5.5), ("STANDARD", 8.0)])
def test_3d_band_lengths_generic_kernels(gpu, conv, W):
    # 96^3 grids (band-FFT lengths) with gridding outside the 3D hot path:
    # atomic NC2C, and a window of 8 cells (the generic kernels)
    say(f"3D 64^3, 3 coils, {conv}, W {W}: every compute mode")
    traj = tj.kooshball_3d(64, 90, seed=32)
    M = traj.shape[1]
    rng = np.random.default_rng(92)
    plan = NFFTPlan([64, 64, 64], None, W=W, os_factor=1.5, conv=getattr(ga.ConvolutionType, conv))
    plan.preprocess(dev(traj))
    op = po.Plan([64, 64, 64], W, alpha=1.5, precision="float")
    op.preprocess(traj)
    C = 3
    smp = tj.complex_normal(rng, (C, 1, M))
    img = tj.complex_normal(rng, (C, 1, 64, 64, 64))
    cases = [
        (CM.FORWARDS_C2NC, img, (C, 1, M), po.FORWARDS_C2NC),
        (CM.BACKWARDS_NC2C, smp, img.shape, po.BACKWARDS_NC2C),
        (CM.FORWARDS_NC2C, smp, img.shape, po.FORWARDS_NC2C),
        (CM.BACKWARDS_C2NC, img, (C, 1, M), po.BACKWARDS_C2NC),
    ]
    for mode, x, oshape, omode in cases:
        err = rel_l2(run(plan, x, oshape, mode), op.compute(x, oshape, omode))
        say(f"  mode {mode.name}: rel err {err:.2e}")
        assert err <= TOL, mode.name


@pytest.mark.parametrize("W", [3.5, 4.5, 6.5])
def test_3d_window_widths(gpu, W):
    # windows of 4 and 5 cells: matrix-core C2NC tiles of 5^3 and 4^3 cells
    # (more new planes per tile than waves); 7 cells: the staged FP32 C2NC
    say(f"3D 32 x 28 x 36, 8 coils, W {W}, per-sample dcw")
    traj = tj.kooshball_3d(64, 260, seed=16)
    check_both_ways([32, 28, 36], traj, 8, tj.ramp_dcf(traj, power=2), seed=77, W=W)


def test_3d_nc2c_repeatable_and_matches_generic(gpu):
    say("3D 40^3, 8 coils: NC2C bitwise repeatable, and equal to the generic kernels")
    N = [40, 40, 40]
    traj = tj.kooshball_3d(80, 500, seed=15)
    M = traj.shape[1]
    rng = np.random.default_rng(76)
    smp = tj.complex_normal(rng, (8, 1, M))
    grid_shape = (8, 1, 60, 60, 60)
    fast = NFFTPlan(N, None, W=5.5, os_factor=1.5)
    fast.preprocess(dev(traj))
    gen = NFFTPlan(N, None, W=5.5, os_factor=1.5)
    gen.set_tiling([8, 8, 8], 0)  # a requested tile selects the generic kernels
    gen.preprocess(dev(traj))
    outs = [conv(fast, smp, grid_shape, NFFT_conv_mode.NC2C) for _ in range(3)]
    assert all(np.array_equal(outs[0].view(np.uint32), o.view(np.uint32)) for o in outs[1:])
    g = conv(gen, smp, grid_shape, NFFT_conv_mode.NC2C)
    assert rel_l2(outs[0], g) <= TOL
    grid = tj.complex_normal(rng, grid_shape)
    assert rel_l2(conv(fast, grid, smp.shape, NFFT_conv_mode.C2NC), conv(gen, grid, smp.shape, NFFT_conv_mode.C2NC)) <= TOL


def test_3d_convolve_accumulate(gpu):
    say("3D 32^3, 3 coils: convolve with accumulate, both directions")
    N = [32, 32, 32]
    traj = tj.kooshball_3d(64, 160, seed=16)
    M = traj.shape[1]
    rng = np.random.default_rng(77)
    plan = NFFTPlan(N, None, W=5.5, os_factor=1.5)
    plan.preprocess(dev(traj))
    op = po.Plan(N, 5.5, alpha=1.5, precision="float")
    op.preprocess(traj)
    grid = tj.complex_normal(rng, (3, 1, 48, 48, 48))
    smp0 = tj.complex_normal(rng, (3, 1, M))
    got = conv(plan, grid, smp0.shape, NFFT_conv_mode.C2NC, acc=smp0)
    ref = op.convolve(grid.reshape(-1), 0, 3, out=smp0.reshape(-1).copy(), accumulate=True)
    assert rel_l2(got.reshape(-1), ref) <= TOL
    smp = tj.complex_normal(rng, (3, 1, M))
    got = conv(plan, smp, grid.shape, NFFT_conv_mode.NC2C, acc=grid)
    ref = op.convolve(smp.reshape(-1), 1, 3, out=grid.reshape(-1).copy(), accumulate=True)
    assert rel_l2(got.reshape(-1), ref) <= TOL


def _dot64(a, b, chunk=1 << 24):
    a, b = a.reshape(-1), b.reshape(-1)
    s = 0j
    for i in range(0, a.size, chunk):
        s += np.vdot(a[i:i + chunk].astype(np.complex128), b[i:i + chunk].astype(np.complex128))
    return s


def test_3d_kooshball_full_size(gpu):
    """3D kooshball at full size: 128^3 (192^3 grid), 20000 spokes x 256, 8 coils, seed 2."""
    say("c3 full size: preprocess (3D kernels and generic kernels)")
    N = [128, 128, 128]
    traj = tj.kooshball_3d(256, 20000, seed=2)
    M = traj.shape[1]
    C = 8
    fast = NFFTPlan(N, None, W=5.5, os_factor=1.5)
    fast.preprocess(dev(traj))
    gen = NFFTPlan(N, None, W=5.5, os_factor=1.5)
    gen.set_tiling([8, 8, 8], 0)
    gen.preprocess(dev(traj))
    rng = np.random.default_rng(2)
    grid_shape = (C, 1, 192, 192, 192)
    grid = tj.complex_normal(rng, grid_shape)
    say("c3 full size: C2NC")
    y = conv(fast, grid, (C, 1, M), NFFT_conv_mode.C2NC)
    e = rel_l2(y, conv(gen, grid, (C, 1, M), NFFT_conv_mode.C2NC))
    say(f"  C2NC vs generic {e:.2e}")
    assert e <= TOL
    say("c3 full size: C2NC vs the oracle on the first 16 spokes")
    sub = 16 * 256
    op = po.Plan(N, 5.5, alpha=1.5, precision="float")
    op.preprocess(np.ascontiguousarray(traj[:, :sub]))
    ref = op.convolve(grid.reshape(-1), 0, C).reshape(C, sub)
    e = rel_l2(y[:, 0, :sub], ref)
    say(f"  C2NC vs oracle (subset) {e:.2e}")
    assert e <= TOL
    say("c3 full size: NC2C")
    s = tj.complex_normal(rng, (C, 1, M))
    g = conv(fast, s, grid_shape, NFFT_conv_mode.NC2C)
    e = rel_l2(g, conv(gen, s, grid_shape, NFFT_conv_mode.NC2C))
    say(f"  NC2C vs generic {e:.2e}")
    assert e <= TOL
    lhs, rhs = _dot64(y, s), _dot64(grid, g)
    e = abs(lhs - rhs) / abs(lhs)
    say(f"  adjoint identity {e:.2e}")
    assert e <= 2e-5
