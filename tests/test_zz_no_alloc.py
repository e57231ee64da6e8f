"""gt_nufft_compute_noalloc against gt_nufft_compute, through the C ABI.

The no-allocation entry point runs the same kernels as gt_nufft_compute with
the grid and the split tiles' partial sums in the library's fixed per-GPU
scratch array, so its results must equal gt_nufft_compute's bit for bit.  Two
calls in a row, the second with a smaller input, check that the array's
leftover contents from the first call do not leak into the second.  The
second input is also checked against the oracle at the parity bar.
"""
import ctypes
import sys

import numpy as np
import pytest

from gadgetron_amd import DeviceArray, NFFT_comp_mode as CM, NFFTPlan
from gadgetron_amd import trajectories as tj
from gadgetron_amd._lib import GTError, lib, sizes
from oracle import pyoracle as po
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-5


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def dims(shape):
    return list(reversed([int(s) for s in shape]))


def abi_compute(fn, plan, d_in, d_out, d_dcw, mode):
    """Calls gt_nufft_compute or gt_nufft_compute_noalloc (fn) by name."""
    di, do = dims(d_in.shape), dims(d_out.shape)
    f = getattr(lib(), fn)
    return f(ctypes.c_void_p(plan._h), ctypes.c_void_p(d_in.data_ptr()), sizes(di), len(di),
             ctypes.c_void_p(d_out.data_ptr()), sizes(do), len(do),
             None if d_dcw is None else ctypes.c_void_p(d_dcw.data_ptr()),
             0 if d_dcw is None else int(np.prod(d_dcw.shape)), int(mode), None)


def both(plan, x, out_shape, mode, dcw):
    d_x = DeviceArray.from_numpy(x)
    d_w = None if dcw is None else DeviceArray.from_numpy(dcw)
    d_a = DeviceArray.zeros(out_shape, np.complex64)
    d_b = DeviceArray.zeros(out_shape, np.complex64)
    assert abi_compute("gt_nufft_compute", plan, d_x, d_a, d_w, mode) == 0, lib().gt_last_error()
    st = abi_compute("gt_nufft_compute_noalloc", plan, d_x, d_b, d_w, mode)
    assert st == 0, lib().gt_last_error()
    plan.synchronize()
    return d_a.numpy(), d_b.numpy()


def test_noalloc_equals_compute_two_calls(gpu):
    rng = np.random.default_rng(11)
    # first call: 256^2 (384^2 grid), 96 golden-angle spokes x 512, 32 coils
    # (matrix-core gridding, split tiles), adjoint then forward
    say("noalloc: 256^2, 96x512, 32 coils, BACKWARDS_NC2C")
    traj = tj.golden_radial_2d(512, 96)
    M = traj.shape[1]
    dcw = tj.ramp_dcf(traj)
    plan = NFFTPlan([256, 256], None, W=5.5, os_factor=1.5)
    plan.preprocess(DeviceArray.from_numpy(traj))
    s32 = tj.complex_normal(rng, (32, 1, M))
    a, b = both(plan, s32, (32, 1, 256, 256), CM.BACKWARDS_NC2C, dcw)
    assert np.array_equal(a, b)
    say("noalloc: 256^2, 32 coils, FORWARDS_C2NC")
    a2, b2 = both(plan, a, (32, 1, M), CM.FORWARDS_C2NC, dcw)
    assert np.array_equal(a2, b2)

    # second call, smaller: the same plan with 3 coils (the FP32 kernels),
    # then a 128^2 plan with one coil, each checked against the oracle too
    say("noalloc: 256^2, 3 coils, BACKWARDS_NC2C then FORWARDS_C2NC")
    s3 = tj.complex_normal(rng, (3, 1, M))
    a, b = both(plan, s3, (3, 1, 256, 256), CM.BACKWARDS_NC2C, dcw)
    assert np.array_equal(a, b)
    a2, b2 = both(plan, b, (3, 1, M), CM.FORWARDS_C2NC, dcw)
    assert np.array_equal(a2, b2)
    op = po.Plan([256, 256], 5.5, alpha=1.5, precision="float")
    op.preprocess(traj)
    e1 = rel_l2(b, op.compute(s3, (3, 1, 256, 256), po.BACKWARDS_NC2C, dcw=dcw))
    e2 = rel_l2(b2, op.compute(b, (3, 1, M), po.FORWARDS_C2NC, dcw=dcw))
    say(f"noalloc 3 coils vs oracle: adjoint {e1:.2e}, forward {e2:.2e}")
    assert e1 <= TOL and e2 <= TOL

    say("noalloc: 128^2, 64x256, 1 coil, every mode")
    t1 = tj.golden_radial_2d(256, 64)
    M1 = t1.shape[1]
    w1 = tj.ramp_dcf(t1)
    p1 = NFFTPlan([128, 128], None, W=5.5, os_factor=1.5)
    p1.preprocess(DeviceArray.from_numpy(t1))
    s1 = tj.complex_normal(rng, (1, 1, M1))
    img1 = tj.complex_normal(rng, (1, 1, 128, 128))
    for mode, x, shp in ((CM.BACKWARDS_NC2C, s1, (1, 1, 128, 128)), (CM.FORWARDS_NC2C, s1, (1, 1, 128, 128)),
                         (CM.FORWARDS_C2NC, img1, (1, 1, M1)), (CM.BACKWARDS_C2NC, img1, (1, 1, M1))):
        a, b = both(p1, x, shp, mode, w1)
        assert np.array_equal(a, b), mode
    o1 = po.Plan([128, 128], 5.5, alpha=1.5, precision="float")
    o1.preprocess(t1)
    a, b = both(p1, s1, (1, 1, 128, 128), CM.BACKWARDS_NC2C, None)
    e = rel_l2(b, o1.compute(s1, (1, 1, 128, 128), po.BACKWARDS_NC2C))
    say(f"noalloc 128^2 vs oracle (no dcw): {e:.2e}")
    assert np.array_equal(a, b) and e <= TOL


def test_noalloc_refusals(gpu):
    say("noalloc: refusals (host checks, no device work)")
    traj = tj.golden_radial_2d(512, 16)
    M = traj.shape[1]
    plan = NFFTPlan([256, 256], None, W=5.5, os_factor=1.5)
    plan.preprocess(DeviceArray.from_numpy(traj))
    # 64 coils of a 384^2 grid need 72 MiB: more than the scratch array holds
    d_s = DeviceArray.zeros((64, 1, M), np.complex64)
    d_i = DeviceArray.zeros((64, 1, 256, 256), np.complex64)
    with pytest.raises(GTError) as e:
        plan.compute_noalloc(d_s, d_i, None, CM.BACKWARDS_NC2C)
    assert e.value.status == 5 and "scratch" in e.value.message
    # 3D plans are not taken
    t3 = np.zeros((1, 8, 3), np.float32)
    p3 = NFFTPlan([16, 16, 16], None, W=5.5, os_factor=1.5)
    p3.preprocess(DeviceArray.from_numpy(t3))
    with pytest.raises(GTError) as e:
        p3.compute_noalloc(DeviceArray.zeros((1, 1, 8), np.complex64),
                           DeviceArray.zeros((1, 1, 16, 16, 16), np.complex64), None, CM.BACKWARDS_NC2C)
    assert e.value.status == 5
    # a plan used through compute_noalloc still works through compute afterwards
    d_s = DeviceArray.from_numpy(tj.complex_normal(np.random.default_rng(1), (2, 1, M)))
    d_a = DeviceArray.zeros((2, 1, 256, 256), np.complex64)
    d_b = DeviceArray.zeros((2, 1, 256, 256), np.complex64)
    plan.compute_noalloc(d_s, d_a, None, CM.BACKWARDS_NC2C)
    plan.compute(d_s, d_b, None, CM.BACKWARDS_NC2C)
    plan.synchronize()
    assert np.array_equal(d_a.numpy(), d_b.numpy())
