"""The matrix-core 2D gridding (17 coils or more) keeps its fp16 operands in
range with powers of two: a per-wave sample scale that is set from the first
non-zero chunk and lowered, with the accumulators rescaled, when a later chunk
is larger (NC2C); a per-tile scale from the staged grid (C2NC).  These cases
drive that logic: sample and image magnitudes spread over twelve orders within
each tile, data scaled by 1e24 and 1e-20 (the range over which the FP32
oracle's own intermediates stay normal floats at this size), and all-zero
coils beside non-zero ones, against the oracle at the usual bar (relative L2
<= 1e-5).
Every case prints its name before touching the GPU.
"""
import sys

import numpy as np
import pytest

from gadgetron_amd import DeviceArray, NFFT_comp_mode as CM, NFFTPlan
from oracle import pyoracle as po
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu
N = [64, 64]
M, C = 4000, 32


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def dev(a, dtype):
    return DeviceArray.from_numpy(np.ascontiguousarray(a, dtype=dtype))


def run(plan, x, out_shape, mode, dcw):
    dout = DeviceArray.zeros(out_shape, np.complex64)
    plan.compute(dev(x, np.complex64), dout, dev(dcw, np.float32), mode)
    plan.synchronize()
    return dout.numpy()


@pytest.mark.parametrize("case", ["log-uniform", "huge", "tiny", "zero coils"])
def test_mfma_gridding_operand_range(gpu, case):
    say(f"matrix-core gridding range: {case}")
    rng = np.random.default_rng(7)
    traj = rng.uniform(-0.5, 0.5, (1, M, 2)).astype(np.float32)
    dcw = rng.uniform(0.2, 1.0, (1, M)).astype(np.float32)
    plan = NFFTPlan(N, None, W=5.5, os_factor=1.5)
    plan.preprocess(dev(traj, np.float32))
    op = po.Plan(N, 5.5, alpha=1.5, precision="float")
    op.preprocess(traj)
    smp = (rng.standard_normal((C, 1, M)) + 1j * rng.standard_normal((C, 1, M)))
    img = (rng.standard_normal((C, 1) + tuple(reversed(N))) + 1j * rng.standard_normal((C, 1) + tuple(reversed(N))))
    if case == "log-uniform":
        smp = smp * 10.0 ** rng.uniform(-6, 6, (1, 1, M))           # per sample, shared by the coils
        img = img * 10.0 ** rng.uniform(-6, 6, img.shape[2:])        # per pixel
    elif case == "huge":
        smp, img = smp * 1e24, img * 1e24  # FP32 sums stay finite (the oracle overflows near 1e30)
    elif case == "tiny":
        smp, img = smp * 1e-20, img * 1e-20  # deapodized grid values stay normal floats
    else:
        smp[::2] = 0.0
        img[1::2] = 0.0
    smp = smp.astype(np.complex64)
    img = img.astype(np.complex64)
    img_shape = (C, 1) + tuple(reversed(N))
    got = run(plan, smp, img_shape, CM.BACKWARDS_NC2C, dcw)
    want = op.compute(smp, img_shape, po.BACKWARDS_NC2C, dcw=dcw)
    assert np.all(np.isfinite(got))
    assert rel_l2(got, want) <= 1e-5, rel_l2(got, want)
    got = run(plan, img, (C, 1, M), CM.FORWARDS_C2NC, dcw)
    want = op.compute(img, (C, 1, M), po.FORWARDS_C2NC, dcw=dcw)
    assert np.all(np.isfinite(got))
    assert rel_l2(got, want) <= 1e-5, rel_l2(got, want)
    if case == "zero coils":
        assert not np.any(run(plan, smp, img_shape, CM.BACKWARDS_NC2C, dcw)[::2])
