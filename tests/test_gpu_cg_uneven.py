"""GPU parity on the batched uneven CG-SENSE workload itself: 64 independent 2D
golden-angle radial reconstructions (seed 3, N in {96,128,160,192,256}, spokes
U[32,512], 2N samples per spoke, coils U[4,32], Gaussian-bump coil maps,
10-ellipse phantom, data = A x + CN(0,1e-3)), 10 CG iterations each.

Bars: the batched solver equals the plain per-problem solver bit for bit on all
64 problems; the three smallest problems and the smallest one with 17 coils or
more at N >= 192 (its gridding runs on the matrix cores) match the oracle's
CG-SENSE (oracle/cgsense.py over the hoNFFT restatement, the same data) at
relative L2 <= 1e-4.  The fused solver (one launch per step for all problems,
every problem's gridding on the matrix cores) matches the plain solver at
relative L2 <= 1e-4 on every problem and the oracle on the same picks.
"""
import sys

import numpy as np
import pytest

import gadgetron_amd as ga
from gadgetron_amd import DeviceArray
from gadgetron_amd import trajectories as tj
from oracle import cgsense as cg
from oracle import pyoracle as po
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def uneven(gpu):
    say("uneven CG-SENSE workload: 64 problems, seed 3 (plans, operators, data)")
    probs = tj.uneven_sense_problems(64, 3)
    keep, ops, data = [], [], []
    for pr in probs:
        N, C, M = pr["N"], pr["C"], pr["M"]
        plan = ga.NFFTPlan([N, N], None, W=5.5, os_factor=1.5)
        d_traj = DeviceArray.from_numpy(pr["traj"])
        plan.preprocess(d_traj)
        d_csm = DeviceArray.from_numpy(pr["csm"])
        d_dcw = DeviceArray.from_numpy(pr["dcw"])
        op = ga.SenseOperator(plan, d_csm, d_dcw)
        d_x = DeviceArray.from_numpy(pr["x"])
        d_y = DeviceArray.zeros((C, 1, M), np.complex64)
        op.mult_M(d_x, d_y)
        plan.synchronize()
        pr["data"] = (d_y.numpy() + pr["noise"]).astype(np.complex64)
        data.append(DeviceArray.from_numpy(pr["data"]))
        ops.append(op)
        keep += [plan, d_traj, d_csm, d_dcw, d_x, d_y]
    return probs, ops, data, keep


def test_uneven_batched_equals_plain_and_oracle(uneven):
    probs, ops, data, _ = uneven
    xp = [DeviceArray.zeros((1, p["N"], p["N"]), np.complex64) for p in probs]
    xb = [DeviceArray.zeros((1, p["N"], p["N"]), np.complex64) for p in probs]
    say("plain per-problem CG, 64 problems x 10 iterations")
    plain = ga.CgBatch(ops, data, xp, iterations=10, tc_tolerance=0.0, mode="plain")
    plain.run()
    say("batched CG (hipGraph), 64 problems x 10 iterations")
    batched = ga.CgBatch(ops, data, xb, iterations=10, tc_tolerance=0.0, mode="batched")
    batched.run()
    ops[0].plan.synchronize()
    got_p = [a.numpy() for a in xp]
    got_b = [b.numpy() for b in xb]
    for i, (a, b) in enumerate(zip(got_p, got_b)):
        assert np.all(np.isfinite(a)), i
        assert np.array_equal(a, b), f"problem {i}: batched differs from plain"
    order = sorted(range(len(probs)), key=lambda i: probs[i]["M"] * probs[i]["C"])
    picks = order[:3]
    big = [i for i in order if probs[i]["C"] >= 17 and probs[i]["N"] >= 192]
    assert big, "the workload has a problem with 17 coils or more at N >= 192"
    picks.append(big[0])
    po.set_num_threads(16)
    say("fused CG (one launch per step for all problems), 64 problems x 10 iterations")
    xf = [DeviceArray.zeros((1, p["N"], p["N"]), np.complex64) for p in probs]
    fused = ga.CgBatch(ops, data, xf, iterations=10, tc_tolerance=0.0, mode="fused")
    fused.run()
    ops[0].plan.synchronize()
    got_f = [a.numpy() for a in xf]
    worst = max(rel_l2(f, a) for f, a in zip(got_f, got_p))
    say(f"fused vs plain: worst rel L2 over 64 problems {worst:.3e}")
    assert all(np.all(np.isfinite(f)) for f in got_f)
    assert worst <= 1e-4
    for i in picks:
        pr = probs[i]
        say(f"oracle CG-SENSE problem {i}: N {pr['N']}, {pr['spokes']} spokes, {pr['C']} coils")
        oplan = po.Plan([pr["N"], pr["N"]], 5.5, alpha=1.5, precision="float")
        oplan.preprocess(pr["traj"])
        oop = cg.SenseOp(oplan, pr["csm"], pr["dcw"])
        xo, _ = cg.cg_solve(oop, pr["data"], 10, tc_tolerance=0.0)
        err = rel_l2(got_p[i], xo)
        errf = rel_l2(got_f[i], xo)
        say(f"problem {i}: rel L2 vs oracle {err:.3e} (plain), {errf:.3e} (fused)")
        assert err <= 1e-4 and errf <= 1e-4, (i, err, errf)
    # a second run of the same batch gives the same result (the graph replays)
    fused.run()
    ops[0].plan.synchronize()
    assert all(np.array_equal(a.numpy(), f) for a, f in zip(xf, got_f))


def test_fused_refusals(uneven):
    """GT_CG_FUSED takes an identity regulariser only: a preconditioner is
    refused on the host, before any device work."""
    import ctypes

    from gadgetron_amd._lib import lib
    from gadgetron_amd.sense import _desc
    probs, ops, data, _ = uneven
    say("fused batch refusals")
    xs = [DeviceArray.zeros((1, p["N"], p["N"]), np.complex64) for p in probs[:2]]
    w = DeviceArray.zeros((1,), np.complex64)
    d = _desc(2, 0.0, 1.0, 0.0, w)
    PA = ctypes.c_void_p * 2
    h = ctypes.c_void_p()
    st = lib().gt_cgsense_batch_create(ctypes.byref(h), PA(*[o._h for o in ops[:2]]),
                                       PA(*[a.data_ptr() for a in data[:2]]), PA(*[a.data_ptr() for a in xs]), 2,
                                       ctypes.byref(d), 2, None)
    assert st == 5 and not h.value
