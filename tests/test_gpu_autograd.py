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


def check(N, C, spokes, S, seed, ncoord=64):
    rng = np.random.default_rng(seed)
    traj = tj.golden_radial_2d(S, spokes)                       # (1, M, 2)
    M = traj.shape[1]
    pert = np.clip(traj + rng.normal(0, 1e-3, traj.shape), -0.5, 0.5).astype(np.float32)
    x = tj.complex_normal(rng, (C, 1, N, N))
    ref = po.Plan([N, N], W, alpha=ALPHA, precision="double")
    ref.preprocess(pert.astype(np.float64))
    y = ref.compute(x.astype(np.complex128), (C, 1, M), po.FORWARDS_C2NC).astype(np.complex64)

    dev = torch.device("cuda", 0)
    xt = torch.tensor(x, device=dev, requires_grad=True)
    kt = torch.tensor(traj, device=dev, requires_grad=True)
    yt = torch.tensor(y, device=dev)
    loss = (nufft(xt, kt, [N, N]) - yt).abs().pow(2).sum()
    loss.backward()
    torch.cuda.synchronize()
    gx = xt.grad.cpu().numpy()
    gk = kt.grad.cpu().numpy()

    # image gradient: 2 A^H (A x - y) in FP64
    ref.preprocess(traj.astype(np.float64))
    Ax = ref.compute(x.astype(np.complex128), (C, 1, M), po.FORWARDS_C2NC)
    resid = Ax - y.astype(np.complex128)
    gx_ref = 2.0 * ref.compute(resid, (C, 1, N, N), po.BACKWARDS_NC2C)
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
            yj = one.compute(x.astype(np.complex128), (C, 1, 1), po.FORWARDS_C2NC)[:, 0, 0]
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


def test_autograd_differentiable_workload(gpu):
    say("autograd: 256^2, 16 coils, 128 golden-angle spokes x 512, seed 4")
    ex, ek = check(256, 16, 128, 512, seed=4)
    print(f"image grad rel err {ex:.2e}, trajectory grad max rel err {ek.max():.2e} over {len(ek)} coords")
    assert len(ek) >= 48
    assert ex <= 1e-3
    assert ek.max() <= 1e-3
