"""The oracle reproduces the committed golden fixtures (tests/golden/make_golden.py),
and the fixtures state the C1 approximation error against the FP64 NDFT."""
import os

import numpy as np

from oracle import pyoracle as po
from tests.helpers import c1_inputs, rel_l2

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_c1_fixture_reproduced_by_oracle():
    g = np.load(os.path.join(HERE, "c1_golden.npz"))
    traj, samples, dcw = c1_inputs()
    assert np.array_equal(samples.reshape(-1), g["samples"])
    p = po.Plan([128, 128], 5.5, alpha=1.5, precision="float")
    p.preprocess(traj)
    adj = p.compute(g["samples"].reshape(1, 1, -1), (1, 1, 128, 128), po.BACKWARDS_NC2C, dcw=dcw)
    assert rel_l2(adj.reshape(128, 128), g["adjoint"]) < 1e-6
    fwd = p.compute(g["adjoint"].reshape(1, 1, 128, 128), (1, 1, traj.shape[1]), po.FORWARDS_C2NC, dcw=dcw)
    assert rel_l2(fwd.reshape(-1), g["forward"]) < 1e-6
    # KB(alpha 1.5, W 5.5) approximation error of the adjoint against the FP64 NDFT
    err = rel_l2(g["adjoint"], g["ndft_adjoint"])
    assert err < 1e-3


def test_small2d_fixture_reproduced_by_oracle():
    g = np.load(os.path.join(HERE, "small2d_golden.npz"))
    p = po.Plan([32, 24], 5.5, alpha=1.5, precision="float")
    p.preprocess(g["traj"])
    C, F, M = g["samples"].shape
    assert rel_l2(p.compute(g["image"], (C, F, M), po.FORWARDS_C2NC, dcw=g["dcw"]), g["fwd_c2nc"]) < 1e-6
    assert rel_l2(p.compute(g["samples"], g["image"].shape, po.BACKWARDS_NC2C, dcw=g["dcw"]), g["bwd_nc2c"]) < 1e-6
