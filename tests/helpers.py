"""Shared helpers for the test suite (inputs of the benchmark workloads, error norms)."""
import numpy as np

from gadgetron_amd import trajectories as tj


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def c1_inputs():
    """Plain 2D NFFT, single coil: 128^2, alpha 1.5, W 5.5, 64 golden-angle spokes x 256
    samples, samples CN(0,1) from seed 0, ramp DCF."""
    rng = np.random.default_rng(0)
    traj = tj.golden_radial_2d(256, 64)            # (1, 16384, 2)
    samples = tj.complex_normal(rng, (1, 1, traj.shape[1]))   # [M, F=1, C=1]
    dcw = tj.ramp_dcf(traj)                        # (1, M)
    return traj, samples, dcw


def random_problem(rng, N, M, F=1, C=1, D=2):
    traj = rng.uniform(-0.5, 0.5, (F, M, D)).astype(np.float32)
    img = tj.complex_normal(rng, (C, F) + tuple(reversed(N)))
    smp = tj.complex_normal(rng, (C, F, M))
    return traj, img, smp
