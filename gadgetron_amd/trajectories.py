"""Trajectories, density compensation and synthetic data for the benchmark workloads.

* 2D golden-angle radial: Gadgetron's GR_ORIGINAL formula
  (toolboxes/core/gpu/radial_utilities.cu:20-44): theta_p = (p + offset) * pi / phi + pi,
  phi = (1 + sqrt 5)/2, k_{p,i} = (i - S/2)/S * (cos theta_p, sin theta_p).
* 3D kooshball: the reference has no 3D radial generator; spoke directions here
  are a generalized spiral on the sphere (Saff & Kuijlaars 1997) with a seeded
  random rotation, samples k = (i - S/2)/S * direction.
* Ramp density compensation |k|/max|k| (2D) and |k|^2/max|k|^2 (3D).
* Complex Gaussian CN(0, 1): real and imaginary parts i.i.d. N(0, 1/2).
"""
import numpy as np

GOLDEN_RATIO = (1.0 + np.sqrt(5.0)) / 2.0


def golden_radial_2d(samples_per_spoke, spokes, frames=1, offset=0, dtype=np.float32):
    """Shape (frames, spokes * samples_per_spoke, 2); x = column 0."""
    S = samples_per_spoke
    i = np.arange(S, dtype=np.float64)
    out = np.empty((frames, spokes * S, 2), dtype=np.float64)
    step = np.pi / GOLDEN_RATIO
    for f in range(frames):
        p = np.arange(spokes, dtype=np.float64) + offset + f * spokes
        theta = p * step + np.pi
        r = (i - S / 2.0) / S
        out[f, :, 0] = (r[None, :] * np.cos(theta)[:, None]).reshape(-1)
        out[f, :, 1] = (r[None, :] * np.sin(theta)[:, None]).reshape(-1)
    return out.astype(dtype)


def kooshball_3d(samples_per_spoke, spokes, seed=2, dtype=np.float32):
    """Shape (1, spokes * samples_per_spoke, 3)."""
    rng = np.random.default_rng(seed)
    n = spokes
    h = -1.0 + 2.0 * (np.arange(n) + 0.5) / n
    theta = np.arccos(h)
    phi = np.zeros(n)
    for k in range(1, n):
        phi[k] = (phi[k - 1] + 3.6 / np.sqrt(n * (1.0 - h[k] ** 2))) % (2.0 * np.pi)
    d = np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], axis=1)
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    d = d @ q.T
    S = samples_per_spoke
    r = (np.arange(S, dtype=np.float64) - S / 2.0) / S
    k = (d[:, None, :] * r[None, :, None]).reshape(-1, 3)
    k = np.clip(k, -0.5, 0.5)
    return k[None].astype(dtype)


def ramp_dcf(traj, power=1):
    k = np.linalg.norm(np.asarray(traj, dtype=np.float64), axis=-1)
    w = k ** power
    m = w.max()
    return (w / m if m > 0 else w).astype(np.float32)


def complex_normal(rng, shape, var=1.0, dtype=np.complex64):
    s = np.sqrt(var / 2.0)
    return (rng.standard_normal(shape) * s + 1j * rng.standard_normal(shape) * s).astype(dtype)
