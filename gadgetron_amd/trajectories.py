"""Trajectories, density compensation and synthetic data for the benchmark workloads.

* 2D golden-angle radial: Gadgetron's GR_ORIGINAL formula
  (toolboxes/core/gpu/radial_utilities.cu:20-44): theta_p = (p + offset) * pi / phi + pi,
  phi = (1 + sqrt 5)/2, k_{p,i} = (i - S/2)/S * (cos theta_p, sin theta_p).
* 3D kooshball: the reference has no 3D radial generator; spoke directions here
  are a generalized spiral on the sphere (Saff & Kuijlaars 1997) with a seeded
  random rotation, samples k = (i - S/2)/S * direction.
* Ramp density compensation |k|/max|k| (2D) and |k|^2/max|k|^2 (3D).
* Complex Gaussian CN(0, 1): real and imaginary parts i.i.d. N(0, 1/2).
* Batched uneven CG-SENSE problems: a Shepp-Logan-like sum of random ellipses,
  normalised complex Gaussian-bump coil maps (sigma N/3) and CN(0, 1e-3) noise,
  drawn from one seeded stream so that every user of the workload sees the
  same 64 problems.
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


def ellipse_phantom(rng, N, n=10):
    """Sum of n random ellipses (Shepp-Logan-like) on an N x N grid."""
    y, x = np.mgrid[-1:1:N * 1j, -1:1:N * 1j]
    img = np.zeros((N, N), np.float64)
    for _ in range(n):
        cx, cy = rng.uniform(-0.6, 0.6, 2)
        a, b = rng.uniform(0.05, 0.45, 2)
        th = rng.uniform(0, np.pi)
        xr = (x - cx) * np.cos(th) + (y - cy) * np.sin(th)
        yr = -(x - cx) * np.sin(th) + (y - cy) * np.cos(th)
        img[(xr / a) ** 2 + (yr / b) ** 2 <= 1.0] += rng.uniform(0.1, 1.0)
    return img


def gaussian_csm(rng, C, N):
    """Normalised complex Gaussian bumps, sigma N/3, at random centres: (C, N, N)."""
    y, x = np.mgrid[0:N, 0:N]
    maps = []
    for _ in range(C):
        cy, cx = rng.uniform(0, N, 2)
        maps.append(np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * (N / 3) ** 2)) * np.exp(1j * rng.uniform(0, 2 * np.pi)))
    maps = np.array(maps)
    maps /= np.sqrt(np.sum(np.abs(maps) ** 2, axis=0, keepdims=True))
    return maps.astype(np.complex64)


def uneven_sense_problems(nprob=64, seed=3, sizes=(96, 128, 160, 192, 256), spokes=(32, 512), coils=(4, 32),
                          noise=1e-3):
    """The batched uneven CG-SENSE workload: nprob independent 2D golden-angle
    radial problems with N drawn from `sizes`, spokes uniform in [32, 512],
    2N samples per spoke and coils uniform in [4, 32].  Per problem (in draw
    order): N, spokes, C, the coil maps, the 10-ellipse phantom and the noise
    CN(0, noise) of shape (C, 1, M) to add to A x.  Returns a list of dicts
    with N, spokes, C, M, traj (1, M, 2), dcw = sqrt(ramp DCF) (1, M), csm
    (C, N, N), x (1, N, N) complex64 and noise (C, 1, M) complex64."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nprob):
        N = int(rng.choice(list(sizes)))
        ns = int(rng.integers(spokes[0], spokes[1] + 1))
        C = int(rng.integers(coils[0], coils[1] + 1))
        traj = golden_radial_2d(2 * N, ns)
        M = traj.shape[1]
        csm = gaussian_csm(rng, C, N)
        x = ellipse_phantom(rng, N).astype(np.complex64)[None]
        nz = (noise * (rng.standard_normal((C, 1, M)) + 1j * rng.standard_normal((C, 1, M))) / np.sqrt(2)).astype(
            np.complex64)
        out.append(dict(N=N, spokes=ns, C=C, M=M, traj=traj, dcw=np.sqrt(ramp_dcf(traj)).astype(np.float32),
                        csm=csm, x=x, noise=nz))
    return out
