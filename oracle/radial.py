"""CPU restatement of the reference's radial trajectory, radial DCW and
iterative SDC functions -- test infrastructure only (the checker of
include/gt_radial.h; the library never calls it).

Follows:
  core/gpu/radial_utilities.cu:13-38   golden-ratio trajectory
  core/gpu/radial_utilities.cu:67-80   variable-angle trajectory (angle per frame)
  core/gpu/radial_utilities.cu:113-130 fixed-angle trajectory
  core/gpu/radial_utilities.cu:201-330 radial Voronoi-cell weights
  mri/sdc/SDC.hpp:36-79, gpu/cuSDC.cu:8-37  estimate_dcw (Pipe-Menon)
  nfft/ConvolutionKernel.h:270-375, .hpp:307-400  Jinc kernel
The radial weights are computed in float32, as the reference does, so that the
discrete neighbour choices match; everything else runs in float64.
"""
import numpy as np

GR_SMALLEST, GR_ORIGINAL = 0, 1
JINC_NORM = 0.96960938
JINC_POLY = (0.999923599661866, 0.014023570134078, -2.969947516849635, 0.891173008769047, 2.326043034495162,
             -1.270644932677047)


def golden_step(mode):
    return np.pi * (3.0 - np.sqrt(5.0)) * 0.5 if mode == GR_SMALLEST else np.pi / ((np.sqrt(5.0) + 1.0) * 0.5)


def _profiles(S, angles):
    r = (np.arange(S, dtype=np.float64) - S / 2.0) / S
    return np.stack([(r[None, :] * np.cos(angles)[:, None]).ravel(), (r[None, :] * np.sin(angles)[:, None]).ravel()],
                    axis=1)


def trajectory_golden_ratio_2d(S, P, F=1, profile_offset=0, mode=GR_ORIGINAL):
    """(F, P*S, 2): profile p over all frames at (p + offset) step + pi (exact
    angles; the reference forms them in float32)."""
    p = np.arange(P * F, dtype=np.float64)
    return _profiles(S, (p + profile_offset) * golden_step(mode) + np.pi).reshape(F, P * S, 2)


def trajectory_fixed_angle_2d(S, P, F=1, angular_offset=0.0):
    """(F, P*S, 2): profile p of frame f at (p + f / F) / P pi + offset + pi."""
    out = []
    for f in range(F):
        lp = np.arange(P, dtype=np.float64)
        out.append(_profiles(S, (lp + f / F) / P * np.pi + angular_offset + np.pi))
    return np.stack(out)


def trajectory_variable_angle_2d(angles, S, P):
    """Every profile of frame f at angles[f] (the reference kernel's indexing)."""
    return np.stack([_profiles(S, np.full(P, a, dtype=np.float64)) for a in np.asarray(angles, np.float64)])


def radial_dcw_2d(S, P, alpha, os_inv, golden=True, profile_offset=0, mode=GR_ORIGINAL):
    """Voronoi-cell weights, (P*S,) float32, computed in float32."""
    f32 = np.float32
    S_, alpha, os_inv = f32(S), f32(alpha), f32(os_inv)
    bias = S_ * f32(0.5)
    step = f32(golden_step(mode))
    off = f32(profile_offset)

    def direction(prof):
        prof = f32(prof)
        a = (prof + off) * step if golden else prof * (f32(1) / f32(P)) * f32(np.pi)
        return f32(np.cos(np.float64(a))), f32(np.sin(np.float64(a)))

    dirs = [direction(i) for i in range(P)]
    out = np.zeros(P * S, np.float32)
    big = np.finfo(np.float32).max
    with np.errstate(over="ignore", invalid="ignore"):
        for prof in range(P):
            cs0, sn0 = dirs[prof]
            normal = np.array([-sn0, cs0], f32)
            for idx in range(S):
                if idx == S // 2:
                    rad = (alpha * os_inv) * f32(0.5)
                    out[prof * S + idx] = rad * rad * f32(np.pi) / f32(P)
                    continue
                si = f32(idx)
                ctr = alpha * ((si - bias) * os_inv)
                ctr_i = alpha * ((S_ - si - bias) * os_inv)
                prv = alpha * ((si - bias - f32(1)) * os_inv)
                prv_i = alpha * ((S_ - (si - f32(1)) - bias) * os_inv)
                nxt = alpha * ((si - bias + f32(1)) * os_inv)
                nxt_i = alpha * ((S_ - (si + f32(1)) - bias) * os_inv)
                pos = np.array([ctr * cs0, ctr * sn0], f32)
                a = [np.array([big, big], f32)] * 3
                b = [np.array([big, big], f32)] * 3
                for i in range(P):
                    if i == prof:
                        continue
                    cs, sn = dirs[i]
                    for sc in ((prv, ctr, nxt), (prv_i, ctr_i, nxt_i)):
                        q = [np.array([s * cs, s * sn], f32) for s in sc]
                        d = q[1] - pos
                        dn = f32(d[0] * normal[0] + d[1] * normal[1])
                        n2 = f32(d[0] * d[0] + d[1] * d[1])
                        if dn > 0:
                            e = a[1] - pos
                            if n2 < f32(e[0] * e[0] + e[1] * e[1]):
                                a = q
                        else:
                            e = b[1] - pos
                            if n2 < f32(e[0] * e[0] + e[1] * e[1]):
                                b = q
                pp = np.array([prv * cs0, prv * sn0], f32)
                pn = np.array([nxt * cs0, nxt * sn0], f32)
                poly = [pp, a[0], a[1], a[2], pn, b[2], b[1], b[0]]
                poly = [f32(0.5) * (pos + v) for v in poly]
                area = f32(0)
                for k in range(8):
                    u, v = poly[k], poly[(k + 1) % 8]
                    area = f32(area + (u[0] * v[1] - v[0] * u[1]))
                out[prof * S + idx] = abs(f32(0.5) * area)
    return out


def jinc(t):
    """The Jinc fit in t = distance / radius; zero from JINC_NORM on, clipped at 0."""
    t = np.asarray(t, np.float64)
    v = sum(c * t ** i for i, c in enumerate(JINC_POLY))
    return np.where(t >= JINC_NORM, 0.0, np.maximum(v, 0.0))


def sdc_grid_size(matrix_size, os_factor):
    """matrix_size * os_factor (truncated), rounded up to multiples of 32."""
    return [((int(np.float32(m) * np.float32(os_factor)) + 31) // 32) * 32 for m in matrix_size]


def _sdc_pairs(p, nos, r0):
    """Per sample: the window cells (wrapped linear index) and their weights."""
    n, D = p.shape
    lo = np.ceil(p - r0).astype(int)
    hi = np.floor(p + r0).astype(int)
    cells, weights = [], []
    for j in range(n):
        axes = [np.arange(lo[j, d], hi[j, d] + 1) for d in range(D)]
        grid = np.meshgrid(*axes, indexing="ij")
        u = np.stack([g.ravel() for g in grid], axis=1)
        dist = np.sqrt(np.sum((u - p[j]) ** 2, axis=1))
        w = jinc(dist / r0)
        lin = np.zeros(len(u), np.int64)
        st = 1
        for d in range(D):
            lin += np.mod(u[:, d], nos[d]) * st
            st *= nos[d]
        cells.append(lin)
        weights.append(w)
    return cells, weights


def estimate_dcw(traj, matrix_size, os_factor=2.1, num_iterations=10, kernel_width=5.5, initial_dcw=None):
    """Pipe-Menon: dcw <- dcw / C2NC(NC2C(dcw)), 0 where the sum is 0."""
    k = np.asarray(traj, np.float64).reshape(-1, len(matrix_size))
    nos = sdc_grid_size(matrix_size, os_factor)
    p = (k + 0.5) * np.array(nos, np.float64)
    r0 = JINC_NORM * kernel_width
    cells, weights = _sdc_pairs(p, nos, r0)
    G = int(np.prod(nos))
    dcw = np.ones(len(k)) if initial_dcw is None else np.asarray(initial_dcw, np.float64).copy()
    for _ in range(num_iterations):
        grid = np.zeros(G)
        for j in range(len(k)):
            np.add.at(grid, cells[j], weights[j] * dcw[j])
        tmp = np.array([np.dot(weights[j], grid[cells[j]]) for j in range(len(k))])
        dcw = np.where(tmp == 0.0, 0.0, dcw / np.where(tmp == 0.0, 1.0, tmp))
    return dcw
