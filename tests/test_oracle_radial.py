"""Pins of the radial trajectory / radial DCW / SDC oracle (oracle/radial.py) on
the CPU, against what the mathematics and the reference's own comments fix:
  * the golden-ratio steps are the angles radial_utilities.h:11-12 states;
  * fixed-angle profiles are pi / P apart, frames interleaved by pi / (P F);
  * an interior fixed-angle Voronoi cell has the closed-form area
    delta * rho * sin(pi / P), the centre sample pi (alpha / 2 os)^2 / P, and
    the cells tile the sampled disc;
  * the Jinc fit starts at 1, stays below it, decreases, and reaches 0 at its cut-off;
  * the SDC iteration equals its dense-matrix form, leaves uniform Cartesian
    sampling uniform, and gives radial sampling weights proportional to |k|.
"""
import numpy as np

from gadgetron_amd import trajectories as tj
from oracle import radial as R


def test_golden_ratio_steps_and_host_generator():
    assert abs(np.degrees(R.golden_step(R.GR_ORIGINAL)) - 111.2461) < 1e-4
    assert abs(np.degrees(R.golden_step(R.GR_SMALLEST)) - 68.7539) < 1e-4
    # the host generator used by the benches and tests follows the same formula
    assert np.allclose(R.trajectory_golden_ratio_2d(64, 10, 2), tj.golden_radial_2d(64, 10, frames=2), atol=1e-7)


def test_fixed_angle_geometry():
    S, P, F = 32, 8, 3
    t = R.trajectory_fixed_angle_2d(S, P, F)
    ang = np.arctan2(t[:, ::S, 1][:, :], t[:, ::S, 0])  # angle of each profile's first sample
    d = np.diff(np.unwrap(ang, axis=1), axis=1)
    assert np.allclose(np.abs(d), np.pi / P, atol=1e-12)
    assert np.allclose(np.abs(np.unwrap(ang[1] - ang[0])), np.pi / (P * F), atol=1e-12)
    r = np.linalg.norm(t, axis=-1)
    assert r.max() <= 0.5 + 1e-12


def test_radial_dcw_closed_forms():
    S, P, alpha, osinv = 64, 16, 1.3, 0.8
    w = R.radial_dcw_2d(S, P, alpha, osinv, golden=False).reshape(P, S)
    delta = alpha * osinv
    assert np.allclose(w[:, S // 2], np.pi * (delta / 2) ** 2 / P, rtol=1e-6)
    rho = np.abs(np.arange(S) - S / 2) * delta
    sel = (np.arange(S) >= 2) & (np.abs(np.arange(S) - S / 2) >= 2) & (np.arange(S) <= S - 3)
    assert np.allclose(w[:, sel], delta * rho[sel] * np.sin(np.pi / P), rtol=2e-5)
    disc = np.pi * (delta * S / 2) ** 2
    for golden in (False, True):
        s = R.radial_dcw_2d(S, P, alpha, osinv, golden=golden).sum()
        assert abs(s / disc - 1.0) < 0.02


def test_jinc_fit():
    assert abs(R.jinc(0.0) - 1.0) < 1e-4
    t = np.linspace(0.01, R.JINC_NORM, 400)  # the fit's small linear term lifts it by < 2e-5 near 0
    assert np.all(np.diff(R.jinc(t)) <= 1e-12) and R.jinc(np.linspace(0, 0.01, 50)).max() < 1.0
    edge = sum(c * (R.JINC_NORM - 1e-9) ** i for i, c in enumerate(R.JINC_POLY))
    assert abs(edge) < 1e-3
    assert R.jinc(R.JINC_NORM) == 0.0 and R.jinc(1.2) == 0.0


def test_sdc_grid_rounding():
    assert R.sdc_grid_size([16, 16], 2.1) == [64, 64]      # 33 -> 64
    assert R.sdc_grid_size([128], 2.1) == [288]            # 268 -> 288


def test_sdc_equals_dense_matrix_form():
    rng = np.random.default_rng(3)
    k = rng.uniform(-0.5, 0.5, (40, 2))
    N, os, W, it = [12, 12], 2.1, 5.5, 4
    nos = R.sdc_grid_size(N, os)
    p = (k + 0.5) * np.array(nos)
    r0 = R.JINC_NORM * W
    # dense G[c, j] over every grid cell, nearest periodic image
    cy, cx = np.meshgrid(np.arange(nos[1]), np.arange(nos[0]), indexing="ij")
    cells = np.stack([cx.ravel(), cy.ravel()], axis=1).astype(np.float64)
    u = cells[:, None, :] - p[None, :, :]
    u -= np.array(nos) * np.round(u / np.array(nos))
    inside = np.all(np.abs(u) <= r0, axis=2)
    G = np.where(inside, R.jinc(np.linalg.norm(u, axis=2) / r0), 0.0)
    d = np.ones(len(k))
    for _ in range(it):
        t = G.T @ (G @ d)
        d = np.where(t == 0, 0.0, d / np.where(t == 0, 1.0, t))
    assert np.allclose(R.estimate_dcw(k, N, os, it, W), d, rtol=1e-12, atol=0)


def test_sdc_uniform_and_radial():
    N = 16
    g = (np.stack(np.meshgrid(np.arange(N), np.arange(N), indexing="ij"), -1).reshape(-1, 2) - N / 2) / N
    d = R.estimate_dcw(g, [N, N], num_iterations=20)
    assert d.std() / d.mean() < 1e-10
    tr = R.trajectory_fixed_angle_2d(32, 24)[0]
    d = R.estimate_dcw(tr, [16, 16], num_iterations=10)
    k = np.linalg.norm(tr, axis=1)
    m = (k > 0.1) & (k < 0.4)
    assert np.corrcoef(d[m], k[m])[0, 1] > 0.98
