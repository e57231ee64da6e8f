"""Pins of the coil-sensitivity oracle (oracle/b1map.py) on the CPU:
  * coil images that are one image times constant coil weights s give
    s / ||s|| exp(-i arg s_0) at every pixel, border included (the smoothed
    correlation is rank one there), in 2D and 3D -- a transposed or
    conjugated correlation, a dropped normalisation or a wrong phase reference
    fails it;
  * the box smoothing equals scipy.ndimage.uniform_filter with zero padding;
  * the two power iterations equal S(p)^2 1 / ||S(p)^2 1|| formed with dense
    per-pixel matrices;
  * target_coils takes the first coils; out-of-range values take all;
  * a pixel whose coils are all zero is NaN, as RSS normalisation makes it.
"""
import numpy as np
import pytest
from scipy.ndimage import uniform_filter

from oracle import b1map as ob


def smooth_phantom(shape, rng):
    grids = np.meshgrid(*[np.linspace(-1, 1, n) for n in shape], indexing="ij")
    m = 1.0 + 0.5 * np.cos(3 * grids[0]) * np.sin(2 * grids[-1])
    return m * np.exp(1j * 2.0 * grids[0] * grids[-1]) + 0.05 * rng.standard_normal(shape)


@pytest.mark.parametrize("shape", [(24, 20), (12, 10, 9)])
def test_constant_coil_weights_closed_form(shape):
    rng = np.random.default_rng(0)
    C = 6
    s = rng.standard_normal(C) + 1j * rng.standard_normal(C)
    m = smooth_phantom(shape, rng)
    x = s.reshape((C,) + (1,) * len(shape)) * m[None]
    csm = ob.estimate_b1_map(x, len(shape))
    want = s / np.linalg.norm(s) * np.exp(-1j * np.angle(s[0]))
    assert np.allclose(csm, want.reshape((C,) + (1,) * len(shape)), atol=1e-12)


def test_box_smoothing_is_zero_padded_uniform_filter():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((3, 17, 13)) + 1j * rng.standard_normal((3, 17, 13))
    got = ob.box_smooth(a, 2)
    want = np.stack([uniform_filter(c.real, 7, mode="constant") + 1j * uniform_filter(c.imag, 7, mode="constant")
                     for c in a])
    assert np.allclose(got, want, atol=1e-13)


def test_power_iterations_equal_dense_matrix_powers():
    rng = np.random.default_rng(2)
    C, shape = 5, (11, 9)
    x = ob.rss_normalize(rng.standard_normal((C,) + shape) + 1j * rng.standard_normal((C,) + shape))
    rs = ob.box_smooth(ob.correlation(x), 2)
    v = ob.power_iterations(rs)
    for p in [(0, 0), (5, 4), (10, 8), (3, 7)]:
        xs = np.pad(x, [(0, 0), (3, 3), (3, 3)])[:, p[0]:p[0] + 7, p[1]:p[1] + 7].reshape(C, -1)
        S = xs @ xs.conj().T / 49.0          # S[j, k] = box(x_j conj x_k)
        w = S @ (S @ np.ones(C))
        assert np.allclose(v[(slice(None),) + p], w / np.linalg.norm(w), atol=1e-12)


def test_target_coils_and_zero_pixels():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((6, 14, 12)) + 1j * rng.standard_normal((6, 14, 12))
    a = ob.estimate_b1_map(x, 2, target_coils=4)
    assert a.shape[0] == 4 and np.allclose(a, ob.estimate_b1_map(x[:4], 2))
    for t in (0, -1, 7):
        assert np.allclose(ob.estimate_b1_map(x, 2, target_coils=t), ob.estimate_b1_map(x, 2))
    x[:, 6, 5] = 0
    c = ob.estimate_b1_map(x, 2)
    nan = np.isnan(c[0])
    assert nan[3:10, 2:9].all() and nan.sum() == 49
