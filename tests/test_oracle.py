"""Pins the oracle (oracle/honfft_oracle.cpp, the CPU restatement of hoNFFT).

No golden vector for the NFFT itself exists in the reference tree (its
hoNFFT_test data is external), so each part is pinned by what the mathematics
fixes, chosen so that a dropped term, a wrong sign, index or transposition
fails at least one check:

* I0 approximant vs scipy.special.i0; beta closed form;
* kernel row of a sample: window, wrap and weights against a direct formula;
* the reference's own fftshift known answers (test/hoNDFFT_test.cpp:60-166)
  and Parseval (test/hoNDFFT_test.cpp:35-56);
* FFT vs brute-force DFT and numpy.fft;
* adjoint identity of the gridding and of the full NFFT;
* exactness of the NFFT for on-grid frequencies (equals the FP64 NDFT);
* the closed forms at k = 0;
* the reference's degenerate input (test/cuNFFT_test.cpp:41-49) in closed form;
* the deapodization ratio sqrt(G) between the two hoNFFT ctors (hoNFFT.cpp:150-197);
* an independent dense numpy restatement (oracle/nufft_numpy.py) in FP64.
"""
import numpy as np
import pytest
import scipy.special

from oracle import nufft_numpy as nn
from oracle import pyoracle as po
from tests.helpers import c1_inputs, rel_l2

W = 5.5
ALPHA = 1.5


def test_bessi0_matches_scipy_i0():
    xs = np.linspace(0.0, 15.0, 1501)
    d = np.array([po.bessi0(x) for x in xs])
    f = np.array([po.bessi0(x, "float") for x in xs])
    ref = scipy.special.i0(xs)
    assert np.max(np.abs(d - ref) / ref) < 1e-13
    assert np.max(np.abs(f - ref) / ref) < 2e-6


def test_beta_closed_form():
    p = po.Plan([128, 128], W, alpha=ALPHA, precision="double")
    expect = np.pi * np.sqrt(W * W / (ALPHA * ALPHA) * (ALPHA - 0.5) ** 2 - 0.8)
    assert p.beta[0] == pytest.approx(expect, rel=1e-14)
    assert p.beta[0] == pytest.approx(11.17119799, rel=1e-8)
    assert p.Nos == [192, 192]


@pytest.mark.parametrize("k", [(0.0, 0.0), (0.1234, -0.377), (-0.5, 0.49), (0.5, -0.5)])
def test_conv_row_window_and_weights(k):
    N, Nos = [16, 12], [24, 18]
    p = po.Plan(N, W, alpha=ALPHA, precision="double")
    p.preprocess(np.array([[k]], dtype=np.float64))
    idx, w = p.conv_row(0)
    pos = [(k[d] + 0.5) * Nos[d] for d in range(2)]
    r = W / 2
    beta = nn.beta_of(W, ALPHA)
    expect = {}
    for cy in range(int(np.ceil(pos[1] - r)), int(np.floor(pos[1] + r)) + 1):
        for cx in range(int(np.ceil(pos[0] - r)), int(np.floor(pos[0] + r)) + 1):
            wt = nn.kb1d(abs(cx - pos[0]), W, Nos[0], beta) * nn.kb1d(abs(cy - pos[1]), W, Nos[1], beta)
            cell = (cx % Nos[0]) + Nos[0] * (cy % Nos[1])
            expect[cell] = expect.get(cell, 0.0) + float(wt)
    got = {}
    for i, v in zip(idx, w):
        got[int(i)] = got.get(int(i), 0.0) + float(v)
    assert set(got) == set(expect)
    for c in expect:
        assert got[c] == pytest.approx(expect[c], rel=1e-12, abs=1e-300)
    # x varies fastest inside each window row
    nx = int(np.floor(pos[0] + r)) - int(np.ceil(pos[0] - r)) + 1
    assert len(idx) % nx == 0


def test_on_node_peak_weight():
    N, Nos = [16, 16], [24, 24]
    p = po.Plan(N, W, alpha=ALPHA, precision="double")
    k = (3.0 / 24.0 - 0.5, 7.0 / 24.0 - 0.5)  # exactly on node (3, 7)
    p.preprocess(np.array([[k]]))
    idx, w = p.conv_row(0)
    beta = nn.beta_of(W, ALPHA)
    peak = (24.0 / W * scipy.special.i0(beta)) ** 2
    j = list(idx).index(3 + 24 * 7)
    assert w[j] == pytest.approx(peak, rel=1e-12)
    assert w.max() == w[j]


def test_periodic_wrap_identity():
    # k = -1/2 and k = +1/2 are the same point on the periodic grid
    p = po.Plan([16, 16], W, alpha=ALPHA, precision="double")
    p.preprocess(np.array([[[-0.5, 0.2], [0.5, 0.2]]]))
    i0, w0 = p.conv_row(0)
    i1, w1 = p.conv_row(1)
    assert sorted(zip(i0.tolist(), w0.tolist())) == pytest.approx(sorted(zip(i1.tolist(), w1.tolist())))


@pytest.mark.parametrize("n", [24, 8, 26])
def test_fftshift_known_answers_1d(n):
    # test/hoNDFFT_test.cpp:60-80: index 2 -> 14 for n = 24
    x = np.zeros(n, np.complex64)
    x[2] = 1
    y = po.fftshift(x, 1)
    assert y[(2 + n // 2) % n] == 1 and np.count_nonzero(y) == 1


def test_fftshift_known_answers_nd():
    # test/hoNDFFT_test.cpp:82-126
    x = np.zeros((8, 24), np.complex64)  # dims (24, 8)
    x[3, 2] = 1
    y = po.fftshift(x, 2)
    assert y[7, 14] == 1 and y[3, 2] == 0
    x = np.zeros((3, 26, 8, 24), np.complex64)  # dims (24, 8, 26, 3)
    x[0, 4, 3, 2] = 1
    y = po.fftshift(x, 3)
    assert y[0, 17, 7, 14] == 1 and y[0, 4, 3, 2] == 0
    # test/hoNDFFT_test.cpp:139-166 round trip
    rng = np.random.default_rng(1)
    r = (rng.random((3, 26, 8, 24)) + 1j * rng.random((3, 26, 8, 24))).astype(np.complex64)
    back = po.fftshift(po.fftshift(r, 3), 3, inverse=True)
    assert np.array_equal(back, r)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 8, 12, 15, 49, 144, 30, 97])
def test_fft_vs_bruteforce_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    k = np.arange(n)
    F = np.exp(-2j * np.pi * np.outer(k, k) / n)
    for fwd in (True, False):
        y = po.plain_fft(x, 1, fwd, precision="double")
        ref = (F if fwd else F.conj()) @ x
        assert rel_l2(y, ref) < 1e-12


@pytest.mark.parametrize("shape", [(24,), (12, 18), (6, 10, 8)])
def test_centred_fft_vs_numpy_and_parseval(shape):
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2,) + shape) + 1j * rng.standard_normal((2,) + shape)
    D = len(shape)
    axes = tuple(range(1, D + 1))
    for fwd in (True, False):
        y = po.centred_fft(x, D, fwd, precision="double")
        f = np.fft.fftn if fwd else np.fft.ifftn
        ref = np.fft.fftshift(f(np.fft.ifftshift(x, axes=axes), axes=axes, norm="ortho"), axes=axes)
        assert rel_l2(y, ref) < 1e-12
        assert np.linalg.norm(y) == pytest.approx(np.linalg.norm(x), rel=1e-12)
        yf = po.centred_fft(x, D, fwd, precision="float")
        assert rel_l2(yf, ref) < 2e-6


def _plan(N, prec="double", Nos=None):
    if Nos is not None:
        return po.Plan(N, W, Nos=Nos, precision=prec)
    return po.Plan(N, W, alpha=ALPHA, precision=prec)


@pytest.mark.parametrize("prec,tol", [("double", 1e-12), ("float", 2e-6)])
def test_gridding_adjoint_identity(prec, tol):
    rng = np.random.default_rng(5)
    N = [16, 20]
    p = _plan(N, prec)
    G = int(np.prod(p.Nos))
    M, nb = 300, 3
    p.preprocess(rng.uniform(-0.5, 0.5, (M, 2)))
    g = rng.standard_normal(nb * G) + 1j * rng.standard_normal(nb * G)
    s = rng.standard_normal(nb * M) + 1j * rng.standard_normal(nb * M)
    As = p.convolve(g, 0, nb)
    AHs = p.convolve(s, 1, nb)
    lhs, rhs = np.vdot(s, As), np.vdot(AHs, g)
    assert abs(lhs - rhs) / abs(lhs) < tol


def test_nfft_adjoint_identity_and_mult_MH_M():
    rng = np.random.default_rng(6)
    N = [16, 12]
    p = _plan(N)
    M, F, C = 200, 2, 3
    p.preprocess(rng.uniform(-0.5, 0.5, (F, M, 2)))
    x = rng.standard_normal((C, F, 12, 16)) + 1j * rng.standard_normal((C, F, 12, 16))
    y = rng.standard_normal((C, F, M)) + 1j * rng.standard_normal((C, F, M))
    dcw = rng.uniform(0.1, 1.0, (F, M))
    Ax = p.compute(x, (C, F, M), po.FORWARDS_C2NC, dcw=dcw)
    AHy = p.compute(y, (C, F, 12, 16), po.BACKWARDS_NC2C, dcw=dcw)
    assert abs(np.vdot(y, Ax) - np.vdot(AHy, x)) / abs(np.vdot(y, Ax)) < 1e-12
    # mult_MH_M = BACKWARDS_NC2C(FORWARDS_C2NC(x)) (NFFT.hpp:89-101)
    mhm = p.mult_MH_M(x, dcw=dcw)
    ref = p.compute(p.compute(x, (C, F, M), po.FORWARDS_C2NC, dcw=dcw), x.shape, po.BACKWARDS_NC2C, dcw=dcw)
    assert rel_l2(mhm, ref) < 1e-14
    # FORWARDS_NC2C and BACKWARDS_C2NC are adjoint to each other as well
    Fx = p.compute(y, (C, F, 12, 16), po.FORWARDS_NC2C, dcw=dcw)
    By = p.compute(x, (C, F, M), po.BACKWARDS_C2NC, dcw=dcw)
    assert abs(np.vdot(x, Fx) - np.vdot(By, y)) / abs(np.vdot(x, Fx)) < 1e-12


@pytest.mark.parametrize("N", [[16], [16, 12], [8, 12, 4], [4, 8, 4, 4]])
def test_nfft_exact_on_grid_frequencies(N):
    """For on-grid positions k = m / N_os the gridding is exact, so FORWARDS_C2NC
    equals NDFT / sqrt(G) to rounding; BACKWARDS_NC2C equals the adjoint NDFT.
    (Even N_os only: for odd N_os the FFT centre N_os//2 and the trajectory
    centre N_os/2 differ by half a cell, a reference quirk.)"""
    rng = np.random.default_rng(7)
    D = len(N)
    p = _plan(N)
    Nos = p.Nos
    G = float(np.prod(Nos))
    M = 40
    m = np.stack([rng.integers(-Nos[d] // 2, Nos[d] // 2, M) for d in range(D)], axis=1)
    k = m / np.array(Nos, dtype=np.float64)
    p.preprocess(k[None])
    x = rng.standard_normal((2,) + tuple(reversed(N))) + 1j * rng.standard_normal((2,) + tuple(reversed(N)))
    y = p.compute(x, (2, M), po.FORWARDS_C2NC)
    ref = po.ndft(k, N, x=x) / np.sqrt(G)
    assert rel_l2(y, ref) < 1e-11
    s = rng.standard_normal((2, M)) + 1j * rng.standard_normal((2, M))
    xa = p.compute(s, (2,) + tuple(reversed(N)), po.BACKWARDS_NC2C)
    refa = po.ndft(k, N, y=s, adjoint=True) / np.sqrt(G)
    assert rel_l2(xa, refa) < 1e-11


def test_nfft_closed_forms_at_k0():
    N = [16, 12]
    p = _plan(N)
    G = float(np.prod(p.Nos))
    p.preprocess(np.zeros((1, 1, 2)))
    # adjoint of a unit sample at k = 0 is the constant 1/sqrt(G)
    img = p.compute(np.ones((1, 1)), (1, 12, 16), po.BACKWARDS_NC2C)
    assert np.max(np.abs(img - 1 / np.sqrt(G))) < 1e-13
    # forward at k = 0 is the image sum / sqrt(G)
    rng = np.random.default_rng(8)
    x = rng.standard_normal((1, 12, 16)) + 1j * rng.standard_normal((1, 12, 16))
    s = p.compute(x, (1, 1), po.FORWARDS_C2NC)
    assert abs(s[0, 0] - x.sum() / np.sqrt(G)) < 1e-12 * abs(x).sum()


def test_degenerate_trajectory_of_reference_test():
    """test/cuNFFT_test.cpp:41-49 feeds every sample at (0.1, 0.1) with value 1 and
    dcw 1; the NC2C grid is then M times the kernel footprint at p."""
    N, Nos = [32, 32], [48, 48]
    p = po.Plan(N, 3.0, Nos=Nos, precision="double")
    M, C = 500, 2
    p.preprocess(np.full((1, M, 2), 0.1))
    g = p.convolve(np.ones(C * M), 1, C).reshape(C, 48, 48)
    pos = (0.1 + 0.5) * 48
    beta = nn.beta_of(3.0, 1.5)
    wx = np.array([nn.kb1d(abs(c - pos), 3.0, 48, beta) for c in range(48)])
    expect = M * np.outer(wx, wx)
    for c in range(C):
        assert np.allclose(g[c].real, expect, rtol=1e-12, atol=0)
        assert np.all(g[c].imag == 0)


def test_deapodization_ctor_ratio_is_sqrt_G():
    """hoNFFT_plan(N, N_os, W) builds its filters with do_scale=true, hoNFFT_plan(N, a, W)
    with do_scale=false (hoNFFT.cpp:161-164 vs :186-189): outputs differ by sqrt(G)."""
    rng = np.random.default_rng(9)
    N = [16, 12]
    pa = _plan(N)
    pn = _plan(N, Nos=[24, 18])
    G = 24 * 18
    k = rng.uniform(-0.5, 0.5, (1, 50, 2))
    pa.preprocess(k)
    pn.preprocess(k)
    x = rng.standard_normal((1, 12, 16)) + 1j * rng.standard_normal((1, 12, 16))
    ya = pa.compute(x, (1, 50), po.FORWARDS_C2NC)
    yn = pn.compute(x, (1, 50), po.FORWARDS_C2NC)
    assert rel_l2(yn, ya * np.sqrt(G)) < 1e-12


@pytest.mark.parametrize("D", [1, 2, 3])
def test_oracle_vs_independent_numpy_restatement(D):
    rng = np.random.default_rng(10 + D)
    N = [12, 10, 8][:D]
    p = _plan(N)
    Nos = p.Nos
    M = 60
    k = rng.uniform(-0.5, 0.5, (M, D))
    p.preprocess(k[None])
    x = rng.standard_normal((2,) + tuple(reversed(N))) + 1j * rng.standard_normal((2,) + tuple(reversed(N)))
    dcw = rng.uniform(0.2, 1, M)
    y = p.compute(x, (2, M), po.FORWARDS_C2NC, dcw=dcw)
    assert rel_l2(y, nn.forward(x, k, N, Nos, W, dcw=dcw)) < 1e-12
    s = rng.standard_normal((2, M)) + 1j * rng.standard_normal((2, M))
    xa = p.compute(s, x.shape, po.BACKWARDS_NC2C, dcw=dcw)
    assert rel_l2(xa, nn.adjoint(s, k, N, Nos, W, dcw=dcw)) < 1e-12


def test_float_oracle_tracks_double_on_c1():
    """FP32 noise floor of the float algorithm itself: positions p = (k + 1/2) N_os
    are rounded to float (|dp| <= 7.6e-6 cells at N_os = 192), which moves kernel
    tail weights by up to ~7e-5 relative; the float oracle ends ~7e-6 from FP64."""
    traj, samples, dcw = c1_inputs()
    pf = _plan([128, 128], "float")
    pd = _plan([128, 128], "double")
    pf.preprocess(traj)
    pd.preprocess(traj.astype(np.float64))
    img_f = pf.compute(samples, (1, 1, 128, 128), po.BACKWARDS_NC2C, dcw=dcw)
    img_d = pd.compute(samples, (1, 1, 128, 128), po.BACKWARDS_NC2C, dcw=dcw)
    assert rel_l2(img_f, img_d) < 1.5e-5
    s_f = pf.compute(img_f, (1, 1, traj.shape[1]), po.FORWARDS_C2NC, dcw=dcw)
    s_d = pd.compute(img_d, (1, 1, traj.shape[1]), po.FORWARDS_C2NC, dcw=dcw)
    assert rel_l2(s_f, s_d) < 1.5e-5


def test_nfft_approximation_error_vs_ndft_small():
    """States the KB(alpha=1.5, W=5.5) approximation error against the FP64 NDFT."""
    rng = np.random.default_rng(11)
    N = [32, 32]
    p = _plan(N)
    G = float(np.prod(p.Nos))
    k = rng.uniform(-0.5, 0.5, (400, 2))
    p.preprocess(k[None])
    x = rng.standard_normal((1, 32, 32)) + 1j * rng.standard_normal((1, 32, 32))
    y = p.compute(x, (1, 400), po.FORWARDS_C2NC)
    err = rel_l2(y, po.ndft(k, N, x=x) / np.sqrt(G))
    assert 1e-6 < err < 1e-3


def test_4d_adjoint_identity_and_ndft_error():
    """D = 4 (cuNFFT_impl<float, 4>, cuNFFT.cu:270-283): NC2C is the adjoint of
    C2NC, and off-grid samples approximate the FP64 NDFT as in 2D."""
    rng = np.random.default_rng(12)
    N = [8, 4, 4, 8]  # even N_os in every dimension
    p = _plan(N)
    G = float(np.prod(p.Nos))
    M = 300
    k = rng.uniform(-0.5, 0.5, (M, 4))
    p.preprocess(k[None])
    shape = (2,) + tuple(reversed(N))
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    y = rng.standard_normal((2, M)) + 1j * rng.standard_normal((2, M))
    Ax = p.compute(x, (2, M), po.FORWARDS_C2NC)
    AHy = p.compute(y, shape, po.BACKWARDS_NC2C)
    assert abs(np.vdot(y, Ax) - np.vdot(AHy, x)) / abs(np.vdot(y, Ax)) < 1e-12
    err = rel_l2(Ax, po.ndft(k, N, x=x) / np.sqrt(G))
    assert 1e-6 < err < 2e-3
