"""Independent dense numpy restatement of the hoNFFT path (test infrastructure only).

Written from the mathematics rather than from the C++ oracle, so the two can
check each other on tiny problems:

* I0 comes from scipy.special.i0 (not the rational approximant of
  ConvolutionKernel.hpp:9-68).
* The interpolation matrix is built densely: every (sample, cell) pair, with
  the periodic images g - N_os, g, g + N_os of each cell tested against the
  kernel radius, instead of the per-sample window walk of
  ConvolutionMatrix.cpp:18-77.
* Centred FFTs are numpy.fft with fftshift/ifftshift and norm="ortho"
  (hoNDFFT.cpp:576-586 semantics).

Arrays are numpy C-order with reversed Gadgetron dims: an image [Nx, Ny] is
shape (Ny, Nx); a trajectory is (M, D) with column 0 = x.
"""
import numpy as np
import scipy.special


def beta_of(W, alpha):
    # ConvolutionKernel.hpp:287-303 (Beatty et al. 2005)
    return np.pi * np.sqrt(W * W / (alpha * alpha) * (alpha - 0.5) ** 2 - 0.8)


def kb1d(u, W, nos, beta):
    """phi(u) = (N_os/W) I0(beta sqrt(1 - (2u/W)^2)) for |u| <= W/2, else 0."""
    u = np.abs(np.asarray(u, dtype=np.float64))
    inside = u <= W / 2
    arg = beta * np.sqrt(np.clip(1.0 - (2.0 * u / W) ** 2, 0.0, None))
    return np.where(inside, nos / W * scipy.special.i0(arg), 0.0)


def interp_matrix(traj, N, Nos, W):
    """Dense (M, G) matrix A with samples = A @ grid (grid flattened x fastest)."""
    traj = np.asarray(traj, dtype=np.float64)
    M, D = traj.shape
    alphas = [Nos[d] / N[d] for d in range(D)]
    betas = [beta_of(W, a) for a in alphas]
    p = (traj + 0.5) * np.asarray(Nos, dtype=np.float64)  # hoGriddingConvolution.cpp:50-56
    per_dim = []
    for d in range(D):
        g = np.arange(Nos[d], dtype=np.float64)
        acc = np.zeros((M, Nos[d]))
        for shift in (-Nos[d], 0, Nos[d]):
            acc += kb1d(g[None, :] + shift - p[:, d:d + 1], W, Nos[d], betas[d])
        per_dim.append(acc)
    # outer product over dims; flattened index = x + Nos_x * (y + Nos_y * z)
    A = per_dim[0]
    for d in range(1, D):
        A = (per_dim[d][:, :, None] * A[:, None, :]).reshape(M, -1)
    return A


def centred_fft(x, D, forward):
    axes = tuple(range(x.ndim - D, x.ndim))
    f = np.fft.fftn if forward else np.fft.ifftn
    return np.fft.fftshift(f(np.fft.ifftshift(x, axes=axes), axes=axes, norm="ortho"), axes=axes)


def deapod_filter(N, Nos, W):
    """Reciprocal deapodization filter of the os-factor / cuNFFT convention
    (hoNFFT.cpp:175-197: centred FFT with do_scale=false)."""
    D = len(N)
    G = int(np.prod(Nos))
    betas = [beta_of(W, Nos[d] / N[d]) for d in range(D)]
    h = np.ones([1] * D)
    for d in range(D):
        v = kb1d(np.arange(Nos[d]) - Nos[d] / 2.0, W, Nos[d], betas[d])
        shape = [1] * D
        shape[D - 1 - d] = Nos[d]
        h = h * v.reshape(shape)
    H = centred_fft(h.astype(np.complex128), D, forward=False) * np.sqrt(G)
    return 1.0 / H


def pad(img, N, Nos):
    D = len(N)
    out = np.zeros(img.shape[:-D] + tuple(reversed(Nos)), dtype=img.dtype)
    sl = tuple(slice(Nos[d] // 2 - N[d] // 2, Nos[d] // 2 - N[d] // 2 + N[d]) for d in reversed(range(D)))
    out[(Ellipsis,) + sl] = img
    return out


def crop(grid, N, Nos):
    D = len(N)
    sl = tuple(slice((Nos[d] - N[d]) >> 1, ((Nos[d] - N[d]) >> 1) + N[d]) for d in reversed(range(D)))
    return grid[(Ellipsis,) + sl]


def forward(img, traj, N, Nos, W, dcw=None):
    """FORWARDS_C2NC for a batch of images (nb, *reversed(N)) -> (nb, M)."""
    D = len(N)
    A = interp_matrix(traj, N, Nos, W)
    g = pad(img.astype(np.complex128), N, Nos) * deapod_filter(N, Nos, W)
    g = centred_fft(g, D, forward=True)
    nb = g.size // int(np.prod(Nos))
    s = g.reshape(nb, -1) @ A.T
    if dcw is not None:
        s = s * dcw
    return s


def adjoint(samples, traj, N, Nos, W, dcw=None):
    """BACKWARDS_NC2C: (nb, M) -> (nb, *reversed(N))."""
    D = len(N)
    A = interp_matrix(traj, N, Nos, W)
    s = samples.astype(np.complex128)
    if dcw is not None:
        s = s * dcw
    g = (s @ A).reshape((s.shape[0],) + tuple(reversed(Nos)))
    g = centred_fft(g, D, forward=False) * deapod_filter(N, Nos, W)
    return crop(g, N, Nos)
