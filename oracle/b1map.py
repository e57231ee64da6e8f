"""CPU restatement of the reference's coil-sensitivity estimate -- test
infrastructure only (the checker of include/gt_b1map.h; the library never
calls it).

Follows toolboxes/mri/pmri/gpu/b1_map.cu:
  :32-108   estimate_b1_map: first target coils, RSS normalisation, correlation
            matrices, box smoothing, dominant eigenvector, phase reference
  :145-187  rss_normalize (x / sqrt(sum_c |x_c|^2); NaN where the sum is 0)
  :189-235  correlation: R[a, b] = x_b conj(x_a)
  :237-640  smooth_correlation_matrices: 7-wide box sum over the taps inside
            the image, times 1 / 7^D (interior and border kernels together)
  :690-760  extract_csm_kernel: two power iterations from ones,
            v_j <- sum_k R[k, j] v_k, then v / ||v||
  :775-800  set_phase_reference: multiply by conj(exp(i arg v_0))
Images are numpy order: (C, *reversed(N)), coil slowest.  Computed in
float64 (the GPU path is float32; the bar is a relative tolerance).
"""
import numpy as np

KERNEL_WIDTH = 7


def rss_normalize(x):
    with np.errstate(divide="ignore", invalid="ignore"):
        return x * (1.0 / np.sqrt(np.sum(np.abs(x) ** 2, axis=0)))


def correlation(x):
    """R[a, b] = x_b conj(x_a), shape (C, C, *image)."""
    return x[None, :] * np.conj(x[:, None])


def box_smooth(a, D):
    """Zero-padded KERNEL_WIDTH^D box sum over the last D axes, / KERNEL_WIDTH^D."""
    h = KERNEL_WIDTH >> 1
    img = a.shape[-D:]
    pad = [(0, 0)] * (a.ndim - D) + [(h, h)] * D
    ap = np.pad(a, pad)
    out = np.zeros_like(a)
    for off in np.ndindex(*([KERNEL_WIDTH] * D)):
        sl = tuple([Ellipsis] + [slice(o, o + n) for o, n in zip(off, img)])
        out += ap[sl]
    return out / float(KERNEL_WIDTH ** D)


def power_iterations(rs, iterations=2):
    """rs: smoothed R, (C, C, *image).  v_j <- sum_k rs[k, j] v_k, normalised."""
    C = rs.shape[0]
    v = np.ones((C,) + rs.shape[2:], dtype=rs.dtype)
    for _ in range(iterations):
        t = np.einsum("kj...,k...->j...", rs, v)
        with np.errstate(divide="ignore", invalid="ignore"):
            v = t * (1.0 / np.sqrt(np.sum(np.abs(t) ** 2, axis=0)))
    return v


def set_phase_reference(v):
    a = np.angle(v[0])
    return v * np.conj(np.cos(a) + 1j * np.sin(a))


def estimate_b1_map(data, D, target_coils=-1):
    """data: (C, *image) complex; returns (T, *image) complex128."""
    x = np.asarray(data, np.complex128)
    if x.ndim - 1 != D:
        raise ValueError("estimate_b1_map:: dimensionality mismatch.")
    C = x.shape[0]
    T = C if (target_coils <= 0 or target_coils > C) else int(target_coils)
    x = rss_normalize(x[:T])
    rs = box_smooth(correlation(x), D)
    return set_phase_reference(power_iterations(rs))
