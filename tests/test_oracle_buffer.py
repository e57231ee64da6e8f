"""Pins of the frame-accumulation buffer oracle (oracle/buffer.py) on the CPU,
against paths of the hoNFFT restatement that do not go through it:
  * one frame per slot, one slot: the readout equals the plan's own
    BACKWARDS_NC2C compute (convolve, FFT, deapodize and crop in one call),
    with and without density weights, including an odd crop offset;
  * the running sum over K cycles of S frames equals, by linearity, the sum of
    compute() over exactly the frames of the last K completed cycles, and it
    stays that way as the oldest cycle leaves;
  * cycle_completed fires on every S-th frame, readout is ones until then,
    and the normalization factor is 1 / (K S);
  * the combined image is sum_c conj(csm_c) image_c;
  * the cuNFFT N_os convention the buffer uses equals the alpha constructor
    where alpha N is whole, differs from hoNFFT's N_os constructor by exactly
    sqrt(prod N_os), and its readout times sqrt(prod N_os) approximates the
    direct adjoint NDFT.
"""
import numpy as np

from gadgetron_amd import trajectories as tj
from oracle import buffer as ob
from oracle import pyoracle as po

N, NOS, W, C = [24, 20], [38, 32], 5.5, 3


def frames(n, M_prof=6, S=48, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for f in range(n):
        traj = tj.golden_radial_2d(S, M_prof, offset=f * M_prof)[0]
        out.append((traj, tj.complex_normal(rng, (C, traj.shape[0]))))
    return out


def direct(traj, smp, dcw=None):
    p = po.Plan(N, W, Nos=NOS, cu=True)
    p.preprocess(traj[None])
    return p.compute(smp, [C] + list(reversed(N)), 3, dcw)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_single_slot_equals_compute():
    (traj, smp), = frames(1)
    for dcw in (None, np.linspace(0.2, 1.0, traj.shape[0]).astype(np.float32)):
        b = ob.Buffer()
        b.setup(N, NOS, W, C, 1, 1)
        b.set_dcw(dcw)
        assert b.add_frame_data(smp, traj)
        assert rel(b.get_accumulated_coil_images(), direct(traj, smp, dcw)) < 2e-6


def test_running_sum_over_last_cycles():
    K, S = 2, 3
    fr = frames(11, seed=1)
    b = ob.Buffer()
    b.setup(N, NOS, W, C, K, S)
    assert b.get_normalization_factor() == np.float32(1.0) / np.float32(K * S)
    singles = [direct(t, s) for t, s in fr]
    for i, (t, s) in enumerate(fr):
        done = b.add_frame_data(s, t)
        assert done == ((i + 1) % S == 0)
        img = b.get_accumulated_coil_images()
        completed = (i + 1) // S
        if completed == 0:
            assert np.all(img == 1)
            continue
        first = max(0, completed - K) * S
        want = sum(singles[first:completed * S])
        assert rel(img, want) < 1e-5, i


def test_clear_and_combined_image():
    fr = frames(2, seed=2)
    b = ob.Buffer()
    b.setup(N, NOS, W, C, 1, 1)
    b.add_frame_data(fr[0][1], fr[0][0])
    b.clear()
    assert np.all(b.get_accumulated_coil_images() == 1)
    b.add_frame_data(fr[1][1], fr[1][0])
    img = direct(*fr[1])
    rng = np.random.default_rng(3)
    csm = tj.complex_normal(rng, img.shape)
    want = np.einsum("c...,c...->...", np.conj(csm), img)
    assert rel(b.get_combined_coil_image(csm), want) < 2e-6


def test_cunfft_nos_convention():
    (traj, smp), = frames(1, seed=4)
    shape = [C] + list(reversed(N))
    out = {}
    for name, kw in (("cu", dict(Nos=[36, 30], cu=True)), ("ho", dict(Nos=[36, 30])), ("alpha", dict(alpha=1.5))):
        p = po.Plan(N, W, **kw)
        p.preprocess(traj[None])
        out[name] = p.compute(smp, shape, 3)
    assert rel(out["cu"], out["alpha"]) < 1e-6
    assert rel(out["ho"] / np.sqrt(36 * 30), out["cu"]) < 1e-6
    # direct adjoint NDFT (FP64) of coil 0
    t = traj.astype(np.float64)
    ny, nx = np.meshgrid(np.arange(N[1]) - N[1] // 2, np.arange(N[0]) - N[0] // 2, indexing="ij")
    E = np.exp(2j * np.pi * (t[:, 0, None, None] * nx + t[:, 1, None, None] * ny))
    ref = np.einsum("m,myx->yx", smp[0].astype(np.complex128), E)
    assert rel(out["cu"][0] * np.sqrt(36 * 30), ref) < 2e-3
