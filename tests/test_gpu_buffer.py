"""GPU parity of the frame-accumulation buffer (include/gt_buffer.h, the
cuBuffer / cuSenseBuffer path) against its oracle (oracle/buffer.py) on the
same frames: relative L2 <= 1e-5 on every readout, identical cycle_completed
flags, ones before the first completed cycle.  Every case prints its name
before touching the GPU.
"""
import sys

import numpy as np
import pytest

from gadgetron_amd import DeviceArray, GTError
from gadgetron_amd import trajectories as tj
from gadgetron_amd.buffer import Buffer, SenseBuffer
from oracle import buffer as ob

pytestmark = pytest.mark.gpu
TOL = 1e-5


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def dev(a, dtype):
    return DeviceArray.from_numpy(np.ascontiguousarray(a, dtype=dtype))


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def frames_2d(n, C, S=64, P=8, seed=0):
    rng = np.random.default_rng(seed)
    return [(tj.golden_radial_2d(S, P, offset=f * P)[0], tj.complex_normal(rng, (C, S * P))) for f in range(n)]


def run_pair(gb, cb, fr, label):
    for i, (traj, smp) in enumerate(fr):
        done_g = gb.add_frame_data(dev(smp, np.complex64), dev(traj, np.float32))
        done_c = cb.add_frame_data(smp, traj)
        assert done_g == done_c, (label, i)
        img = gb.get_accumulated_coil_images().numpy()
        want = cb.get_accumulated_coil_images()
        if cb.empty:
            assert np.all(img == 1), (label, i)
        else:
            assert rel(img, want) <= TOL, (label, i, rel(img, want))


@pytest.mark.parametrize("with_dcw", [False, True])
def test_buffer_2d_cycles(gpu, with_dcw):
    say(f"buffer 2D 48x40 (os 72x64), 4 coils, 2 cycles x 3 frames, dcw={with_dcw}")
    N, NOS, C = [48, 40], [72, 64], 4
    fr = frames_2d(10, C)
    gb, cb = Buffer(), ob.Buffer()
    gb.setup(N, NOS, 5.5, C, 2, 3)
    cb.setup(N, NOS, 5.5, C, 2, 3)
    if with_dcw:
        w = np.abs(fr[0][0][:, 0]) + np.abs(fr[0][0][:, 1]) + 0.05
        d_w = dev(w, np.float32)
        gb.set_dcw(d_w)
        cb.set_dcw(w)
    assert gb.get_normalization_factor() == cb.get_normalization_factor()
    run_pair(gb, cb, fr, "2d")


def test_buffer_3d_and_1d(gpu):
    say("buffer 3D 16^3 (os 24^3), 2 coils, 1 cycle x 2 frames")
    rng = np.random.default_rng(4)
    N, NOS, C = [16, 16, 16], [24, 24, 24], 2
    fr = [(rng.uniform(-0.5, 0.5, (900, 3)).astype(np.float32), tj.complex_normal(rng, (C, 900))) for _ in range(5)]
    gb, cb = Buffer(), ob.Buffer()
    gb.setup(N, NOS, 5.5, C, 1, 2)
    cb.setup(N, NOS, 5.5, C, 1, 2)
    run_pair(gb, cb, fr, "3d")
    say("buffer 1D 64 (os 96), 3 coils, 3 cycles x 1 frame")
    N, NOS, C = [64], [96], 3
    fr = [(rng.uniform(-0.5, 0.5, (200, 1)).astype(np.float32), tj.complex_normal(rng, (C, 200))) for _ in range(6)]
    gb, cb = Buffer(), ob.Buffer()
    gb.setup(N, NOS, 4.5, C, 3, 1)
    cb.setup(N, NOS, 4.5, C, 3, 1)
    run_pair(gb, cb, fr, "1d")


def test_buffer_setup_clear_and_combined_image(gpu):
    say("buffer setup again / clear / combined coil image")
    N, NOS, C = [32, 32], [48, 48], 3
    fr = frames_2d(7, C, S=48, P=6, seed=5)
    gb, cb = SenseBuffer(), ob.Buffer()
    gb.setup(N, NOS, 5.5, C, 2, 2)
    cb.setup(N, NOS, 5.5, C, 2, 2)
    run_pair(gb, cb, fr[:3], "before")
    # the same sizes again keep the slots and the position in the cycle
    gb.setup(N, NOS, 5.5, C, 2, 2)
    cb.setup(N, NOS, 5.5, C, 2, 2)
    run_pair(gb, cb, fr[3:5], "same setup")
    # another cycle count starts over
    gb.setup(N, NOS, 5.5, C, 1, 2)
    cb.setup(N, NOS, 5.5, C, 1, 2)
    assert np.all(gb.get_accumulated_coil_images().numpy() == 1)
    run_pair(gb, cb, fr[5:], "new cycles")
    rng = np.random.default_rng(6)
    csm = tj.complex_normal(rng, (C, 32, 32))
    with pytest.raises(RuntimeError):
        gb.get_combined_coil_image()
    gb.set_csm(dev(csm, np.complex64))
    got = gb.get_combined_coil_image().numpy()
    assert rel(got, cb.get_combined_coil_image(csm)) <= TOL
    gb.clear()
    cb.clear()
    assert np.all(gb.get_accumulated_coil_images().numpy() == 1)
    run_pair(gb, cb, fr[:2], "after clear")


def test_buffer_errors(gpu):
    say("buffer argument errors")
    b = Buffer()
    with pytest.raises(RuntimeError):
        b.add_frame_data(dev(np.zeros((1, 4), np.complex64), np.complex64), dev(np.zeros((4, 2)), np.float32))
    with pytest.raises(GTError):
        b.setup([16, 16], [24, 24], 5.5, 0, 1, 1)
    with pytest.raises(GTError):
        b.setup([16, 16, 16, 16], [24, 24, 24, 24], 5.5, 1, 1, 1)
    b.setup([16, 16], [24, 24], 5.5, 2, 1, 1)
    with pytest.raises(ValueError):
        b.add_frame_data(dev(np.zeros((3, 4), np.complex64), np.complex64), dev(np.zeros((4, 2)), np.float32))
    with pytest.raises(GTError):  # out-of-range trajectory, caught by preprocess
        b.add_frame_data(dev(np.zeros((2, 4), np.complex64), np.complex64),
                         dev(np.full((4, 2), 0.75), np.float32))
