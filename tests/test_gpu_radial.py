"""GPU parity of include/gt_radial.h against the oracle (oracle/radial.py):
trajectories to within the float32 rounding of their angles, radial Voronoi weights and SDC weights to
relative L2 1e-5 / 1e-4 (SDC runs 10 divisions in FP32 against FP64), the
kooshball against the host generator.  Every case prints its name first."""
import sys

import numpy as np
import pytest

import gadgetron_amd as ga
from gadgetron_amd import DeviceArray
from gadgetron_amd import radial as rd
from gadgetron_amd import trajectories as tj
from oracle import radial as R
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def angle_tol(max_angle):
    """Position error bound from forming the angle in float32 (as the reference
    does): a few ulps of the largest angle, times the largest radius 1/2."""
    return 2.0 * float(np.spacing(np.float32(max_angle))) + 2e-7


def test_trajectories(gpu):
    say("radial trajectories: golden ratio (both steps), fixed angle, variable angle")
    for mode in (rd.GoldenRatioStep.GR_ORIGINAL, rd.GoldenRatioStep.GR_SMALLEST):
        got = rd.trajectory_golden_ratio_2d(96, 12, frames=3, profile_offset=5, mode=mode).numpy()
        tol = angle_tol((36 + 5) * R.golden_step(int(mode)) + np.pi)
        assert np.abs(got - R.trajectory_golden_ratio_2d(96, 12, 3, 5, int(mode))).max() < tol
    got = rd.trajectory_fixed_angle_2d(64, 10, frames=4, angular_offset=0.3).numpy()
    assert np.abs(got - R.trajectory_fixed_angle_2d(64, 10, 4, 0.3)).max() < angle_tol(2 * np.pi + 0.3)
    ang = np.array([0.1, 1.3, -2.0], np.float32)
    got = rd.trajectory_variable_angle_2d(DeviceArray.from_numpy(ang), 32, 5).numpy()
    assert np.abs(got - R.trajectory_variable_angle_2d(ang, 32, 5)).max() < angle_tol(2.0)


@pytest.mark.parametrize("golden,S,P", [(False, 64, 16), (True, 128, 32), (True, 64, 7)])
def test_radial_dcw(gpu, golden, S, P):
    say(f"radial dcw golden={golden} S={S} P={P}")
    alpha, osinv = 1.5, 1.0 / 1.5
    if golden:
        got = rd.dcw_golden_ratio_2d(S, P, alpha, osinv, profile_offset=3).numpy()
        ref = R.radial_dcw_2d(S, P, alpha, osinv, golden=True, profile_offset=3)
    else:
        got = rd.dcw_fixed_angle_2d(S, P, alpha, osinv).numpy()
        ref = R.radial_dcw_2d(S, P, alpha, osinv, golden=False)
    assert rel_l2(got, ref) <= 1e-5


def test_kooshball_matches_host_generator(gpu):
    say("3D kooshball 40 x 300 against the host generator")
    rng = np.random.default_rng(7)
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    got = rd.kooshball_3d(40, 300, rotation=q).numpy()
    ref = tj.kooshball_3d(40, 300, seed=7)
    assert np.abs(got - ref).max() < 1e-6


@pytest.mark.parametrize("case", ["fixed2d", "golden2d_init", "koosh3d"])
def test_sdc_matches_oracle(gpu, case):
    say(f"sdc {case}")
    init = None
    if case == "fixed2d":
        traj, N = R.trajectory_fixed_angle_2d(32, 24)[0], [16, 16]
    elif case == "golden2d_init":
        traj, N = R.trajectory_golden_ratio_2d(48, 20)[0], [24, 20]
        init = np.linalg.norm(traj, axis=1).astype(np.float32) + 0.05
    else:
        traj, N = tj.kooshball_3d(16, 60, seed=3)[0].astype(np.float64), [8, 8, 8]
    t32 = np.ascontiguousarray(traj, np.float32)
    d_init = None if init is None else DeviceArray.from_numpy(init)
    got = rd.estimate_dcw(DeviceArray.from_numpy(t32), N, 2.1, 10, 5.5, initial_dcw=d_init).numpy()
    ref = R.estimate_dcw(t32.astype(np.float64), N, 2.1, 10, 5.5, initial_dcw=init)
    err = rel_l2(got, ref)
    say(f"  rel err {err:.2e}")
    assert err <= 1e-4


def test_errors(gpu):
    say("radial / sdc argument errors")
    with pytest.raises(ga.GTError):
        rd.dcw_fixed_angle_2d(63, 16, 1.0, 1.0)  # odd samples per profile
    with pytest.raises(ga.GTError):
        rd.dcw_fixed_angle_2d(64, 3, 1.0, 1.0)  # fewer than four profiles
    bad = DeviceArray.from_numpy(np.array([[0.2, 0.7], [0.0, 0.1]], np.float32))
    with pytest.raises(ga.GTError) as e:
        rd.estimate_dcw(bad, [16, 16])
    assert "trajectory" in str(e.value)
