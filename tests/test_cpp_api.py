"""The C++ front end (include/gt_nufft.hpp): the cuNFFT_plan<float,2> shape a
Gadgetron C++ caller binds to.

CPU: the header and the test program compile (g++ -fsyntax-only).
GPU: tests/cpp/hpp_roundtrip (built by __graft_entry__.build()) runs
BACKWARDS_NC2C, FORWARDS_C2NC and mult_MH_M through the C++ API on the plain
2D workload; its outputs match the oracle at relative L2 <= 1e-5, and its
error paths throw std::runtime_error.
"""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import c1_inputs, rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
BIN = os.path.join(CPP, "hpp_roundtrip")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_cpp_front_end_compiles():
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        "-I" + os.path.join(ROOT, "include"), os.path.join(CPP, "hpp_roundtrip.cpp")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_cpp_front_end_round_trip(gpu, tmp_path):
    print("[gpu-case] C++ front end: plain 2D 128^2 through cuNFFT_plan<float,2>", flush=True)
    assert os.path.exists(BIN), "tests/cpp/hpp_roundtrip is not built (run __graft_entry__.build())"
    traj, samples, dcw = c1_inputs()  # (1, M, 2), (1, 1, M), (1, M)
    M = traj.shape[1]
    C = 2
    rng = np.random.default_rng(11)
    samples = np.concatenate([samples, (rng.standard_normal((1, 1, M)) + 1j * rng.standard_normal((1, 1, M)))
                              .astype(np.complex64) / np.sqrt(2)], axis=0)
    N = 128
    (tmp_path / "meta.txt").write_text(f"{N} {M} 1 {C}\n")
    np.ascontiguousarray(traj, np.float32).tofile(tmp_path / "traj.bin")
    np.ascontiguousarray(samples, np.complex64).tofile(tmp_path / "samples.bin")
    np.ascontiguousarray(dcw, np.float32).tofile(tmp_path / "dcw.bin")
    r = subprocess.run([BIN, str(tmp_path)], capture_output=True, text=True, timeout=120)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0, r.stderr
    img = np.fromfile(tmp_path / "adjoint.bin", np.complex64).reshape(C, 1, N, N)
    fwd = np.fromfile(tmp_path / "forward.bin", np.complex64).reshape(C, 1, M)
    mhm = np.fromfile(tmp_path / "mhm.bin", np.complex64).reshape(C, 1, N, N)

    p = po.Plan([N, N], 5.5, alpha=1.5, precision="float")
    p.preprocess(traj)
    ref_img = p.compute(samples, (C, 1, N, N), po.BACKWARDS_NC2C, dcw=dcw)
    ref_fwd = p.compute(img, (C, 1, M), po.FORWARDS_C2NC, dcw=dcw)
    ref_mhm = p.mult_MH_M(img, dcw=dcw)
    assert rel_l2(img, ref_img) <= 1e-5
    assert rel_l2(fwd, ref_fwd) <= 1e-5
    assert rel_l2(mhm, ref_mhm) <= 1e-5
    # NFFTOperator<float, 2>: mult_MH twice with accumulate, mult_M, mult_MH_M twice
    op_mh2 = np.fromfile(tmp_path / "op_mh2.bin", np.complex64).reshape(C, 1, N, N)
    op_m = np.fromfile(tmp_path / "op_m.bin", np.complex64).reshape(C, 1, M)
    op_mhm2 = np.fromfile(tmp_path / "op_mhm2.bin", np.complex64).reshape(C, 1, N, N)
    assert rel_l2(op_mh2, 2 * ref_img) <= 1e-5
    assert rel_l2(op_m, ref_fwd) <= 1e-5
    assert rel_l2(op_mhm2, 2 * ref_mhm) <= 1e-5


PMRI_BIN = os.path.join(CPP, "pmri_roundtrip")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_cpp_pmri_front_end_compiles():
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        "-I" + os.path.join(ROOT, "include"), os.path.join(CPP, "pmri_roundtrip.cpp")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_cpp_pmri_pipeline(gpu, tmp_path):
    """include/gt_pmri.hpp end to end (trajectory, DCW, SDC, cuSenseBuffer,
    estimate_b1_map, SENSE operator, CG), each output against its oracle given
    the previous stage's device output."""
    print("[gpu-case] C++ front end: golden-angle SENSE pipeline, 64^2, 4 coils, 3 frames", flush=True)
    from oracle import b1map as ob1
    from oracle import buffer as obuf
    from oracle import cgsense as ocg
    from oracle import radial as orad
    assert os.path.exists(PMRI_BIN), "tests/cpp/pmri_roundtrip is not built (run __graft_entry__.build())"
    N, C, S, P, F = 64, 4, 128, 16, 3
    M = S * P
    rng = np.random.default_rng(17)
    samples = ((rng.standard_normal((F, C, M)) + 1j * rng.standard_normal((F, C, M))) / np.sqrt(2)).astype(np.complex64)
    (tmp_path / "meta.txt").write_text(f"{N} {C} {S} {P} {F}\n")
    samples.tofile(tmp_path / "samples.bin")
    r = subprocess.run([PMRI_BIN, str(tmp_path)], capture_output=True, text=True, timeout=120)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0, r.stderr

    def load(name, dt, shape):
        return np.fromfile(tmp_path / name, dt).reshape(shape)

    traj = load("traj.bin", np.float32, (F, M, 2))
    dcw = load("dcw.bin", np.float32, (M,))
    sdc = load("sdc.bin", np.float32, (M,))
    acc = load("acc.bin", np.complex64, (C, N, N))
    csm = load("csm.bin", np.complex64, (C, N, N))
    comb = load("comb.bin", np.complex64, (N, N))
    mh = load("mh.bin", np.complex64, (1, N, N))
    x = load("cg.bin", np.complex64, (1, N, N))
    flags = [int(v) for v in (tmp_path / "flags.txt").read_text().split()]

    # positions within the float32 rounding of angles up to P F golden steps
    assert np.abs(traj - orad.trajectory_golden_ratio_2d(S, P, F)).max() <= 1e-5
    assert rel_l2(dcw, orad.radial_dcw_2d(S, P, 1.5, 1.0 / 1.5, golden=True)) <= 1e-5
    assert rel_l2(sdc, orad.estimate_dcw(traj[0], [N, N], 2.1, 10, 5.5)) <= 1e-4
    assert flags == [1] * F
    b = obuf.Buffer()
    b.setup([N, N], [N * 3 // 2] * 2, 5.5, C, 2, 1)
    b.set_dcw(dcw)
    for f in range(F):
        b.add_frame_data(samples[f], traj[f])
    assert rel_l2(acc, b.get_accumulated_coil_images()) <= 1e-5
    assert rel_l2(csm, ob1.estimate_b1_map(acc, 2)) <= 1e-5
    assert rel_l2(comb, np.sum(np.conj(csm) * acc, axis=0)) <= 1e-5
    plan = po.Plan([N, N], 5.5, alpha=1.5, precision="float")
    plan.preprocess(traj[:1])
    op = ocg.SenseOp(plan, csm, np.sqrt(dcw)[None])
    assert rel_l2(mh, op.mult_MH(samples[0][:, None])) <= 1e-5
    xo, _ = ocg.cg_solve(op, samples[0][:, None], 5, tc_tolerance=1e-12)
    assert rel_l2(x, xo) <= 1e-4
    # gpuCgSenseGadget's order: R from the combined image, its image scaled by
    # kappa in place (the reference's get() aliases it), precond 1/sqrt(sum |csm|^2 + kappa R)
    kappa = 0.05
    xr = load("cgreg.bin", np.complex64, (1, N, N))
    rimg, _ = ocg.image_operator_compute(comb)
    reg_img = (np.float32(kappa) * rimg).astype(np.float32)
    w = ocg.precond_weights(csm, reg_img, 1.0)
    xro, _ = ocg.cg_solve(op, samples[0][:, None], 5, tc_tolerance=1e-12, precond=w, regs=[(kappa, reg_img)])
    assert rel_l2(xr, xro) <= 1e-4
