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
