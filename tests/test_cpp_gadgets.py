"""The gadget restatements (include/gt_gadgets.hpp): radial profiles streamed
through gpuRadialSensePrepGadget into gpuCgSenseGadget, and the gridding
reconstruction gadget, checked against the oracle stage by stage.

CPU: the header and its test program compile.
GPU: tests/cpp/gadget_chain (built by __graft_entry__.build()) streams one
acquisition profile by profile, in golden-ratio mode (2, profiles per frame
given) and in fixed-angle mode (0, profiles per frame and frames per
rotation found from the profile numbering).  Every job must carry the
profiles of its frame, the trajectory and DCW the reference's formulas give
for that frame, and coil maps and regularisation image from an accumulation
buffer fed the same frames (recomputed when a buffer cycle completes, as the
gadget does; relative L2 <= 1e-4 through the two stages); the CG-SENSE image
of every job must match the oracle's regularised, preconditioned CG on that
job (<= 1e-4), and the gridding gadget's coil-combined image the oracle's
(<= 1e-5).
"""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
BIN = os.path.join(CPP, "gadget_chain")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_cpp_gadgets_compile():
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        "-I" + os.path.join(ROOT, "include"), os.path.join(CPP, "gadget_chain.cpp")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


def _round_up(v, m):
    return ((v + m - 1) // m) * m


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [2, 0])
def test_gadget_chain(gpu, tmp_path, mode):
    from oracle import b1map as ob1
    from oracle import buffer as obuf
    from oracle import cgsense as ocg
    from oracle import radial as orad
    print(f"[gpu-case] gadget chain, radial mode {mode}: 64^2, 4 coils, 128 samples x 16 profiles", flush=True)
    assert os.path.exists(BIN), "tests/cpp/gadget_chain is not built (run __graft_entry__.build())"
    N, C, S, P = 64, 4, 128, 16
    F, fpr = (3, 1) if mode == 2 else (5, 4)
    M = S * P
    rng = np.random.default_rng(23 + mode)
    samples = ((rng.standard_normal((F * P, C, S)) + 1j * rng.standard_normal((F * P, C, S))) / np.sqrt(2)).astype(
        np.complex64)
    (tmp_path / "meta.txt").write_text(f"{N} {C} {S} {P} {F} {mode} {fpr}\n")
    samples.tofile(tmp_path / "samples.bin")
    r = subprocess.run([BIN, str(tmp_path)], capture_output=True, text=True, timeout=300)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0, r.stderr
    njobs = int((tmp_path / "njobs.txt").read_text())
    assert njobs == F

    # the prep gadget's sizes at warp size 64: recon 64, buffer oversampling 1.25 -> 128 (factor 2)
    Nos_buf = _round_up(int(np.ceil(N * 1.25)), 64)
    alpha = Nos_buf / N
    # the CG gadget's plan: oversampling 1.5 -> 128
    Nos_cg = _round_up(int(np.ceil(N * 1.5)), 64)
    os_inv = 1.0 / (S / N)
    if mode == 2:
        dcw_ref = orad.radial_dcw_2d(S, P, alpha, os_inv, golden=True)
        sub = Nos_buf // P  # buffer frames per rotation (golden ratio)
    else:
        dcw_ref = orad.radial_dcw_2d(S, P, alpha, os_inv, golden=False)
        sub = fpr
    buf = obuf.Buffer()
    buf.setup([N, N], [Nos_buf] * 2, 5.5, C, 1, sub)
    buf.set_dcw(dcw_ref)
    update, csm_ref, reg_ref = True, None, None
    kappa = 0.3

    def load(j, name, dt, shape):
        return np.fromfile(tmp_path / f"job{j}_{name}.bin", dt).reshape(shape)

    for f in range(F):
        if mode == 2:
            traj_ref = orad.trajectory_golden_ratio_2d(S, P, 1, profile_offset=f * P)[0]
        else:
            traj_ref = orad.trajectory_fixed_angle_2d(S, P, 1, angular_offset=np.pi / P * (f % fpr) / fpr)[0]
        frame = np.ascontiguousarray(samples[f * P:(f + 1) * P].transpose(1, 0, 2)).reshape(C, M)
        update |= buf.add_frame_data(frame, traj_ref)

        tra = load(f, "tra", np.float32, (M, 2))
        dcw = load(f, "dcw", np.float32, (M,))
        dat = load(f, "dat", np.complex64, (C, M))
        csm = load(f, "csm", np.complex64, (C, N, N))
        reg = load(f, "reg", np.complex64, (N, N))
        img = load(f, "img", np.complex64, (N, N))
        hdr = [[int(v) for v in line.split()] for line in (tmp_path / f"job{f}_hdr.txt").read_text().splitlines()]
        assert np.array_equal(dat, frame)
        assert np.abs(tra - traj_ref).max() <= 1e-5
        assert rel_l2(dcw, dcw_ref) <= 1e-5
        if update:
            acc = buf.get_accumulated_coil_images()
            csm_ref = ob1.estimate_b1_map(acc, 2)
            reg_ref = np.sum(np.conj(csm_ref) * acc, axis=0).astype(np.complex64)
            update = False
        # two stages from the oracle (its buffer, then its coil-map estimate): 1e-4
        assert rel_l2(csm, csm_ref) <= 1e-4
        assert rel_l2(reg, reg_ref) <= 1e-4
        # job header: image index f, matrix (N, N, 1), C channels; output image: one frame, one channel
        assert hdr[0] == [f, N, N, 1, C, 0, 0]
        assert hdr[1] == [f, N, N, 1, 1, 0, 0]

        # the CG gadget on the job's own arrays
        # an explicit oversampled size in the cuNFFT convention, as the GPU plan
        plan = po.Plan([N, N], 5.5, Nos=[Nos_cg] * 2, precision="float", cu=True)
        plan.preprocess(tra[None])
        w = np.sqrt(dcw).astype(np.float32)
        op = ocg.SenseOp(plan, csm, w[None])
        rimg, _ = ocg.image_operator_compute(reg)
        reg_img = (np.float32(kappa) * rimg).astype(np.float32)
        pw = ocg.precond_weights(csm, reg_img, 1.0)
        data = (dat * w[None]).astype(np.complex64)[:, None]
        xo, _ = ocg.cg_solve(op, data, 5, tc_tolerance=1e-6, precond=pw, regs=[(kappa, reg_img)])
        err = rel_l2(img, xo[0])
        print(f"  job {f}: CG-SENSE image against the oracle {err:.2e}", flush=True)
        assert err <= 1e-4

    # the gridding gadget on job 0 (warp size 32 there: 1.5 x 64 -> 96)
    tra0 = load(0, "tra", np.float32, (M, 2))
    dcw0 = load(0, "dcw", np.float32, (M,))
    dat0 = load(0, "dat", np.complex64, (C, M))
    Nos_g = _round_up(int(np.ceil(N * 1.5)), 32)
    scaled = (dcw0 * np.float32(Nos_g * Nos_g / np.sum(np.abs(dcw0.astype(np.float64))))).astype(np.float32)
    gplan = po.Plan([N, N], 5.5, Nos=[Nos_g] * 2, precision="float", cu=True)
    gplan.preprocess(tra0[None])
    images = gplan.compute(dat0[:, None], (C, 1, N, N), po.BACKWARDS_NC2C, dcw=scaled[None])[:, 0]
    gcsm = ob1.estimate_b1_map(images, 2)
    comb = np.sum(np.conj(gcsm) * images, axis=0)
    grid = np.fromfile(tmp_path / "grid.bin", np.complex64).reshape(N, N)
    assert rel_l2(grid, comb) <= 1e-5
