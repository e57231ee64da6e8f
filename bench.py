#!/usr/bin/env python3
"""NUFFT adjoint+forward throughput on one MI355X.

Workload (default): the 32-coil 2D radial configuration -- 256^2 image,
alpha 1.5 (384^2 grid), Kaiser-Bessel W 5.5, 400 golden-angle spokes x 512
samples, 32 coils batched in one launch, complex64 samples CN(0,1) from seed 1,
ramp DCF.  One step = BACKWARDS_NC2C (samples -> 32 coil images) followed by
FORWARDS_C2NC (images -> samples), DCF applied in both, inputs resident in HBM.

value = M*C / step time, in Msamples*coil/s (one direction's work units per
adjoint+forward pair).  The timed steps record no events except around the two
gridding stages of every 10th step (an event waits for the work before it, so
recording every stage of every step would slow the step by about a sixth).
The roofline object covers the dominant gridding stage (NC2C or C2NC, its mean
HIP-event time on the compute stream over those sampled steps) against its
algorithmic bytes 8*C*(M+G) + 4*D*M + 4*M per launch (samples, grid,
trajectory, DCF each touched once).  The CPU baseline is the oracle (the
hoNFFT restatement) timed on this host's cores on the same workload.

The default run also reports the plain 128^2 one-coil latency, the 3D
kooshball workload, the differentiable workload (forward + backward through
torch autograd) and the batched uneven CG-SENSE, each with its own figures.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2|c1|c3|c5]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "NUFFT adjoint+fwd Msamples·coil/s on MI355X; batched vs plain at ≤1e-4 rel err"
HBM_PEAK_GBS = 8000.0


def workload(name):
    from gadgetron_amd import trajectories as tj
    if name == "c2":
        traj = tj.golden_radial_2d(512, 400)
        return dict(name="2D radial 256^2, alpha 1.5 (384^2 grid), W 5.5, 400 spokes x 512, 32 coils",
                    N=[256, 256], traj=traj, C=32, seed=1, dcf=tj.ramp_dcf(traj))
    if name == "c1":
        traj = tj.golden_radial_2d(256, 64)
        return dict(name="2D radial 128^2, alpha 1.5 (192^2 grid), W 5.5, 64 spokes x 256, 1 coil",
                    N=[128, 128], traj=traj, C=1, seed=0, dcf=tj.ramp_dcf(traj))
    if name == "c3":
        traj = tj.kooshball_3d(256, 20000, seed=2)
        return dict(name="3D kooshball 128^3, alpha 1.5 (192^3 grid), W 5.5, 20000 spokes x 256, 8 coils",
                    N=[128, 128, 128], traj=traj, C=8, seed=2, dcf=tj.ramp_dcf(traj, power=2))
    if name == "c5":
        traj = tj.golden_radial_2d(512, 128)
        return dict(name="2D radial 256^2, alpha 1.5, W 5.5, 128 spokes x 512, 16 coils",
                    N=[256, 256], traj=traj, C=16, seed=4, dcf=tj.ramp_dcf(traj))
    raise SystemExit(f"unknown config {name}")


def grid_bytes(N, M, C, D):
    G = int(np.prod([int(n * 1.5) for n in N]))
    return 8 * C * (M + G) + 4 * D * M + 4 * M, G


def host_cpu():
    """CPU model and core counts of this host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count(), "usable": usable}


def cpu_baseline(wl, samples, budget_s):
    """The oracle on this host's cores: OMP_NUM_THREADS threads where the
    environment sets it (the share of the host this process is given), else
    every core this process may run on."""
    from oracle import pyoracle as po
    host = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS") or host["usable"])
    po.set_num_threads(threads)
    N, traj, C = wl["N"], wl["traj"], wl["C"]
    M = traj.shape[1]
    D = len(N)
    p = po.Plan(N, 5.5, alpha=1.5, precision="float")
    t0 = time.perf_counter()
    p.preprocess(traj)
    prep_ms = (time.perf_counter() - t0) * 1e3
    img_shape = (C, 1) + tuple(reversed(N))
    times = []
    t_end = time.perf_counter() + budget_s
    while len(times) < 2 or (time.perf_counter() < t_end and len(times) < 5):
        t0 = time.perf_counter()
        img = p.compute(samples, img_shape, po.BACKWARDS_NC2C, dcw=wl["dcf"])
        p.compute(img, (C, 1, M), po.FORWARDS_C2NC, dcw=wl["dcf"])
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    _ = D
    return {
        "value": M * C / t / 1e6,
        "unit": "Msamples·coil/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle (hoNFFT restatement, FP32, OpenMP over batches and samples, a textbook radix-2/mixed "
                  f"recursive FFT in place of FFTW) on the full workload, {len(times)} adjoint+forward pairs, "
                  f"median {t * 1e3:.1f} ms per pair, preprocess excluded",
        "cpu_model": host["model"],
        "host_nproc": host["nproc"],
        "host_usable_cores": host["usable"],
        "preprocess_ms": round(prep_ms, 1),
        "ms_per_pair": round(t * 1e3, 2),
    }


# 17 coils or more: the 2D gridding runs on the matrix cores unless GT_NUFFT_MFMA=0 (in 3D at
# every coil count, for windows of up to 6 cells)
_FP32_GRIDDING = os.environ.get("GT_NUFFT_MFMA", "1").startswith("0")
KERNELS_2D = {"nc2c": ("nc2c_reg2d_kernel" if _FP32_GRIDDING else "nc2c_mfma2d_kernel") + " + nc2c_reduce2d_kernel",
              "c2nc": "c2nc_multi2d_kernel" if _FP32_GRIDDING else "c2nc_mfma2d_kernel"}
KERNELS_3D = {"nc2c": ("nc2c_reg3d_kernel" if _FP32_GRIDDING else "nc2c_mfma3d_kernel") + " + nc2c_reduce3d_kernel",
              "c2nc": "c2nc_multi3d_kernel" if _FP32_GRIDDING else "c2nc_mfma3d_kernel"}
KERNELS_GENERIC = {"nc2c": "nc2c_tile_kernel + nc2c_reduce_kernel", "c2nc": "c2nc_gather_kernel"}
FP32_PEAK_TFLOPS = 157.3


def gridding_kernels(N):
    # the library takes its 2D and 3D hot paths for every workload here
    # (W 5.5, default tiles); 1D uses the generic kernels
    return {2: KERNELS_2D, 3: KERNELS_3D}.get(len(N), KERNELS_GENERIC)


def gridding_flops(N, M, C, W=5.5):
    """Algorithmic FP32 flops of one gridding direction: W^D cells per sample
    and coil, a complex-times-real multiply-add (4 flops) each."""
    return 4.0 * W ** len(N) * M * C


def copy_probe(mib, reps):
    """A plain device copy of known size, timed and traced so that rocprofv3's
    FETCH_SIZE / WRITE_SIZE counters can be scaled to bytes."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.randn(n, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "reps": reps}),
          flush=True)


def ndft_error(img, out, wl, rng, nsamp=512):
    """NUFFT approximation error of the run's last FORWARDS_C2NC against an FP64
    direct NDFT: coil 0, nsamp random samples, y = dcw * NDFT(x) / sqrt(G)."""
    from oracle import pyoracle as po
    N, traj, dcf = wl["N"], wl["traj"], wl["dcf"]
    G = float(np.prod([int(n * 1.5) for n in N]))
    idx = np.sort(rng.choice(traj.shape[1], size=min(nsamp, traj.shape[1]), replace=False))
    k = traj[0, idx].astype(np.float64)
    exact = po.ndft(k, N, x=img[0, 0].astype(np.complex128))[0] / np.sqrt(G) * dcf[0, idx]
    got = out[0, 0, idx].astype(np.complex128)
    return {"rel_l2": float(np.linalg.norm(got - exact) / np.linalg.norm(exact)), "samples": int(len(idx)),
            "coil": 0, "reference": "FP64 direct NDFT / sqrt(G), times dcw"}


def batched_cgsense(nprob=64, iterations=10, reps=3, seed=3):
    """Batched uneven CG-SENSE: 64 independent 2D radial reconstructions,
    N in {96,128,160,192,256}, spokes U[32,512], 2N samples per spoke, coils
    U[4,32], Gaussian-bump coil maps, 10-ellipse phantom, data = A x + CN(0,1e-3),
    10 CG iterations; the plain per-problem solver (dot products read back to
    the host) against the hipGraph-batched solver."""
    import gadgetron_amd as ga
    from gadgetron_amd import DeviceArray
    from gadgetron_amd import trajectories as tj
    keep, ops, data, xp, xb = [], [], [], [], []
    work = 0
    for pr in tj.uneven_sense_problems(nprob, seed):
        N, C, M = pr["N"], pr["C"], pr["M"]
        plan = ga.NFFTPlan([N, N], None, W=5.5, os_factor=1.5)
        d_traj = DeviceArray.from_numpy(pr["traj"])
        plan.preprocess(d_traj)
        d_csm = DeviceArray.from_numpy(pr["csm"])
        d_dcw = DeviceArray.from_numpy(pr["dcw"])
        op = ga.SenseOperator(plan, d_csm, d_dcw)
        d_x = DeviceArray.from_numpy(pr["x"])
        d_y = DeviceArray.zeros((C, 1, M), np.complex64)
        op.mult_M(d_x, d_y)
        plan.synchronize()
        d_data = DeviceArray.from_numpy((d_y.numpy() + pr["noise"]).astype(np.complex64))
        keep += [plan, d_traj, d_csm, d_dcw, d_x, d_y]
        ops.append(op)
        data.append(d_data)
        xp.append(DeviceArray.zeros((1, N, N), np.complex64))
        xb.append(DeviceArray.zeros((1, N, N), np.complex64))
        work += M * C
    xf = [DeviceArray.zeros((1, a.shape[1], a.shape[2]), np.complex64) for a in xp]
    sync = ops[0].plan.synchronize

    def timed(b):
        b.run()
        sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            b.run()
        sync()
        return (time.perf_counter() - t0) / reps

    # plain and the graph-branch mode first: the fused mode's autotuner re-tiles
    # the plans' NC2C work items (which makes batches recorded before it stale,
    # and would hand the plain loop the fused launch's item lengths)
    plain = ga.CgBatch(ops, data, xp, iterations=iterations, tc_tolerance=0.0, mode="plain")
    tp = timed(plain)
    batched = ga.CgBatch(ops, data, xb, iterations=iterations, tc_tolerance=0.0, mode="batched")
    tb = timed(batched)
    fused = ga.CgBatch(ops, data, xf, iterations=iterations, tc_tolerance=0.0, mode="fused")
    tf = timed(fused)

    def rel_to_plain(xs):
        num = sum(float(np.linalg.norm((a.numpy() - b.numpy()).ravel())) ** 2 for a, b in zip(xp, xs))
        den = sum(float(np.linalg.norm(a.numpy().ravel())) ** 2 for a in xp)
        return (num / den) ** 0.5

    # one E^H d, then per iteration one E and one E^H application
    units = work * (2 * iterations + 1) / 1e6
    best = "fused" if tf <= tb else "graph_branches"
    return {
        "problems": nprob, "iterations": iterations,
        "plain_ms": round(tp * 1e3, 3), "batched_ms": round(min(tb, tf) * 1e3, 3),
        "speedup": round(tp / min(tb, tf), 3), "batched_method": best,
        "graph_branches_ms": round(tb * 1e3, 3), "graph_branches_speedup": round(tp / tb, 3),
        "graph_branches": batched.branches,
        "fused_ms": round(tf * 1e3, 3), "fused_speedup": round(tp / tf, 3),
        "fused_nc2c_item_length": fused.item_length,
        "batched_Msamples_coil_per_s": round(units / min(tb, tf), 1),
        "plain_Msamples_coil_per_s": round(units / tp, 1),
        "rel_err_graph_branches_vs_plain": rel_to_plain(xb),
        "rel_err_fused_vs_plain": rel_to_plain(xf),
        "note": "plain: the reference's per-problem loop (dot products read back each iteration); "
                "graph_branches: each problem's whole solve in one hipGraph, problems on parallel branches; "
                "fused: one launch per step for all 64 problems (descriptor tables), captured in one hipGraph",
    }


def atomic_comparison(ga, N, d_traj, d_s, d_w, img_shape, reps=10):
    """The float-atomic NC2C (ConvolutionType::ATOMIC, ConvolverNC2C_atomic.cuh)
    kept only as a measured comparison: its NC2C stage time against the
    deterministic tile gather on the same samples, and how far its (run to run
    varying) sums differ."""
    from gadgetron_amd import DeviceArray, NFFT_comp_mode as CM
    out = {}
    imgs = {}
    for name, conv in (("standard", ga.ConvolutionType.STANDARD), ("atomic", ga.ConvolutionType.ATOMIC)):
        plan = ga.NFFTPlan(N, None, W=5.5, os_factor=1.5, conv=conv)
        plan.preprocess(d_traj)
        d_img = DeviceArray.zeros(img_shape, np.complex64)
        plan.compute(d_s, d_img, d_w, CM.BACKWARDS_NC2C)
        plan.synchronize()
        plan.profile_read()
        plan.profile(True)
        for _ in range(reps):
            plan.compute(d_s, d_img, d_w, CM.BACKWARDS_NC2C)
        plan.synchronize()
        plan.profile(False)
        st = plan.profile_read()
        out[f"{name}_nc2c_ms"] = round(st["nc2c"][0] / max(st["nc2c"][1], 1), 5)
        imgs[name] = d_img.numpy()
    a, b = imgs["standard"].ravel(), imgs["atomic"].ravel()
    out["rel_diff"] = float(np.linalg.norm(a - b) / np.linalg.norm(a))
    return out


def density_timing(ga, N, d_traj, samples_per_profile, profiles, reps=3):
    """The step before the path, on the same trajectory: iterative SDC
    (estimate_dcw: 10 Pipe-Menon iterations, Jinc kernel, os 2.1, width 5.5)
    and the golden-ratio radial Voronoi weights."""
    from gadgetron_amd import DeviceArray
    from gadgetron_amd import radial as rd
    n = samples_per_profile * profiles
    out = DeviceArray.zeros((n,), np.float32)
    rd.estimate_dcw(d_traj, N, 2.1, 10, 5.5, out=out)  # warm
    ga.lib().gt_device_synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        rd.estimate_dcw(d_traj, N, 2.1, 10, 5.5, out=out)
    ga.lib().gt_device_synchronize()
    sdc_ms = (time.perf_counter() - t0) * 1e3 / reps
    rd.dcw_golden_ratio_2d(samples_per_profile, profiles, 1.5, 1.0 / 1.5, out=out)
    ga.lib().gt_device_synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        rd.dcw_golden_ratio_2d(samples_per_profile, profiles, 1.5, 1.0 / 1.5, out=out)
    ga.lib().gt_device_synchronize()
    vor_ms = (time.perf_counter() - t0) * 1e3 / reps
    return {"samples": n, "sdc_ms": round(sdc_ms, 3), "sdc_iterations": 10,
            "radial_voronoi_dcw_ms": round(vor_ms, 3)}


def coil_map_timing(ga, N, d_traj, d_s, d_w, C, reps=3):
    """The step before CG-SENSE on the same frame: the frame buffer
    (add_frame_data = NC2C preprocess + accumulating gridding; readout =
    backwards FFT, deapodization, crop) and the coil maps estimated from its
    coil images (estimate_b1_map, 7x7 box, 2 power iterations)."""
    from gadgetron_amd import DeviceArray
    from gadgetron_amd.b1map import estimate_b1_map
    from gadgetron_amd.buffer import Buffer
    sync = ga.lib().gt_device_synchronize
    buf = Buffer()
    buf.setup(N, [int(n * 1.5) for n in N], 5.5, C, 4, 1)
    buf.set_dcw(d_w)
    imgs = DeviceArray.zeros((C,) + tuple(reversed(N)), np.complex64)
    csm = DeviceArray.zeros((C,) + tuple(reversed(N)), np.complex64)

    def timed(fn):
        fn()
        sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        sync()
        return round((time.perf_counter() - t0) * 1e3 / reps, 3)

    add_ms = timed(lambda: buf.add_frame_data(d_s, d_traj))
    read_ms = timed(lambda: buf.get_accumulated_coil_images(out=imgs))
    b1_ms = timed(lambda: estimate_b1_map(imgs, out=csm))
    if not np.all(np.isfinite(csm.numpy())):
        raise SystemExit("bench.py: non-finite coil maps")
    return {"image": list(N), "coils": C, "buffer_add_frame_ms": add_ms, "buffer_readout_ms": read_ms,
            "b1_map_ms": b1_ms}


def load_traffic(config):
    path = os.path.join(ROOT, "profiles", f"traffic_{config}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


GRID_STAGES = ("nc2c", "c2nc")


class NufftRun:
    """One workload's plan and device buffers, and the step BACKWARDS_NC2C
    (samples -> coil images) then FORWARDS_C2NC (images -> samples), DCF both
    ways, on the default stream."""

    def __init__(self, ga, wl, max_items=0):
        from gadgetron_amd import DeviceArray
        from gadgetron_amd import trajectories as tj
        self.ga, self.wl = ga, wl
        N, traj, C = wl["N"], wl["traj"], wl["C"]
        self.M = traj.shape[1]
        rng = np.random.default_rng(wl["seed"])
        self.rng = rng
        self.samples = tj.complex_normal(rng, (C, 1, self.M))
        self.plan = ga.NFFTPlan(N, None, W=5.5, os_factor=1.5)
        if max_items:
            self.plan.set_tiling(None, max_items)
        self.d_traj = DeviceArray.from_numpy(traj)
        sync = ga.lib().gt_device_synchronize
        sync()
        t0 = time.perf_counter()
        self.plan.preprocess(self.d_traj)
        sync()
        self.prep_ms = (time.perf_counter() - t0) * 1e3
        self.d_s = DeviceArray.from_numpy(self.samples)
        self.d_w = DeviceArray.from_numpy(wl["dcf"])
        self.img_shape = (C, 1) + tuple(reversed(N))
        self.d_img = DeviceArray.zeros(self.img_shape, np.complex64)
        self.d_out = DeviceArray.zeros((C, 1, self.M), np.complex64)

    def step(self):
        from gadgetron_amd import NFFT_comp_mode as CM
        self.plan.compute(self.d_s, self.d_img, self.d_w, CM.BACKWARDS_NC2C)
        self.plan.compute(self.d_img, self.d_out, self.d_w, CM.FORWARDS_C2NC)

    def timed(self, steps, warmup, sample_every=10):
        """steps timed steps between device synchronisations; HIP events on
        the compute stream around the two gridding stages of every
        sample_every-th step.  Returns (ms per step, {stage: mean ms}, launches
        sampled per stage)."""
        sync = self.ga.lib().gt_device_synchronize
        for _ in range(warmup):
            self.step()
        sync()
        self.plan.profile_read()
        t0 = time.perf_counter()
        for i in range(steps):
            rec = i % sample_every == 0
            if rec:
                self.plan.profile_stages(GRID_STAGES)
            self.step()
            if rec:
                self.plan.profile_stages(())
        sync()
        t1 = time.perf_counter()
        st = self.plan.profile_read()
        mean = {k: st[k][0] / max(st[k][1], 1) for k in GRID_STAGES}
        return (t1 - t0) * 1e3 / steps, mean, st["nc2c"][1]

    def breakdown(self, steps=10):
        """Every stage recorded on every step (each event waits for the work
        before it: these add up to more than an unrecorded step)."""
        self.plan.profile_read()
        self.plan.profile(True)
        for _ in range(steps):
            self.step()
        self.plan.profile(False)
        st = self.plan.profile_read()
        return {k: round(v[0] / steps, 5) for k, v in st.items()}


def roofline_of(wl, M, stage_ms, traffic=None):
    N, C = wl["N"], wl["C"]
    D = len(N)
    B, _ = grid_bytes(N, M, C, D)
    kernels = gridding_kernels(N)
    flops = gridding_flops(N, M, C)
    dom = "nc2c" if stage_ms["nc2c"] >= stage_ms["c2nc"] else "c2nc"
    achieved = B / (stage_ms[dom] * 1e-3) / 1e9
    out = {
        "bound": "hbm",
        "kernel": kernels[dom],
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None if traffic is None else traffic.get(dom, {}).get("bytes_per_launch"),
        "algorithmic_bytes_per_launch": B,
        "stage_ms": {k: round(v, 5) for k, v in stage_ms.items()},
        "stage_frac": {k: round(B / (v * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) for k, v in stage_ms.items()},
        "stage_kernels": kernels,
        # the FP32 vector side of the same stages (the 3D window makes gridding
        # compute-heavy: 665.5 flop per sample and coil against 121 in 2D)
        "stage_valu_tflops": {**{k: round(flops / (v * 1e-3) / 1e12, 2) for k, v in stage_ms.items()},
                              "peak": FP32_PEAK_TFLOPS},
    }
    if traffic is not None:
        out["traffic_source"] = traffic.get("measured", "profiles/traffic json")
        out["traffic_ratio"] = {k: round(traffic[k]["bytes_per_launch"] / B, 3) for k in GRID_STAGES if k in traffic}
    return out


def secondary(ga, name, steps, warmup, sample_every, with_roofline):
    """A non-headline workload's figures: step time, stage times, NDFT error."""
    wl = workload(name)
    run = NufftRun(ga, wl)
    ms, stage_ms, nsamp = run.timed(steps, warmup, sample_every)
    out = run.d_out.numpy()
    if not np.all(np.isfinite(out)):
        raise SystemExit(f"bench.py: non-finite output ({name})")
    res = {
        "workload": wl["name"],
        "ms_per_step": round(ms, 5),
        "us_per_step": round(ms * 1e3, 2),
        "Msamples_coil_per_s": round(run.M * wl["C"] / (ms * 1e-3) / 1e6, 1),
        "preprocess_ms": round(run.prep_ms, 2),
        "stage_ms": {k: round(v, 5) for k, v in stage_ms.items()},
        "stages_ms_per_step_all_recorded": run.breakdown(min(steps, 10)),
        "ndft_error": ndft_error(run.d_img.numpy(), out, wl, run.rng),
    }
    if with_roofline:
        res["roofline"] = roofline_of(wl, run.M, stage_ms, load_traffic(name))
    del run
    return res


def differentiable(iters=20):
    """The differentiable workload: loss = ||A(x; k) - y||^2 at 256^2, 16 coils,
    128 golden-angle spokes x 512 (seed 4), y from the trajectory perturbed by
    N(0, 1e-3); one iteration = forward NUFFT + backward (image gradient by the
    adjoint, trajectory gradient by the derivative-of-Kaiser-Bessel gather)."""
    import torch
    from gadgetron_amd import autograd as ag
    from gadgetron_amd import trajectories as tj
    wl = workload("c5")
    traj, C, N = wl["traj"], wl["C"], wl["N"]
    rng = np.random.default_rng(wl["seed"])
    pert = np.clip(traj + rng.normal(0, 1e-3, traj.shape), -0.5, 0.5).astype(np.float32)
    dev = torch.device("cuda", 0)
    x = torch.tensor(tj.complex_normal(rng, (C, 1, N[1], N[0])), device=dev, requires_grad=True)
    k = torch.tensor(traj, device=dev, requires_grad=True)
    with torch.no_grad():
        y = ag.nufft(x.detach(), torch.tensor(pert, device=dev), N)

    def it(kk):
        x.grad = None
        kk.grad = None
        loss = (ag.nufft(x, kk, N) - y).abs().pow(2).sum()
        loss.backward()

    def timed(fn, n):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / n

    kd = k.detach()  # the same trajectory without a gradient
    both = timed(lambda: it(k), iters)
    x_only = timed(lambda: it(kd), iters)
    fresh = timed(lambda: it(k.detach().clone().requires_grad_()), 5)
    M = traj.shape[1]
    return {
        "workload": wl["name"] + ", loss ||A(x;k) - y||^2",
        "fwd_bwd_ms": round(both, 3),
        "fwd_bwd_Msamples_coil_per_s": round(M * C / (both * 1e-3) / 1e6, 1),
        "fwd_bwd_image_grad_only_ms": round(x_only, 3),
        "fwd_bwd_new_trajectory_ms": round(fresh, 3),
        "note": "fwd_bwd: image and trajectory gradients, trajectory unchanged between iterations (its plan "
                "is reused); image_grad_only: the trajectory without a gradient; new_trajectory: a new "
                "trajectory tensor per iteration, so each pays preprocess; gradient accuracy is tested in "
                "tests/test_gpu_autograd.py",
    }


def differentiable_child():
    """differentiable() in a child process.  torch brings its own HIP runtime
    with the same soname as the one libgtnufft.so links, and a process keeps
    whichever loads first; this process loaded the library's, so torch runs in
    a process of its own that imports it first."""
    import subprocess
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--only-c5"], capture_output=True, text=True,
                       timeout=600)
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
    if r.returncode != 0 or not lines:
        return {"error": f"child exited {r.returncode}: {r.stderr[-400:]}"}
    return json.loads(lines[-1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cg", action="store_true", help="skip the batched CG-SENSE comparison")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the other workloads and side timings (profiling runs)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run only a device copy of this many MiB (counter calibration)")
    ap.add_argument("--max-items", type=int, default=0,
                    help="entries per gridding work item (gt_nufft_set_tiling); 0 keeps the default")
    ap.add_argument("--only-c5", action="store_true",
                    help="print only the differentiable workload's figures (run as a child process)")
    args = ap.parse_args()
    if args.only_c5:
        import torch  # noqa: F401  (torch's HIP runtime first; see differentiable_child)
        print(json.dumps(differentiable()), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the NUFFT shards by coil "
              "with no exchange between GPUs).  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib, 5)
        return

    import gadgetron_amd as ga

    if ga.device_count() < 1:
        raise SystemExit("bench.py: no GPU visible")
    wl = workload(args.config)
    N, C = wl["N"], wl["C"]
    D = len(N)
    run = NufftRun(ga, wl, args.max_items)
    M = run.M
    # 3D steps take milliseconds: an event pair per gridding stage of every step costs ~0.1 %
    every = 1 if D == 3 else 10
    ms_per_step, stage_ms, sampled = run.timed(args.steps, args.warmup, every)
    value = M * C / (ms_per_step * 1e-3) / 1e6
    roofline = roofline_of(wl, M, stage_ms, load_traffic(args.config))
    roofline["stage_launches_timed"] = sampled
    stages_all = run.breakdown(10)

    out = run.d_out.numpy()
    if not np.all(np.isfinite(out)):
        raise SystemExit("bench.py: non-finite output")
    ndft_err = ndft_error(run.d_img.numpy(), out, wl, run.rng)

    extras = not args.no_extras
    atomic = atomic_comparison(ga, N, run.d_traj, run.d_s, run.d_w, run.img_shape) if D == 2 and extras else None
    density = density_timing(ga, N, run.d_traj, 512, 400) if args.config == "c2" and extras else None
    coil_maps = coil_map_timing(ga, N, run.d_traj, run.d_s, run.d_w, C) if args.config == "c2" and extras else None

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(wl, run.samples, args.cpu_seconds)
    prep_ms = run.prep_ms
    del run

    others = {}
    if args.config == "c2" and extras:
        for name, fn in (("c1_latency", lambda: secondary(ga, "c1", 2000, 500, 10, False)),
                         ("c3", lambda: secondary(ga, "c3", 10, 2, 1, True)),
                         ("c5_differentiable", differentiable_child)):
            try:
                others[name] = fn()
            except Exception as e:  # reported, never hides the headline number
                others[name] = {"error": f"{type(e).__name__}: {e}"}

    cgb = None
    if not args.no_cg:
        try:
            cgb = batched_cgsense()
        except Exception as e:  # reported, never hides the headline number
            cgb = {"error": f"{type(e).__name__}: {e}"}

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Msamples·coil/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": wl["name"], "N": N, "N_os": [int(n * 1.5) for n in N], "W": 5.5,
                   "samples": M, "coils": C, "step": "BACKWARDS_NC2C + FORWARDS_C2NC, DCF both ways"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched_cgsense": cgb,
        "nc2c_atomic_comparison": atomic,
        "density_compensation": density,
        "coil_maps": coil_maps,
        "ndft_error": ndft_err,
        "stages_ms_per_step_all_recorded": stages_all,
        "preprocess_ms": round(prep_ms, 2),
        **others,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
