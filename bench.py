#!/usr/bin/env python3
"""Benchmark of the RdfPlan / RmsdPlan hot path on one MI355X.

One step = one pass of the hot path over one batch of synthetic input, with the
inputs already resident in HBM:
  * RdfPlan, OW-OW: 100,000 waters (the OW atoms of a 300,000-atom system) in a
    cubic 144.3 A box, 1,000 float32 frames (47^3 lattice sites, a seeded choice
    of 100,000, sigma = 0.3 A jitter per frame), 500 bins, r_max 15 A;
  * batched RmsdPlan: 5,000 frames of a 20,000-atom chain, fit on the 2,000 CA
    atoms, RMSD over all atoms, float32 coordinates, FP64 QCP.
value = reference-equivalent RDF pair evaluations (frames x (N_a N_b - common))
per second of the whole step.  The RMSD frames/s, the triclinic RDF rate and
the per-kernel roofline figures are reported beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import ctypes
import json
import math
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "warp-md_amd"))

VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12  # lane-ops/s: 256 CU x 4 SIMD x 16 lanes x 2.4 GHz
HBM_PEAK_GBS = 8000.0
MFMA_F64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 matrix
RDF_OPS_PER_PAIR = 24  # algorithmic VALU lane-ops per unordered in-range pair


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    """The host CPU's model string (/proc/cpuinfo, else lscpu)."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.lower().startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor()


def host_threads():
    """Threads for the multi-core CPU lines: the process's CPU share (the
    OMP_NUM_THREADS the host sets for one GPU's job, within the affinity
    mask), with nproc and the mask size recorded beside it."""
    nproc = os.cpu_count() or 1
    try:
        mask = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        mask = nproc
    share = int(os.environ.get("OMP_NUM_THREADS", mask) or mask)
    threads = max(1, min(share, mask))
    return threads, {"nproc": nproc, "affinity_cpus": mask, "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
                     "threads_used": threads, "cpu": cpu_model()}


def traffic_per_launch(kernel, frames):
    """HBM bytes per launch of `kernel` over `frames` frames, from the committed
    FETCH_SIZE / WRITE_SIZE counter runs (profiles/hbm_traffic.json: KiB per
    dispatch on a known frame count, scaled by what the counters read for a
    plain copy of known size); None when the file is absent."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as fh:
            t = json.load(fh)
        k = t[kernel]
    except (OSError, KeyError, ValueError):
        return None
    per_frame = (k["fetch_kib"] * t["fetch_scale"] + k["write_kib"] * t["write_scale"]) * 1024.0 / k["frames"]
    return per_frame * frames


def traffic_stamp():
    """Commit and date of the counter runs behind traffic_per_launch (the
    counters need rocprofv3 around the process, so they are not re-read in
    a bench run); None when unrecorded."""
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as fh:
            t = json.load(fh)
        return t.get("measured_at")
    except (OSError, ValueError):
        return None


def cpu_baseline(n_side=47, box=144.3, n_waters=100000):
    """The oracle (a C restatement of the reference's single-threaded CPU pair
    loop) timed on a bounded sample: frame 0 of the same water box, the first
    rows of sel_a against all of sel_b."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle
    from warp_md_amd import workloads

    w = workloads.water_lattice(n_side=n_side, box=box, n_frames=1, n_waters=n_waters, seed=1,
                                with_h=False)
    n = w["coords"].shape[1]
    allat = np.arange(n)
    # calibrate to ~10 s of single-thread work
    probe = 500
    t0 = time.perf_counter()
    oracle.rdf_counts(w["coords"], allat[:probe], allat, 500, 15.0, boxes=w["box"])
    dt = max(time.perf_counter() - t0, 1e-6)
    rate = probe * (n - 1) / dt
    rows = int(min(n, max(probe, 10.0 * rate / n)))
    t0 = time.perf_counter()
    oracle.rdf_counts(w["coords"], allat[:rows], allat, 500, 15.0, boxes=w["box"])
    dt = time.perf_counter() - t0
    evals = rows * (n - 1)
    out = {"value": evals / dt, "unit": "pair-evals/s", "cores": 1, "kind": "port",
           "sample": f"frame 0 of the 100,000-OW box, sel_a = first {rows} OW x all {n} OW "
                     f"({evals:.3g} pair-evals, {dt:.1f} s, 1 thread, brute force in the "
                     f"reference loop order)",
           "cpu": cpu_model()}
    # multi-core line of the same loop (OpenMP over sel_a) on the process's
    # CPU share, labelled with the thread count and the host's core counts
    threads, host = host_threads()
    rows2 = int(min(n, rows * threads))
    t0 = time.perf_counter()
    oracle.rdf_counts(w["coords"], allat[:rows2], allat, 500, 15.0, boxes=w["box"], method="omp",
                      threads=threads)
    dt2 = time.perf_counter() - t0
    out["host"] = host
    out["all_cores"] = {"value": rows2 * (n - 1) / dt2, "threads": threads,
                        "label": f"{threads} threads (the job's CPU share) of {host['nproc']} host CPUs",
                        "sample": f"sel_a = first {rows2} OW x all {n} OW ({rows2 * (n - 1):.3g} pair-evals, "
                                  f"{dt2:.1f} s)"}
    # RMSD frames/s of the oracle's FP64 Kabsch (fit on CA, RMSD over all atoms)
    c = workloads.chain(n_frames=100, seed=2)
    t0 = time.perf_counter()
    oracle.fit_rmsd(c["coords"], c["ca"], np.arange(c["ref"].shape[0]), c["ref"])
    out["rmsd_frames_per_s"] = 100 / (time.perf_counter() - t0)
    return out


def small_box_leg(ctx, dev, reps=20):
    """RdfPlan OW-OW on 1,000 rigid waters (3,000 atoms, cubic 31.04 A, 100
    frames, 200 bins, r_max 10): the device path on a resident store against
    the oracle's reference loop over all 100 frames, single-threaded and over
    the host cores; counts compared."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import torch

    import oracle
    from warp_md_amd import workloads
    from warp_md_amd._native import check, lib, ptr

    L = lib()
    w = workloads.water_lattice(n_side=10, box=31.04, n_frames=100, seed=0)
    coords, ow = w["coords"], np.asarray(w["ow"])
    F, n = coords.shape[0], ow.size
    stride = (n + 3) // 4 * 4
    soa = np.zeros((F, 3, stride), np.float32)
    soa[:, :, :n] = np.transpose(coords[:, ow, :], (0, 2, 1))
    store = torch.as_tensor(soa).to(dev)
    boxes = np.repeat(np.diag(np.asarray(w["box"], np.float32)).reshape(1, 9), F, 0).astype(np.float32)
    sel = np.arange(n, dtype=np.uint32)
    h = ctypes.c_void_p()
    check(L.wmd_rdf_create(ctx.handle, ptr(sel), n, ptr(sel), n, 200, ctypes.c_float(10.0), 1, 0,
                           ctypes.byref(h)))
    V = ctypes.c_void_p(store.data_ptr())
    check(L.wmd_rdf_accumulate_device(h, V, stride, F, ptr(boxes), None))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        check(L.wmd_rdf_reset(h))
        check(L.wmd_rdf_accumulate_device(h, V, stride, F, ptr(boxes), None))
    torch.cuda.synchronize()
    gpu_s = (time.perf_counter() - t0) / reps
    counts = np.zeros(200, np.uint64)
    check(L.wmd_rdf_read(h, ptr(counts), None))
    L.wmd_rdf_destroy(h)
    evals = F * (n * n - n)
    t0 = time.perf_counter()
    ref = oracle.rdf_counts(coords, ow, ow, 200, 10.0, boxes=w["box"])
    cpu_s = time.perf_counter() - t0
    threads, _ = host_threads()
    t0 = time.perf_counter()
    ref_omp = oracle.rdf_counts(coords, ow, ow, 200, 10.0, boxes=w["box"], method="omp", threads=threads)
    omp_s = time.perf_counter() - t0
    same = bool(np.array_equal(counts, ref) and np.array_equal(ref_omp, ref))
    log(f"1,000 waters x 100 frames: GPU {gpu_s * 1e3:.3f} ms ({evals / gpu_s:.3g} pair-evals/s), "
        f"oracle 1 thread {cpu_s:.2f} s, {threads} threads {omp_s:.2f} s, counts equal: {same}")
    return {"pair_evals": evals, "gpu_ms": gpu_s * 1e3, "gpu_pair_evals_per_s": evals / gpu_s,
            "cpu_1thread_pair_evals_per_s": evals / cpu_s,
            "cpu_all_cores": {"pair_evals_per_s": evals / omp_s, "threads": threads},
            "counts_equal": same, "in_range_pairs": int(counts.sum())}


def soft_rdf_ratio(ctx, dev, n=20000, frames=64, reps=20):
    """Soft-RDF forward + backward against the plain RDF forward on the same
    input: 20,000 beads at 0.08 / A^3 (cubic 62.996 A), 64 frames of a uniform
    base configuration plus N(0, 0.2^2) jitter, 300 bins, r_max 12 A, kernel
    sigma 0.05 A; the loss is sum_k (g_k - g_t(r_k))^2 plus the fitted RMSD to
    frame 0 (whose forward + backward is reported beside)."""
    import numpy as np
    import torch

    from warp_md_amd import autograd
    from warp_md_amd._native import check, lib, ptr

    L = lib()
    V = lambda t: ctypes.c_void_p(t.data_ptr())
    box = (n / 0.08) ** (1.0 / 3.0)
    gen = torch.Generator(device=dev)
    gen.manual_seed(4)
    base = torch.rand((n, 3), generator=gen, device=dev, dtype=torch.float32) * box
    x = base[None] + 0.2 * torch.randn((frames, n, 3), generator=gen, device=dev, dtype=torch.float32)
    stride = (n + 3) // 4 * 4
    soa = torch.zeros((frames, 3, stride), dtype=torch.float32, device=dev)
    soa[:, :, :n] = x.permute(0, 2, 1)
    boxes = np.repeat(np.diag([box] * 3).reshape(1, 9), frames, 0).astype(np.float32)
    sel = np.arange(n, dtype=np.uint32)
    h = ctypes.c_void_p()
    check(L.wmd_rdf_create(ctx.handle, ptr(sel), n, ptr(sel), n, 300, ctypes.c_float(12.0), 1, 0,
                           ctypes.byref(h)))
    plan = autograd.SoftRdf(bins=300, r_max=12.0, sigma=0.05, pbc="orthorhombic")
    norm = torch.as_tensor(plan.normalisation(n, frames, np.array([box] * 3)), device=dev)
    target = torch.as_tensor(autograd.g_target(plan.bin_centers), device=dev)
    rmsd_plan = autograd.FittedRmsd(x[0])

    def plain():
        check(L.wmd_rdf_reset(h))
        check(L.wmd_rdf_accumulate_device(h, V(soa), stride, frames, ptr(boxes), None))

    two_over_norm, neg_target = 2.0 / norm, -target

    def soft_step():
        # L = sum (c / norm - g_t)^2, dL/dc = 2 (c / norm - g_t) / norm; the
        # loss stays on the device (read after the timed steps)
        c = plan.forward_store(soa, stride, n, boxes)
        d = torch.addcdiv(neg_target, c, norm)
        return plan.backward_store(soa, stride, n, boxes, d * two_over_norm), torch.dot(d, d)

    def rmsd_step():
        xr = x.clone().requires_grad_(True)
        r = rmsd_plan(xr)
        r.sum().backward()
        return xr.grad

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps * 1e3, out

    plain_ms, _ = timed(plain)
    soft_ms, (grad, loss) = timed(soft_step)
    loss = float(loss)
    rmsd_ms, _ = timed(rmsd_step)
    path = plan.last_path
    # kernel breakdown of one more soft step (HIP events)
    ctx.profile(True)
    soft_step()
    torch.cuda.synchronize()
    parts = {k: ctx.profile_read(k)[0] for k in ("softrdf_fwd", "softrdf_bwd", "cell_build")}
    ctx.profile(False)
    L.wmd_rdf_destroy(h)
    log(f"soft rdf: plain forward {plain_ms:.2f} ms, soft fwd+bwd {soft_ms:.2f} ms "
        f"(ratio {soft_ms / plain_ms:.2f}), fitted RMSD fwd+bwd {rmsd_ms:.2f} ms, path {path}")
    return {"ratio": soft_ms / plain_ms, "plain_fwd_ms": plain_ms, "soft_fwd_bwd_ms": soft_ms,
            "fitted_rmsd_fwd_bwd_ms": rmsd_ms, "loss": loss,
            "grad_norm": float(grad.norm()), "path": path[0], "cells": path[1],
            "kernel_ms": parts,
            "workload": "20,000 beads, cubic 62.996 A, 64 frames, 300 bins, r_max 12 A, sigma 0.05 A"}


def pairwise_bench(ctx, dev, F=5000, reps=3, cpu_frames=600):
    """PairwiseRmsdPlan (metric rms) over the 5,000 chain frames, fit on the
    2,000 CA atoms: the 5000 x 5000 RMSD matrix from a resident store.  The
    Gram kernel's algorithmic work is 9 cross-covariance entries x n atoms x
    2 flops per frame pair i < j; the CPU sample is the oracle's
    PairwiseRmsdPlan on the first `cpu_frames` frames (one thread)."""
    import numpy as np
    import torch

    from warp_md_amd import workloads
    from warp_md_amd._native import check, lib, ptr

    L = lib()
    c = workloads.chain(n_frames=0, seed=2)
    ca = c["ca"]
    n = int(ca.size)
    store = workloads.chain_store_device(c["ref"], ca, F, seed=2, device=dev)
    sel = np.arange(n, dtype=np.uint32)
    h = ctypes.c_void_p()
    check(L.wmd_pairwise_create(ctx.handle, ptr(sel), n, 0, ctypes.byref(h)))
    out = torch.empty((F, F), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream

    def run():
        check(L.wmd_pairwise_run_device(h, ctypes.c_void_p(store.data_ptr()), store.shape[2], F,
                                        ctypes.c_void_p(out.data_ptr()), None,
                                        ctypes.c_void_p(stream)))

    run()
    torch.cuda.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        run()
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) / reps * 1e3
    gram_ms, gram_n = ctx.profile_read("pairwise_gram")
    cen_ms, _ = ctx.profile_read("pairwise_center")
    ctx.profile(False)
    L.wmd_pairwise_destroy(h)
    gram_avg = gram_ms / max(gram_n, 1)
    pairs = F * (F - 1) / 2
    flops = pairs * 9 * n * 2
    ach = flops / (gram_avg / 1e3) / 1e12
    m = out.cpu().numpy()
    frames = np.transpose(store[:cpu_frames].cpu().numpy(), (0, 2, 1))
    del store, out
    torch.cuda.empty_cache()
    cpu = None
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        t0 = time.perf_counter()
        oracle.pairwise_rmsd(frames, np.arange(n), "rms")
        dt = time.perf_counter() - t0
        cpu = {"value": cpu_frames * (cpu_frames - 1) / 2 / dt, "unit": "frame pairs/s",
               "cores": 1, "kind": "port",
               "sample": f"first {cpu_frames} frames ({cpu_frames * (cpu_frames - 1) // 2} pairs, "
                         f"{dt:.1f} s, 1 thread, SVD Kabsch per pair as the reference)"}
    except Exception as e:  # the oracle is a checker; its absence leaves the GPU line intact
        log(f"pairwise cpu baseline skipped: {e}")
    log(f"pairwise rmsd: {F} x {F} matrix, {n} atoms: {wall_ms:.2f} ms wall, gram kernel "
        f"{gram_avg:.2f} ms ({ach:.1f} TFLOP/s f64), centring {cen_ms / max(reps, 1):.3f} ms")
    return {"frames": F, "atoms": n, "ms": wall_ms, "pairs_per_s": pairs / (wall_ms / 1e3),
            "checks": {"mean": float(m[np.triu_indices(F, 1)].mean()),
                       "symmetric": bool(np.array_equal(m, m.T))},
            "roofline": {"kernel": "pw_gram_kernel", "bound": "mfma", "achieved": ach,
                         "peak": MFMA_F64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": ach / MFMA_F64_PEAK_TFLOPS,
                         "traffic": traffic_per_launch("pw_gram_kernel_chain_5000f_2000ca", F),
                         "algorithmic": f"9 x {n} x 2 f64 flops per frame pair x {pairs:.4g} pairs",
                         "avg_launch_ms": gram_avg},
            "cpu_baseline": cpu,
            "workload": "PairwiseRmsdPlan metric rms, 5,000 chain frames, fit on 2,000 CA"}


def align_bench(ctx, dev, F=5000, reps=3, cpu_frames=40):
    """SuperposePlan over the 5,000 frames of the 20,000-atom chain, fit on the
    2,000 CA atoms, from device frames [F][n][4]: one workgroup per frame
    fits in f64 and writes R x + t for every atom from the frame held in
    registers and LDS (al_superpose_reg_kernel; HBM-bound,
    algorithmic 16 B read + 12 B written per atom)."""
    import numpy as np
    import torch

    from warp_md_amd import workloads
    from warp_md_amd._native import check, lib, ptr

    L = lib()
    c = workloads.chain(n_frames=0, seed=2)
    ref, ca = c["ref"], c["ca"]
    n = ref.shape[0]
    store = workloads.chain_store_device(ref, np.arange(n, dtype=np.uint32), F, seed=2, device=dev)
    x4 = torch.ones((F, n, 4), dtype=torch.float32, device=dev)
    x4[:, :, :3] = store.permute(0, 2, 1)
    del store
    ref4 = np.ones((n, 4), np.float32)
    ref4[:, :3] = ref
    h = ctypes.c_void_p()
    check(L.wmd_align_create(ctx.handle, ptr(ca), ca.size, None, ptr(ref4), n, 0, ctypes.byref(h)))
    rt64 = torch.zeros((F, 12), dtype=torch.float64, device=dev)
    xyz = torch.empty((F, n, 3), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream

    def run():
        check(L.wmd_align_run_device(h, ctypes.c_void_p(x4.data_ptr()), n, F,
                                     ctypes.c_void_p(rt64.data_ptr()), None,
                                     ctypes.c_void_p(xyz.data_ptr()), ctypes.c_void_p(stream)))

    run()
    torch.cuda.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        run()
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) / reps * 1e3
    app_ms, app_n = ctx.profile_read("align_superpose")
    ctx.profile(False)
    L.wmd_align_destroy(h)
    app_avg = app_ms / max(app_n, 1)
    algo = 28.0 * n * F
    ach = algo / (app_avg / 1e3) / 1e9
    sample = x4[:cpu_frames, :, :3].cpu().numpy()
    del x4, xyz
    torch.cuda.empty_cache()
    cpu = None
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        t0 = time.perf_counter()
        oracle.align(sample, ca, ref, superpose=True)
        dt = time.perf_counter() - t0
        cpu = {"value": cpu_frames / dt, "unit": "frames/s", "cores": 1, "kind": "port",
               "sample": f"first {cpu_frames} frames ({dt:.2f} s, 1 thread, SVD Kabsch + f64 transform)"}
    except Exception as e:
        log(f"align cpu baseline skipped: {e}")
    log(f"superpose: {F} frames x {n} atoms (fit {ca.size}): {wall_ms:.2f} ms wall, fused "
        f"fit + transform kernel {app_avg:.3f} ms ({ach:.0f} GB/s)")
    return {"frames": F, "atoms": n, "fit_atoms": int(ca.size), "ms": wall_ms,
            "frames_per_s": F / (wall_ms / 1e3),
            "kernel_ms": {"align_superpose": app_avg},
            "roofline": {"kernel": "al_superpose_reg_kernel (fit + transform)", "bound": "hbm", "achieved": ach,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                         "traffic": traffic_per_launch("al_superpose_reg_kernel_chain_5000f_20k", F),
                         "traffic_measured_at": traffic_stamp(),
                         "algorithmic": f"28 B per atom (16 read, 12 written) x {n * F:.4g} atoms",
                         "avg_launch_ms": app_avg},
            "cpu_baseline": cpu,
            "workload": "SuperposePlan, 5,000 chain frames x 20,000 atoms, fit on 2,000 CA"}


def host_ingest_bench(ctx, frames=96, rmsd_frames=1000, reps=2):
    """The host-buffer entry points end to end (PCIe and host copies included):
    RdfPlan's full chunks of the 100,000-water system (300,000 atoms, OW
    selected on the device) through wmd_rdf_accumulate, and RmsdPlan's
    20,000-atom chain chunks through wmd_rmsd_run.  These rates are reported
    beside the bench value, never as it."""
    import numpy as np

    from warp_md_amd import workloads
    from warp_md_amd._native import check, lib, ptr

    L = lib()
    w = workloads.water_lattice(n_side=47, box=144.3, n_frames=8, n_waters=100000, seed=1)
    N = w["coords"].shape[1]
    x4 = np.ones((frames, N, 4), np.float32)
    for f in range(frames):
        x4[f, :, :3] = w["coords"][f % 8]
    boxes = np.repeat(np.diag([144.3] * 3).reshape(1, 9), frames, 0).astype(np.float32)
    ow = w["ow"]
    h = ctypes.c_void_p()
    check(L.wmd_rdf_create(ctx.handle, ptr(ow), ow.size, ptr(ow), ow.size, 500, ctypes.c_float(15.0),
                           1, 0, ctypes.byref(h)))
    check(L.wmd_rdf_accumulate(h, ptr(x4[:8]), N, 8, ptr(boxes[:8])))  # warm-up
    ctx.profile(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        check(L.wmd_rdf_accumulate(h, ptr(x4), N, frames, ptr(boxes)))
    rdf_s = (time.perf_counter() - t0) / reps
    # the device-side selection gather (300,000-atom AoS frames -> the
    # 100,000 OW SoA store the pair kernels read), timed on its own
    gather_ms, gather_n = ctx.profile_read("rdf_gather")
    ctx.profile(False)
    gather_frame_ms = gather_ms / (reps * frames)
    gather = {"ms_per_frame": gather_frame_ms, "ms_per_1000_frames": gather_frame_ms * 1000.0,
              "launches": gather_n,
              "algorithmic_GBps": (16.0 * ow.size * 3 + 12.0 * ow.size) / (gather_frame_ms / 1e3) / 1e9,
              "note": "reads the 300,000-atom frame's rows (OW every third atom: every line) and writes "
                      "the OW SoA store; beside the headline, which starts from that store"}
    host_counts = np.zeros(500, np.uint64)
    check(L.wmd_rdf_read(h, ptr(host_counts), None))
    L.wmd_rdf_destroy(h)
    # The pipeline's counts (the 8 distinct frames once in the warm-up, then
    # frames / 8 times per rep) must equal the device-resident path's counts
    # of the same 8 frames times that multiplicity, bit for bit.
    import torch
    stride = (ow.size + 3) // 4 * 4
    st8 = torch.zeros((8, 3, stride), dtype=torch.float32, device="cuda:0")
    st8[:, :, :ow.size] = torch.as_tensor(np.ascontiguousarray(w["coords"][:, ow].transpose(0, 2, 1))).cuda()
    hd = ctypes.c_void_p()
    check(L.wmd_rdf_create(ctx.handle, ptr(ow), ow.size, ptr(ow), ow.size, 500, ctypes.c_float(15.0),
                           1, 0, ctypes.byref(hd)))
    dev_counts = np.zeros(500, np.uint64)
    check(L.wmd_rdf_accumulate_device(hd, ctypes.c_void_p(st8.data_ptr()), stride, 8, ptr(boxes[:8]),
                                      ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    check(L.wmd_rdf_read(hd, ptr(dev_counts), None))
    L.wmd_rdf_destroy(hd)
    del st8
    mult = 1 + reps * (frames // 8)
    rdf_match = frames % 8 == 0 and np.array_equal(host_counts, dev_counts * np.uint64(mult))
    if not rdf_match:
        raise AssertionError("host-ingest RDF counts differ from the device-resident path")
    # the same through the Python plan API (Trajectory holds the FrameChunk layout)
    from warp_md_amd import RdfPlan, System, Trajectory
    traj = Trajectory(x4, boxes)
    sysm = System.from_arrays({"name": np.array(w["names"])})
    plan = RdfPlan(ow, ow, bins=500, r_max=15.0, pbc="orthorhombic")
    plan.run(traj, sysm, device="cuda:0")
    t0 = time.perf_counter()
    plan.run(traj, sysm, device="cuda:0")
    plan_s = time.perf_counter() - t0
    del x4, traj
    c = workloads.chain(n_frames=rmsd_frames, seed=2)
    nc = c["ref"].shape[0]
    cx4 = np.ones((rmsd_frames, nc, 4), np.float32)
    cx4[:, :, :3] = c["coords"]
    r4 = np.ones((nc, 4), np.float32)
    r4[:, :3] = c["ref"]
    allat = np.arange(nc, dtype=np.uint32)
    hr = ctypes.c_void_p()
    check(L.wmd_rmsd_create(ctx.handle, ptr(c["ca"]), c["ca"].size, ptr(allat), allat.size, ptr(r4), nc,
                            1, ctypes.byref(hr)))
    out = np.zeros(rmsd_frames, np.float32)
    check(L.wmd_rmsd_run(hr, ptr(cx4[:8]), nc, 8, ptr(out[:8]), None))
    t0 = time.perf_counter()
    for _ in range(reps):
        check(L.wmd_rmsd_run(hr, ptr(cx4), nc, rmsd_frames, ptr(out), None))
    rmsd_s = (time.perf_counter() - t0) / reps
    # the pipeline's RMSD series against the device-resident path (same kernels
    # on a store built on the host; only the tiling of the sums may differ)
    n_store = ctypes.c_size_t()
    rstride = ctypes.c_size_t()
    check(L.wmd_rmsd_store_layout(hr, None, 0, ctypes.byref(n_store), ctypes.byref(rstride)))
    slots = np.zeros(rstride.value, np.uint32)
    check(L.wmd_rmsd_store_layout(hr, ptr(slots), slots.size, ctypes.byref(n_store), ctypes.byref(rstride)))
    valid = slots != 0xFFFFFFFF
    host_store = np.zeros((rmsd_frames, 3, slots.size), np.float32)
    host_store[:, :, valid] = np.transpose(c["coords"][:, slots[valid].astype(np.int64)], (0, 2, 1))
    dstore = torch.as_tensor(host_store).cuda()
    del host_store
    dout = torch.zeros(rmsd_frames, dtype=torch.float32, device="cuda:0")
    check(L.wmd_rmsd_run_device(hr, ctypes.c_void_p(dstore.data_ptr()), slots.size, rmsd_frames,
                                ctypes.c_void_p(dout.data_ptr()), None, None, None, None,
                                ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    torch.cuda.synchronize()
    rmsd_diff = float(np.max(np.abs(dout.cpu().numpy().astype(np.float64) - out.astype(np.float64))))
    del dstore, dout
    if not rmsd_diff <= 2e-6:
        raise AssertionError(f"host-ingest RMSD differs from the device-resident path by {rmsd_diff:.3e} A")
    L.wmd_rmsd_destroy(hr)
    # SuperposePlan's host path: the same chain chunks in, every atom superposed out
    ha = ctypes.c_void_p()
    check(L.wmd_align_create(ctx.handle, ptr(c["ca"]), c["ca"].size, None, ptr(r4), nc, 0, ctypes.byref(ha)))
    rt = np.zeros((rmsd_frames, 12), np.float32)
    xyz = np.zeros((rmsd_frames, nc, 3), np.float32)
    check(L.wmd_align_run(ha, ptr(cx4[:8]), nc, 8, ptr(rt[:8]), ptr(xyz[:8])))
    t0 = time.perf_counter()
    for _ in range(reps):
        check(L.wmd_align_run(ha, ptr(cx4), nc, rmsd_frames, ptr(rt), ptr(xyz)))
    sup_s = (time.perf_counter() - t0) / reps
    L.wmd_align_destroy(ha)
    # PairwiseRmsdPlan's host path: the chunks' CA appended to the device store
    hp = ctypes.c_void_p()
    check(L.wmd_pairwise_create(ctx.handle, ptr(c["ca"]), c["ca"].size, 0, ctypes.byref(hp)))
    check(L.wmd_pairwise_append(hp, ptr(cx4[:8]), nc, 8))
    t0 = time.perf_counter()
    for _ in range(reps):
        check(L.wmd_pairwise_append(hp, ptr(cx4), nc, rmsd_frames))
    app_s = (time.perf_counter() - t0) / reps
    L.wmd_pairwise_destroy(hp)
    pair_evals = frames * (float(ow.size) * ow.size - ow.size)
    res = {"rdf": {"frames": frames, "atoms_per_frame": N, "s": rdf_s,
                   "host_GB_per_s": frames * N * 16 / rdf_s / 1e9,
                   "pair_evals_per_s": pair_evals / rdf_s,
                   "rdfplan_run_s": plan_s,
                   "counts_equal_device_path": bool(rdf_match),
                   "device_selection_gather": gather},
           "rmsd": {"frames": rmsd_frames, "atoms_per_frame": nc, "s": rmsd_s,
                    "max_abs_diff_vs_device_path_A": rmsd_diff,
                    "host_GB_per_s": rmsd_frames * nc * 16 / rmsd_s / 1e9,
                    "frames_per_s": rmsd_frames / rmsd_s},
           "superpose": {"frames": rmsd_frames, "atoms_per_frame": nc, "s": sup_s,
                         "host_GB_per_s": rmsd_frames * nc * 28 / sup_s / 1e9},
           "pairwise_append": {"frames": rmsd_frames, "atoms_per_frame": nc, "s": app_s,
                               "host_GB_per_s": rmsd_frames * nc * 16 / app_s / 1e9}}
    log(f"host ingest: RDF {frames} x {N} atoms {rdf_s * 1e3:.1f} ms (RdfPlan.run {plan_s * 1e3:.1f} ms, "
        f"device selection gather {gather_frame_ms * 1e3:.1f} us per frame) "
        f"({res['rdf']['host_GB_per_s']:.1f} GB/s of chunks, {res['rdf']['pair_evals_per_s']:.3g} pair-evals/s); "
        f"RMSD {rmsd_frames} x {nc} atoms {rmsd_s * 1e3:.1f} ms ({res['rmsd']['host_GB_per_s']:.1f} GB/s); "
        f"superpose {rmsd_frames} x {nc} atoms {sup_s * 1e3:.1f} ms ({res['superpose']['host_GB_per_s']:.1f} GB/s in + out); "
        f"pairwise append {rmsd_frames} x {nc} atoms {app_s * 1e3:.1f} ms ({res['pairwise_append']['host_GB_per_s']:.1f} GB/s)")
    return res


def sf_bymol_bench(ctx, dev, reps=2):
    """Rates beside the bench value: StructureFactorPlan's
    f64 pair counts on 20,000 atoms x 8 frames from a resident store
    (rdf_brute64_kernel), and RdfPlan bymol (centres of 100,000 3-site waters)
    from host chunks of the full 300,000-atom system (group centres, cell-list
    pairs, PCIe included)."""
    import numpy as np
    import torch

    from warp_md_amd import RdfPlan, System, Trajectory, workloads
    from warp_md_amd._native import check, lib, ptr

    L = lib()
    out = {}
    n, F, box = 20000, 8, 62.996
    gen = torch.Generator(device=dev)
    gen.manual_seed(9)
    store = torch.zeros((F, 3, n), dtype=torch.float32, device=dev)
    store[:] = torch.rand((F, 3, n), generator=gen, device=dev) * box
    sel = np.arange(n, dtype=np.uint32)
    boxes = np.repeat(np.diag([box] * 3).reshape(1, 9), F, 0).astype(np.float32)
    h = ctypes.c_void_p()
    check(L.wmd_rdf_create_f64(ctx.handle, ptr(sel), n, 300, ctypes.c_double(12.0), 1, ctypes.c_double(1.0),
                               ctypes.byref(h)))
    check(L.wmd_rdf_accumulate_device(h, ctypes.c_void_p(store.data_ptr()), n, F, ptr(boxes), None))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        check(L.wmd_rdf_accumulate_device(h, ctypes.c_void_p(store.data_ptr()), n, F, ptr(boxes), None))
    torch.cuda.synchronize()
    sf_s = (time.perf_counter() - t0) / reps
    L.wmd_rdf_destroy(h)
    out["structure_factor_f64"] = {"atoms": n, "frames": F, "ms": sf_s * 1e3,
                                   "pair_evals_per_s": F * n * (n - 1.0) / sf_s}
    del store
    w = workloads.water_lattice(n_side=47, box=144.3, n_frames=4, n_waters=100000, seed=1)
    nw = w["coords"].shape[1]
    x = np.concatenate([w["coords"]] * 5)  # 20 frames
    sysm = System.from_arrays({"name": w["names"], "resid": [1 + i // 3 for i in range(nw)],
                               "molecule_id": [1 + i // 3 for i in range(nw)],
                               "mass": [15.999, 1.008, 1.008] * (nw // 3)})
    traj = Trajectory(x, np.repeat(np.diag([144.3] * 3).reshape(1, 9), x.shape[0], 0).astype(np.float32))
    allat = np.arange(nw)
    plan = RdfPlan(allat, allat, bins=500, r_max=15.0, pbc="orthorhombic", bymol1=True, bymol2=True)
    plan.run(traj, sysm, device=dev)
    t0 = time.perf_counter()
    plan.run(traj, sysm, device=dev)
    bm_s = time.perf_counter() - t0
    out["rdf_bymol_host"] = {"molecules": nw // 3, "frames": int(x.shape[0]), "ms": bm_s * 1e3,
                             "frames_per_s": x.shape[0] / bm_s}
    log(f"structure factor f64 {n} atoms x {F} frames {sf_s * 1e3:.2f} ms "
        f"({out['structure_factor_f64']['pair_evals_per_s']:.3g} pair-evals/s); RdfPlan bymol "
        f"{nw // 3} molecules x {x.shape[0]} frames from host {bm_s * 1e3:.1f} ms")
    return out


def counter_sequence(ctx, dev):
    """One launch of each measured kernel at its bench size, for hardware
    counter runs (rocprofv3 --pmc): a plain 1 GiB device copy (the calibration
    of FETCH_SIZE / WRITE_SIZE), one RDF accumulate of 100 frames of the
    100,000-water box (500 bins, r_max 15), one RMSD run over 5,000 frames of
    the 20,000-atom chain (fit on 2,000 CA), one 5,000 x 5,000 pairwise RMSD
    matrix on the CA, one superposition of 5,000 x 20,000 atoms and one soft-RDF
    forward and backward (20,000 beads x 64 frames)."""
    import numpy as np
    import torch

    from warp_md_amd import workloads
    from warp_md_amd._native import check, lib, ptr

    L = lib()
    V = lambda t: ctypes.c_void_p(t.data_ptr())
    x = torch.ones(1 << 28, dtype=torch.float32, device=dev)  # 1 GiB
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    del x, y
    log("counters: 1 GiB copy done")
    F = 100
    store, n, stride, _ = workloads.water_lattice_device(47, 144.3, F, n_waters=100000, seed=1,
                                                         device=dev)
    boxes = np.repeat(np.diag([144.3] * 3).reshape(1, 9), F, 0).astype(np.float32)
    sel = np.arange(n, dtype=np.uint32)
    h = ctypes.c_void_p()
    check(L.wmd_rdf_create(ctx.handle, ptr(sel), n, ptr(sel), n, 500, ctypes.c_float(15.0), 1, 0,
                           ctypes.byref(h)))
    check(L.wmd_rdf_accumulate_device(h, V(store), stride, F, ptr(boxes), None))
    torch.cuda.synchronize()
    L.wmd_rdf_destroy(h)
    del store
    log("counters: RDF 100 frames done")
    c = workloads.chain(n_frames=0, seed=2)
    nc = c["ref"].shape[0]
    r4 = np.ones((nc, 4), np.float32)
    r4[:, :3] = c["ref"]
    allat = np.arange(nc, dtype=np.uint32)
    hr = ctypes.c_void_p()
    check(L.wmd_rmsd_create(ctx.handle, ptr(c["ca"]), c["ca"].size, ptr(allat), allat.size, ptr(r4),
                            nc, 1, ctypes.byref(hr)))
    ns, st = ctypes.c_size_t(), ctypes.c_size_t()
    check(L.wmd_rmsd_store_layout(hr, None, 0, ctypes.byref(ns), ctypes.byref(st)))
    atoms = np.zeros(st.value, np.uint32)
    check(L.wmd_rmsd_store_layout(hr, ptr(atoms), atoms.size, ctypes.byref(ns), ctypes.byref(st)))
    RF = 5000
    cs = workloads.chain_store_device(c["ref"], atoms, RF, seed=2, device=dev)
    o = torch.zeros(RF, dtype=torch.float32, device=dev)
    check(L.wmd_rmsd_run_device(hr, V(cs), st.value, RF, V(o), None, None, None, None, None))
    torch.cuda.synchronize()
    L.wmd_rmsd_destroy(hr)
    del cs
    log("counters: RMSD 5000 frames done")
    store = workloads.chain_store_device(c["ref"], c["ca"], 5000, seed=2, device=dev)
    selp = np.arange(c["ca"].size, dtype=np.uint32)
    hp = ctypes.c_void_p()
    check(L.wmd_pairwise_create(ctx.handle, ptr(selp), selp.size, 0, ctypes.byref(hp)))
    outp = torch.empty((5000, 5000), dtype=torch.float32, device=dev)
    check(L.wmd_pairwise_run_device(hp, V(store), store.shape[2], 5000, V(outp), None, None))
    torch.cuda.synchronize()
    L.wmd_pairwise_destroy(hp)
    del store, outp
    log("counters: pairwise 5000 x 5000 done")
    st2 = workloads.chain_store_device(c["ref"], allat, 5000, seed=2, device=dev)
    x4 = torch.ones((5000, nc, 4), dtype=torch.float32, device=dev)
    x4[:, :, :3] = st2.permute(0, 2, 1)
    del st2
    ha = ctypes.c_void_p()
    check(L.wmd_align_create(ctx.handle, ptr(c["ca"]), c["ca"].size, None, ptr(r4), nc, 0,
                             ctypes.byref(ha)))
    rt64 = torch.zeros((5000, 12), dtype=torch.float64, device=dev)
    xyz = torch.empty((5000, nc, 3), dtype=torch.float32, device=dev)
    check(L.wmd_align_run_device(ha, V(x4), nc, 5000, V(rt64), None, V(xyz), None))
    torch.cuda.synchronize()
    L.wmd_align_destroy(ha)
    log("counters: superposition 5000 x 20000 done")
    del x4, xyz
    # soft RDF forward and backward: 20,000 beads at 0.08 / A^3, 64 frames,
    # 300 bins, r_max 12, sigma 0.05 (the soft-RDF ratio's workload)
    from warp_md_amd import autograd

    n, frames = 20000, 64
    box = (n / 0.08) ** (1.0 / 3.0)
    gen = torch.Generator(device=dev)
    gen.manual_seed(4)
    base = torch.rand((n, 3), generator=gen, device=dev, dtype=torch.float32) * box
    xs = base[None] + 0.2 * torch.randn((frames, n, 3), generator=gen, device=dev, dtype=torch.float32)
    stride = (n + 3) // 4 * 4
    soa = torch.zeros((frames, 3, stride), dtype=torch.float32, device=dev)
    soa[:, :, :n] = xs.permute(0, 2, 1)
    boxes = np.repeat(np.diag([box] * 3).reshape(1, 9), frames, 0).astype(np.float32)
    plan = autograd.SoftRdf(bins=300, r_max=12.0, sigma=0.05, pbc="orthorhombic")
    c = plan.forward_store(soa, stride, n, boxes)
    plan.backward_store(soa, stride, n, boxes, torch.ones_like(c))
    torch.cuda.synchronize()
    log("counters: soft RDF forward + backward done")
    # the plain RDF forward on the same frames (the soft-RDF ratio's
    # denominator), then its pair evaluations under edge diagnostics
    sel = np.arange(n, dtype=np.uint32)
    for flags in (0, 1):
        h = ctypes.c_void_p()
        check(L.wmd_rdf_create(ctx.handle, ptr(sel), n, ptr(sel), n, 300, ctypes.c_float(12.0), 1,
                               flags, ctypes.byref(h)))
        check(L.wmd_rdf_accumulate_device(h, V(soa), stride, frames, ptr(boxes), None))
        torch.cuda.synchronize()
        if flags:
            ev, cnt = ctypes.c_uint64(), np.zeros(300, np.uint64)
            check(L.wmd_rdf_pair_evaluations(h, ctypes.byref(ev)))
            check(L.wmd_rdf_read(h, ptr(cnt), None))
            log(f"counters: soft workload plain RDF: {ev.value} evaluations, "
                f"{int(cnt.sum()) // 2} unordered in-range pairs")
        L.wmd_rdf_destroy(h)
    log("counters: plain RDF on the soft workload done")
    del soa, xs, base
    # the uneven batch (256 trajectories, 50 frames, default tuning), one
    # eager step: its pair kernel, cell build and RMSD kernels
    from warp_md_amd import workloads as wl
    from warp_md_amd.batch import UnevenBatch

    spec = wl.uneven_batch(sizes_only=True)
    ns = np.array([s["n"] for s in spec], np.int64)
    b = UnevenBatch(ns, [50] * len(spec), bins=300, r_max=12.0, graph=False, device=dev)
    b.set_boxes([s["box"] for s in spec])
    store = wl.uneven_store_device(b, spec, seed=3)
    rm = torch.zeros(50 * len(spec), dtype=torch.float32, device=dev)
    b.run(store, rm)
    torch.cuda.synchronize()
    log(f"counters: uneven batch step done ({float(b.counts().sum()) / 2:.4g} unordered in-range pairs)")


def uneven_batch_bench(ctx, dev, reps=3, autotune=True, cpu_sample=True):
    """256 trajectories, atom counts log-uniform in [500, 50000], cubic boxes at
    0.1 atoms/A^3, 50 uniform-random frames each; RDF (300 bins, r_max 12 A)
    and RMSD to frame 0 of every trajectory in one captured-graph step, against
    the same work as 256 single-trajectory plan calls."""
    import numpy as np
    import torch

    from warp_md_amd import workloads
    from warp_md_amd._native import check, lib, ptr
    from warp_md_amd.batch import UnevenBatch

    spec = workloads.uneven_batch(sizes_only=True)
    F = 50
    ns = np.array([s["n"] for s in spec], np.int64)
    b = UnevenBatch(ns, [F] * len(spec), bins=300, r_max=12.0, device=dev)
    b.set_boxes([s["box"] for s in spec])
    store = workloads.uneven_store_device(b, spec, seed=3)
    rm = torch.zeros(F * len(spec), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    tune_s = None
    if autotune:
        t0 = time.perf_counter()
        b.autotune(store, reps=2)
        torch.cuda.synchronize()
        tune_s = time.perf_counter() - t0
    kinds, units = b.plan()
    b.run(store, rm)  # capture + first replay
    torch.cuda.synchronize()
    b.reset()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.run(store, rm)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    counts = b.counts()
    pairs_in = float(counts.sum()) / 2.0 / reps  # unordered in-range pairs per step
    ref_evals = float(np.sum(F * ns * (ns - 1)))
    # the pair kernel's own duration from its per-slice wall-clock marks (first
    # wave start to last wave end), over further replays of the same graph
    marks = []
    for _ in range(reps):
        b.run(store, rm)
        marks.append(b.slice_times())
    marks = np.array(marks)
    pair_ms = float(np.mean(marks[:, :, 2].max(axis=1)))
    drained = marks[:, :, 1]
    ach = RDF_OPS_PER_PAIR * pairs_in / (pair_ms / 1e3) / 1e12
    k_t, bh_t, g_t = b.plan_tuning()
    tuning = {}
    for k, bh, g in zip(k_t.tolist(), bh_t.tolist(), g_t.tolist()):
        key = f"k={k}" + (f" home={bh}" if k == 0 else "") + f" grab={g}"
        tuning[key] = tuning.get(key, 0) + 1
    # the same work as one plan call per trajectory (RDF accumulate + RMSD run)
    V = lambda t: ctypes.c_void_p(t.data_ptr())
    handles = []
    for t, s in enumerate(spec):
        n = int(s["n"])
        sel = np.arange(n, dtype=np.uint32)
        h = ctypes.c_void_p()
        check(lib().wmd_rdf_create(ctx.handle, ptr(sel), n, ptr(sel), n, 300, ctypes.c_float(12.0), 1,
                                   0, ctypes.byref(h)))
        x0 = b.store_view(store, t)[0, :, :n].t().cpu().numpy()
        r4 = np.ones((n, 4), np.float32)
        r4[:, :3] = x0
        hr = ctypes.c_void_p()
        check(lib().wmd_rmsd_create(ctx.handle, ptr(sel), n, None, 0, ptr(r4), n, 1, ctypes.byref(hr)))
        bx = np.repeat(np.diag(s["box"]).reshape(1, 9), F, 0).astype(np.float32)
        handles.append((h, hr, bx))

    def single_pass():
        for t, (h, hr, bx) in enumerate(handles):
            v = b.store_view(store, t)
            check(lib().wmd_rdf_accumulate_device(h, V(v), v.shape[2], F, ptr(bx), None))
            check(lib().wmd_rmsd_run_device(hr, V(v), v.shape[2], F, V(rm[t * F:(t + 1) * F]), None,
                                            None, None, None, None))

    single_pass()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    single_pass()
    torch.cuda.synchronize()
    single_ms = (time.perf_counter() - t0) * 1e3
    for h, hr, _ in handles:
        lib().wmd_rdf_destroy(h)
        lib().wmd_rmsd_destroy(hr)
    log(f"uneven batch: {ms:.2f} ms per graph step ({units} pair units, "
        f"{int((kinds == 0).sum())} brute-force trajectories), pair kernel {pair_ms:.2f} ms "
        f"({ach / VALU_PEAK_TOPS:.3f} of VALU peak), 256 single-plan calls {single_ms:.2f} ms")
    # the oracle's reference loop on a bounded sample: frame 0 of the
    # trajectories at the 25th, 50th, 75th and 90th size percentiles
    cpu = None
    if cpu_sample:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        order = np.argsort(ns, kind="stable")
        picks = [int(order[int(q * (len(ns) - 1))]) for q in (0.25, 0.5, 0.75, 0.9)]
        evals, secs = 0.0, 0.0
        for t in picks:
            n = int(ns[t])
            x = b.store_view(store, t)[0, :, :n].t().contiguous().cpu().numpy()[None]
            t0 = time.perf_counter()
            oracle.rdf_counts(x, np.arange(n), np.arange(n), 300, 12.0, boxes=spec[t]["box"])
            secs += time.perf_counter() - t0
            evals += float(n) * (n - 1)
        cpu = {"value": evals / secs, "unit": "pair-evals/s", "cores": 1, "kind": "port",
               "sample": f"frame 0 of the trajectories of {', '.join(str(int(ns[t])) for t in picks)} atoms "
                         f"({evals:.3g} pair-evals, {secs:.1f} s, 1 thread, brute force in the reference loop order)"}
    return {"ms": ms, "pair_evals_per_s": ref_evals / (ms / 1e3), "cpu_baseline": cpu,
            "rmsd_frames_per_s": F * len(spec) / (ms / 1e3),
            "unordered_pairs_in_range": pairs_in,
            "single_plan_calls_ms": single_ms, "speedup_vs_single_calls": single_ms / ms,
            "autotune_s": tune_s, "cell_k_histogram": {str(k): int((kinds == k).sum()) for k in sorted(set(kinds.tolist()))},
            "tuning_histogram": tuning,
            "roofline": {"kernel": "batch_pair_kernel", "bound": "valu", "achieved": ach, "peak": VALU_PEAK_TOPS,
                         "unit": "TOP/s", "frac": ach / VALU_PEAK_TOPS,
                         "traffic": traffic_per_launch("batch_pair_kernel_256_uneven_50f", 1),
                         "traffic_measured_at": traffic_stamp(),
                         "algorithmic": f"{RDF_OPS_PER_PAIR} VALU lane-ops per unordered in-range pair x "
                                        f"{pairs_in:.4g} pairs per step",
                         "avg_launch_ms": pair_ms,
                         "timing": "wall-clock marks inside the kernel (first wave start to last wave end), "
                                   f"mean of {reps} graph replays"},
            "slices": {"drained_ms": [round(float(x), 3) for x in drained.mean(axis=0)],
                       "end_ms": [round(float(x), 3) for x in marks[:, :, 2].mean(axis=0)],
                       "tasks_helped": [int(x) for x in marks[:, :, 3].mean(axis=0)]},
            "workload": "256 trajectories, N log-uniform [500, 50000] (seed 3, sum N = 2,612,540), "
                        "density 0.1 /A^3, 50 frames, RDF 300 bins r_max 12 A + RMSD to frame 0"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=1000)
    ap.add_argument("--rmsd-frames", type=int, default=5000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-triclinic", action="store_true")
    ap.add_argument("--no-soft", action="store_true", help="skip the soft-RDF ratio")
    ap.add_argument("--no-batch", action="store_true", help="skip the uneven batch")
    ap.add_argument("--only-batch", action="store_true", help="time only the uneven batch")
    ap.add_argument("--no-pairwise", action="store_true", help="skip the pairwise RMSD matrix")
    ap.add_argument("--only-pairwise", action="store_true", help="time only the pairwise RMSD matrix")
    ap.add_argument("--no-align", action="store_true", help="skip the superposition leg")
    ap.add_argument("--no-ingest", action="store_true", help="skip the host-buffer (PCIe-inclusive) leg")
    ap.add_argument("--only-ingest", action="store_true", help="time only the host-buffer leg")
    ap.add_argument("--no-sf-bymol", action="store_true", help="skip the structure-factor and bymol legs")
    ap.add_argument("--only-sf-bymol", action="store_true", help="time only the structure-factor and bymol legs")
    ap.add_argument("--only-superpose", action="store_true", help="time only the pairwise and superposition legs")
    ap.add_argument("--counters", action="store_true",
                    help="launch each measured kernel once (for rocprofv3 --pmc runs) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for "
              f"--gpus {args.gpus}); frames shard across GPUs by contiguous ranges, which is not built here.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from warp_md_amd import _native, workloads
    from warp_md_amd._native import check, lib, ptr

    dev = "cuda:0"
    ctx = _native.context(0)
    L = lib()
    V = lambda t: ctypes.c_void_p(t.data_ptr())

    if args.counters:
        counter_sequence(ctx, dev)
        return
    if args.only_batch:
        print(json.dumps({"uneven_batch": uneven_batch_bench(ctx, dev, cpu_sample=not args.no_cpu)}), flush=True)
        return
    if args.only_sf_bymol:
        print(json.dumps({"sf_bymol": sf_bymol_bench(ctx, dev)}), flush=True)
        return
    if args.only_ingest:
        print(json.dumps({"host_ingest": host_ingest_bench(ctx)}), flush=True)
        return
    if args.only_pairwise or args.only_superpose:
        out = {"pairwise_rmsd": pairwise_bench(ctx, dev)}
        if args.only_superpose:
            out["superpose"] = align_bench(ctx, dev)
        print(json.dumps(out), flush=True)
        return

    # ---- inputs resident in HBM ------------------------------------------
    F = args.frames
    box_len = 144.3
    log(f"generating {F} frames of the 100,000-water box on the GPU")
    store, n, stride, _ = workloads.water_lattice_device(47, box_len, F, n_waters=100000, seed=1,
                                                         device=dev)
    boxes = np.repeat(np.diag([box_len] * 3).reshape(1, 9), F, 0).astype(np.float32)
    sel = np.arange(n, dtype=np.uint32)
    h_rdf = ctypes.c_void_p()
    check(L.wmd_rdf_create(ctx.handle, ptr(sel), n, ptr(sel), n, 500, ctypes.c_float(15.0), 1, 0,
                           ctypes.byref(h_rdf)))

    c = workloads.chain(n_frames=0, seed=2)
    n_chain = c["ref"].shape[0]
    r4 = np.ones((n_chain, 4), np.float32)
    r4[:, :3] = c["ref"]
    allat = np.arange(n_chain, dtype=np.uint32)
    h_rmsd = ctypes.c_void_p()
    check(L.wmd_rmsd_create(ctx.handle, ptr(c["ca"]), c["ca"].size, ptr(allat), allat.size, ptr(r4),
                            n_chain, 1, ctypes.byref(h_rmsd)))
    ns, st = ctypes.c_size_t(), ctypes.c_size_t()
    check(L.wmd_rmsd_store_layout(h_rmsd, None, 0, ctypes.byref(ns), ctypes.byref(st)))
    atoms = np.zeros(st.value, np.uint32)
    check(L.wmd_rmsd_store_layout(h_rmsd, ptr(atoms), atoms.size, ctypes.byref(ns), ctypes.byref(st)))
    RF = args.rmsd_frames
    log(f"generating {RF} chain frames on the GPU")
    chain_store = workloads.chain_store_device(c["ref"], atoms, RF, seed=2, device=dev)
    o32 = torch.zeros(RF, dtype=torch.float32, device=dev)
    o64 = torch.zeros(RF, dtype=torch.float64, device=dev)
    oerr = torch.zeros(RF, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    def step():
        check(L.wmd_rdf_reset(h_rdf))
        check(L.wmd_rdf_accumulate_device(h_rdf, V(store), stride, F, ptr(boxes), None))
        check(L.wmd_rmsd_run_device(h_rmsd, V(chain_store), st.value, RF, V(o32), V(o64), V(oerr),
                                    None, None, None))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    pair_ms, pair_n = ctx.profile_read("rdf_pair")
    build_ms, build_n = ctx.profile_read("cell_build")
    mom_ms, mom_n = ctx.profile_read("rmsd_moments")
    fin_ms, fin_n = ctx.profile_read("rmsd_finalize")
    ctx.profile(False)
    counts = np.zeros(500, np.uint64)
    check(L.wmd_rdf_read(h_rdf, ptr(counts), None))
    rmsd = o64.cpu().numpy()
    path = ctypes.c_int()
    ncs = (ctypes.c_int * 3)()
    check(L.wmd_rdf_last_path(h_rdf, ctypes.byref(path), ncs))
    K = args.steps
    ms_per_step = elapsed / K * 1e3
    pair_evals = F * (n * n - n)  # reference-equivalent ordered pair evaluations
    value = pair_evals / (ms_per_step / 1e3)
    rdf_ms = (pair_ms + build_ms) / K
    rmsd_ms = (mom_ms + fin_ms) / K
    p_in = float(counts.sum()) / 2.0  # unordered in-range pairs per step
    rdf_ach = RDF_OPS_PER_PAIR * p_in * K / (pair_ms / 1e3) / 1e12
    rmsd_bytes = 12.0 * n_chain * RF + 12.0 * n_chain + 4.0 * RF
    rmsd_ach = rmsd_bytes * K / (mom_ms / 1e3) / 1e9
    log(f"rdf path={path.value} cells={list(ncs)} pair {pair_ms / K:.2f} ms/step over "
        f"{pair_n // K} launches, build {build_ms / K:.2f} ms; rmsd moments {mom_ms / K:.3f} ms, "
        f"finalize {fin_ms / K:.3f} ms; step {ms_per_step:.2f} ms")

    # candidate efficiency of the cell enumeration (untimed): the pair
    # evaluations of 100 frames under edge diagnostics, against the in-range pairs
    h_diag = ctypes.c_void_p()
    check(L.wmd_rdf_create(ctx.handle, ptr(sel), n, ptr(sel), n, 500, ctypes.c_float(15.0), 1, 1,
                           ctypes.byref(h_diag)))
    FD = min(F, 100)
    check(L.wmd_rdf_accumulate_device(h_diag, V(store), stride, FD, ptr(boxes[:FD]), None))
    dcounts = np.zeros(500, np.uint64)
    n_edge = np.zeros(1, np.uint64)
    check(L.wmd_rdf_read(h_diag, ptr(dcounts), ptr(n_edge)))
    n_eval = ctypes.c_uint64()
    check(L.wmd_rdf_pair_evaluations(h_diag, ctypes.byref(n_eval)))
    L.wmd_rdf_destroy(h_diag)
    candidates = {"frames": FD, "pair_evaluations": int(n_eval.value),
                  "in_range_pairs": int(dcounts.sum()) // 2,
                  "efficiency": (int(dcounts.sum()) / 2.0) / max(1, int(n_eval.value)),
                  "edge_pairs": int(n_edge[0])}
    log(f"cell enumeration: {candidates['pair_evaluations']} evaluations for "
        f"{candidates['in_range_pairs']} in-range pairs over {FD} frames "
        f"(efficiency {candidates['efficiency']:.3f})")

    # triclinic variant of the RDF workload (10 deg xy tilt), timed separately
    tric = None
    if not args.no_triclinic:
        del store
        torch.cuda.empty_cache()
        log("triclinic variant")
        tstore, tn, tstride, M = workloads.water_lattice_device(47, box_len, F, n_waters=100000,
                                                                seed=1, tilt_deg=10.0, device=dev)
        tboxes = np.repeat(M.reshape(1, 9), F, 0).astype(np.float32)
        h_tri = ctypes.c_void_p()
        check(L.wmd_rdf_create(ctx.handle, ptr(sel), n, ptr(sel), n, 500, ctypes.c_float(15.0), 2, 0,
                               ctypes.byref(h_tri)))
        torch.cuda.synchronize()
        check(L.wmd_rdf_accumulate_device(h_tri, V(tstore), tstride, F, ptr(tboxes), None))
        torch.cuda.synchronize()
        check(L.wmd_rdf_reset(h_tri))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        check(L.wmd_rdf_accumulate_device(h_tri, V(tstore), tstride, F, ptr(tboxes), None))
        torch.cuda.synchronize()
        tdt = time.perf_counter() - t1
        tcounts = np.zeros(500, np.uint64)
        check(L.wmd_rdf_read(h_tri, ptr(tcounts), None))
        L.wmd_rdf_destroy(h_tri)
        tric = {"pair_evals_per_s": pair_evals / tdt, "ms": tdt * 1e3,
                "total_pairs": int(tcounts.sum())}
        del tstore
        torch.cuda.empty_cache()

    L.wmd_rdf_destroy(h_rdf)
    L.wmd_rmsd_destroy(h_rmsd)
    soft = None if args.no_soft else soft_rdf_ratio(ctx, dev)
    batch = None if args.no_batch else uneven_batch_bench(ctx, dev, cpu_sample=not args.no_cpu)
    pairwise = None if args.no_pairwise else pairwise_bench(ctx, dev)
    superpose = None if args.no_align else align_bench(ctx, dev)
    ingest = None if args.no_ingest else host_ingest_bench(ctx)
    sf_bymol = None if args.no_sf_bymol else sf_bymol_bench(ctx, dev)
    small = None if args.no_cpu else small_box_leg(ctx, dev)
    cpu = None if args.no_cpu else cpu_baseline()

    line = {
        "metric": "RDF pair-evals/s + RMSD frames/s on 1 MI355X; soft-RDF fwd+bwd vs plain",
        "value": value,
        "unit": "pair-evals/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "RdfPlan OW-OW, 100,000 waters, cubic 144.3 A, 1,000 frames, 500 bins, "
                        "r_max 15 A + batched RmsdPlan 5,000 frames x 20,000 atoms, fit on 2,000 CA, "
                        "RMSD over all atoms",
            "rdf_frames": F, "rdf_atoms": n, "bins": 500, "r_max": 15.0,
            "rmsd_frames": RF, "rmsd_atoms": n_chain, "rmsd_fit_atoms": int(c["ca"].size),
            "parallelism": "single GPU",
        },
        "rdf_pair_evals_per_s": pair_evals / (rdf_ms / 1e3),
        "rmsd_frames_per_s": RF / (rmsd_ms / 1e3),
        "rdf_triclinic": tric,
        "softrdf_fwd_bwd_over_plain": None if soft is None else soft["ratio"],
        "softrdf": soft,
        "uneven_batch": batch,
        "pairwise_rmsd": pairwise,
        "superpose": superpose,
        "host_ingest": ingest,
        "selection_gather_300k_atoms": None if ingest is None else {
            "ms_per_1000_frames": ingest["rdf"]["device_selection_gather"]["ms_per_1000_frames"],
            "note": "device gather of the 100,000 OW from 300,000-atom frames, not in the timed step "
                    "(the headline starts from the selected store, the reference's selected-chunk contract)"},
        "sf_bymol": sf_bymol,
        "rdf_1000_waters": small,
        "roofline": {
            "kernel": "rdf_cell_kernel",
            "bound": "valu",
            "achieved": rdf_ach,
            "peak": VALU_PEAK_TOPS,
            "unit": "TOP/s",
            "frac": rdf_ach / VALU_PEAK_TOPS,
            "traffic": traffic_per_launch("rdf_cell_kernel_100k_waters", F / max(pair_n // K, 1)),
            "traffic_measured_at": traffic_stamp(),
            "algorithmic": f"{RDF_OPS_PER_PAIR} VALU lane-ops per unordered in-range pair x "
                           f"{p_in:.4g} pairs per step",
            "avg_launch_ms": pair_ms / max(pair_n, 1),
            "candidates": candidates,
        },
        "roofline_rmsd": {
            "kernel": "rmsd_moments_kernel",
            "bound": "hbm",
            "achieved": rmsd_ach,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": rmsd_ach / HBM_PEAK_GBS,
            "traffic": traffic_per_launch("rmsd_moments_kernel_chain_20k", RF),
            "traffic_measured_at": traffic_stamp(),
            "avg_launch_ms": mom_ms / max(mom_n, 1),
        },
        "checks": {
            "rdf_total_pairs_per_step": int(counts.sum()),
            "rmsd_mean": float(np.mean(rmsd)),
            "rmsd_min": float(np.min(rmsd)),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
