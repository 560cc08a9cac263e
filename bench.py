"""WFBP throughput benchmark on one MI355X.

Workload (the configuration the headline metric is quoted on): GE-like
888 channels x 64 rows, 984 views/rotation, pitch 0.984, 10 rotations (9840
views), seeded ellipsoid phantom with Poisson noise at I0 = 5e4, reconstructed
into 512 x 512 x 320 slices at 0.625 mm, 500 mm FOV, ramp filter.

A step is one whole reconstruction (rebin + filter -> backprojection ->
thickening) with the raw scan already resident in HBM.  The chosen path
stores filtered projections as FP16 2 x 2 bilinear quads with one FP32 scale
per scan, accumulates in FP32 and runs as a Batch: the step recorded once
into a hipGraph, the tile depth from the autotuner (a batch of this one job
backprojects with the single-plan grid kernel; several jobs share one
persistent launch).
The plain FP32 single-plan path (and the FP16 single-plan path) are timed
beside it.  value is voxel-views/s of the chosen path, where a voxel-view is
a (native voxel, view) pair selected by the reconstruction rule (counted
exactly on the GPU).

--workload batch measures the 24 uneven jobs instead (aggregate
voxel-views/s of one mixed-precision Batch against the same jobs run one by
one on the plain FP32 path).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
XCDS = 8                   # MI355X accelerator complex dies (GRBM_GUI_ACTIVE sums over them)
BYTES_PER_VV = {"fp32": 16, "fp16": 8, "bf16": 8}   # 4 bilinear taps


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_path(plan, raw, steps, warmup, torch):
    """One plan's whole chain (fct_wfbp_gpu_reconstruct on device buffers:
    rebin+filter of the lines some slab reads, backprojection, thickening)
    per step; then the backprojection alone, timed with HIP events on the
    plan's stream over the filtered projections of a stage call."""
    vol = plan.alloc_volume()
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        plan.reconstruct(raw, vol)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.reconstruct(raw, vol)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    filt, scale = plan.alloc_filtered()
    native = plan.alloc_native()
    plan.rebin_filter(raw, filt, scale)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    for a, b in evs:
        a.record(stream)
        plan.backproject(filt, scale, native)
        b.record(stream)
    torch.cuda.synchronize()
    bp_ms = sum(a.elapsed_time(b) for a, b in evs) / steps
    del filt, scale, native
    return (t1 - t0) * 1e3 / steps, bp_ms, vol


def nominal_voxel_views(w, info):
    """The closed form in-FOV voxels x native slices x ppr / pitch (every
    in-FOV voxel seen for 1/pitch rotations); the exact count comes from the
    GPU selection rule."""
    import numpy as np
    rc, sc = w.recon, w.scan
    x = (np.arange(rc.nx) - 0.5 * (rc.nx - 1)) * rc.recon_fov / rc.nx + rc.x_origin
    y = (np.arange(rc.ny) - 0.5 * (rc.ny - 1)) * rc.recon_fov / rc.ny + rc.y_origin
    X, Y = np.meshgrid(x, y)
    in_fov = int(np.count_nonzero(X * X + Y * Y <= (0.5 * info.acquisition_fov) ** 2))
    pitch = abs(sc.feed) / (sc.n_rows * info.collimated_slice_width)
    return float(in_fov * info.n_slices_native * sc.views_per_rotation / pitch)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _lscpu():
    """CPU model, sockets and threads per core (lscpu), for the baseline record."""
    import subprocess
    info = {"model": _cpu_model()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("CPU(s)", "Thread(s) per core", "Core(s) per socket", "Socket(s)",
                             "Model name", "CPU max MHz"):
                info[k.strip()] = v.strip()
    except Exception:
        pass
    return info


def cpu_baseline(w, raw_np, n_vv_total, n_native, n_rf_views=960,
                 slice_groups=((0, 2), (None, 4), (-2, None)), repeats=3):
    """The CPU restatement on a bounded sample of the full job, FP64 (the
    oracle) and FP32 (the same templated backprojection loop in float):
    rebin + filter of n_rf_views central views (scaled to all views) and the
    backprojection of native slices at both ends and in the middle of the
    volume (scaled by the exact voxel-view count).  One warm-up, then the
    median of `repeats` runs of each stage; the spread of the per-voxel-view
    backprojection time between the three slice groups is the sampling error
    of the extrapolation.  OpenMP on every host core the process may use
    (OMP_PROC_BIND=close)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import statistics
    import numpy as np
    import pyoracle

    sc, rc = w.scan, w.recon
    d = pyoracle.geometry(sc, rc)
    cores = _cores()
    h = pyoracle.filter_kernel(d.nfft, d.ds, rc.filter_kind, rc.filter_c, rc.filter_a)
    va = sc.n_views // 2 - n_rf_views // 2

    def rf_sample():
        par = pyoracle.rebin(sc, raw_np, va, va + n_rf_views)
        return pyoracle.filter_rows(par, h, d.nfft, use_fft=True)

    groups = []
    for a, n in slice_groups:
        if a is None:
            s0 = n_native // 2 - n // 2
            s1 = s0 + n
        elif n is None:
            s0, s1 = n_native + a, n_native
        else:
            s0, s1 = a, a + n
        v0, v1 = pyoracle.view_range(sc, rc, s0, s1)
        par = pyoracle.filter_rows(pyoracle.rebin(sc, raw_np, v0, v1), h, d.nfft, use_fft=True)
        groups.append((s0, s1, v0, par))

    def timed(fn):
        fn()   # warm-up
        ts = []
        for _ in range(repeats):
            t = time.perf_counter()
            out = fn()
            ts.append(time.perf_counter() - t)
        return statistics.median(ts), ts, out

    t_rf, t_rf_all, _ = timed(rf_sample)
    res = {}
    refs = []
    for prec in ("fp64", "fp32"):
        per_vv, t_groups, n_groups = [], [], []
        for gi, (s0, s1, v0, par) in enumerate(groups):
            if prec == "fp64":
                fn = lambda: pyoracle.backproject(sc, rc, par, v0, s0, s1)   # noqa: E731
            else:
                p32 = par.astype(np.float32)
                fn = lambda: pyoracle.backproject_f32(sc, rc, p32, v0, s0, s1)   # noqa: E731
            tm, _, (vol, n_sel) = timed(fn)
            per_vv.append(tm / max(n_sel, 1))
            t_groups.append(tm)
            n_groups.append(n_sel)
            if prec == "fp64":
                refs.append((s0, vol))
        n_s = sum(n_groups)
        t_bp_vv = sum(t_groups) / max(n_s, 1)
        t_bp = t_bp_vv * n_vv_total
        t_rf_total = t_rf * sc.n_views / float(n_rf_views)
        total = t_rf_total + t_bp
        res[prec] = {"value": n_vv_total / total, "s_per_volume_est": total,
                     "stage_s": {"rebin_filter": t_rf_total, "backproject": t_bp},
                     "bp_s_per_voxel_view_by_group": per_vv,
                     "extrapolation_spread": (max(per_vv) - min(per_vv)) / t_bp_vv}
    groups_txt = ", ".join(f"{s0}..{s1 - 1}" for s0, s1, _, _ in groups)
    out = dict(res["fp64"])
    out.update({
        "unit": "voxel-views/s", "cores": cores, "kind": "port",
        "sample": (f"FP64 OpenMP restatement on {cores} threads ({_cpu_model()}), warm-up + "
                   f"median of {repeats}: rebin+filter of {n_rf_views} of {sc.n_views} views "
                   f"({t_rf:.2f} s) scaled to all views, plus backprojection of native slices "
                   f"{groups_txt} scaled by the exact voxel-view count"),
        "fp32": res["fp32"],
        "rebin_filter_runs_s": t_rf_all,
        "omp": {"OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"),
                "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND")},
        "lscpu": _lscpu(),
    })
    return out, refs


def batch_path(ctx, jobs, raws, precision, steps, warmup, torch, candidates=(4, 8, 16)):
    """Chosen path: one Batch (per-job rebin+filter, one backprojection
    launch over every job's tiles -- persistent for several jobs, the grid
    kernel for one -- per-job thicken) recorded into a hipGraph and replayed
    once per step, tile depth from the autotuner.  Returns ms/step, the
    backprojection's mean ms (HIP events around BP-only launches on the batch
    stream), the tuner's table and the volumes."""
    from freect_amd import wfbp
    stream = torch.cuda.Stream()
    b = wfbp.Batch(ctx, jobs, precision=precision, stream=stream)
    vols = [p.alloc_volume() for p in b.plans]
    for j, (r, v) in enumerate(zip(raws, vols)):
        b.submit(j, r, v)
    torch.cuda.synchronize()
    tuned = b.autotune(candidates, reps=2)
    log("autotune", precision, tuned, "->", b.info()["tile_z"])
    for _ in range(max(warmup, 1)):
        b.run(wfbp.Batch.GRAPH)
    b.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev[0].record(stream)
    for _ in range(steps):
        b.run(wfbp.Batch.GRAPH)
    ev[1].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = ev[0].elapsed_time(ev[1]) / steps
    wall_ms = (t1 - t0) * 1e3 / steps
    # The backprojection alone, same stream, same buffers.
    bev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    for a, c in bev:
        a.record(stream)
        b.backproject()
        c.record(stream)
    torch.cuda.synchronize()
    bp_ms = sum(a.elapsed_time(c) for a, c in bev) / steps
    # Restore the volumes of a full run (the BP-only launches rewrote native).
    b.run(wfbp.Batch.GRAPH)
    b.synchronize()
    return max(ms, wall_ms), bp_ms, tuned, b, vols


def _affinity():
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


# Host cores this process may use, read before the OpenMP runtime of the CPU
# restatement binds the main thread (after which the affinity shows one core).
_START_CORES = _affinity()


def _cores():
    """Threads the CPU restatement runs on: OMP_NUM_THREADS if set, else the
    cores the process could use at start."""
    try:
        n = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        n = 0
    return min(n, _START_CORES) if n > 0 else _START_CORES


def bench_batch(args, torch):
    """24 uneven jobs on the 736x16 detector: sequential plain FP32 single-plan
    reconstructions vs one mixed-precision Batch (graph + persistent BP)."""
    import numpy as np
    from freect_amd import scans as S
    from freect_amd import wfbp
    ctx = wfbp.Context(0)
    ws = S.batch_jobs()
    raws = []
    for w in ws:
        raws.append(ctx.phantom_project(w.scan, w.phantom, w.i0, w.seed))
    torch.cuda.synchronize()
    plans = [wfbp.Plan(ctx, w.scan, w.recon, precision="fp32") for w in ws]
    n_vv = sum(p.count_voxel_views() for p in plans)
    log("batch: jobs", len(ws), "voxel-views", n_vv)
    vols = [p.alloc_volume() for p in plans]

    def seq():
        for p, r, v in zip(plans, raws, vols):
            p.reconstruct(r, v)

    for _ in range(max(args.warmup, 1)):
        seq()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        seq()
    torch.cuda.synchronize()
    seq_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    log("sequential fp32 ms", seq_ms)
    jobs = [(w.scan, w.recon) for w in ws]
    ms, bp_ms, tuned, b, bvols = batch_path(ctx, jobs, raws, args.precision, args.steps,
                                            args.warmup, torch)
    log("batched", args.precision, "ms", ms, "bp ms", bp_ms)
    # Per-job HU difference between the two paths (FP16 storage vs FP32).
    hu = []
    for v0, v1 in zip(vols, bvols):
        d = (v1.double() - v0.double()) / (S.MU_WATER / 1000.0)
        hu.append(float(d.pow(2).mean().sqrt()))
    bpv = BYTES_PER_VV[args.precision]
    out = {
        "metric": "aggregate voxel-views/s, 24 uneven jobs, batched mixed vs sequential plain FP32",
        "value": n_vv / (ms * 1e-3), "unit": "voxel-views/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "storage": args.precision,
        "data": "synthetic (24 seeded jobs, Poisson noise, generated on the GPU)",
        "config": {"workload": "batch: 24 jobs on 736x16 (1152 views/rot), 2-20 rotations, "
                               "pitch {0.516,0.984,1.375}, 256^2/512^2, 16-240 slices cut to "
                               "the fully covered z-range",
                   "voxel_views": n_vv, "tile_z": b.info()["tile_z"], "autotune_ms": tuned},
        "sequential_plain_fp32": {"ms_per_step": seq_ms,
                                  "voxel_views_per_s": n_vv / (seq_ms * 1e-3)},
        "speedup_vs_sequential_plain_fp32": seq_ms / ms,
        "gather_rate": {"kernel": "bp_batch_kernel", "GBps": n_vv * bpv / (bp_ms * 1e-3) / 1e9,
                        "bytes_per_voxel_view": bpv, "bp_ms": bp_ms,
                        "note": "algorithmic tap bytes (served from L1/L2) per backprojection "
                                "time; not an HBM roofline (the full workload's line carries "
                                "the counter-measured HBM fraction)"},
        "hu_rmse_batched_vs_sequential_per_job_max": max(hu),
    }
    print(json.dumps(out), flush=True)


def bench_adjoint(args, torch):
    """Differentiable WFBP on the adjoint workload (the 736 x 16 detector at
    984 views/rot, 3 rotations = 2952 views, 512^2 x 64 slices, noiseless):
    WFBP volume x0, then 10 gradient steps on 0.5 ||m (k P x - p)||^2 with
    step 1 / (k^2 lambda_max) from power iteration, keeping the sinogram-sized
    residual ("store") or recomputing it in the backward pass ("recompute");
    P the ray-driven line-integral projector (k = 1, the default), the
    matched voxel-driven projector R^T M^T (k fitted) or the exact transpose
    A^T of the whole WFBP chain.  HU error of every iterate against the
    noiseless phantom, sampled at the voxel centres and averaged over each
    voxel (3 x 3 x 3 points)."""
    import numpy as np
    from freect_amd import autograd as AG
    from freect_amd import scans as S
    from freect_amd import wfbp
    w = S.adjoint()
    sc, rc = w.scan, w.recon
    ctx = wfbp.Context(0)
    raw = ctx.phantom_project(sc, w.phantom, 0.0, w.seed)
    plan = wfbp.Plan(ctx, sc, rc, precision="fp32")
    x0 = AG.wfbp(raw, plan).detach()
    torch.cuda.synchronize()
    info = plan.info
    n = info.n_slices_native
    # Truth on the native grid: central slices (fully covered by the helix),
    # voxels inside 0.9 of the acquisition FOV.
    csw = info.collimated_slice_width
    zs = [rc.start_pos + csw * i for i in range(n)]
    sl = list(range(n // 4, 3 * n // 4, 4))
    dx = rc.recon_fov / rc.nx
    xy = (np.arange(rc.nx) - 0.5 * (rc.nx - 1)) * dx
    X, Y = np.meshgrid(xy, xy)
    fov = X * X + Y * Y <= (0.45 * info.acquisition_fov) ** 2
    off = (np.arange(3) + 0.5) / 3 - 0.5
    truth = {i: w.phantom.mu_at(X, Y, zs[i]) for i in sl}
    truth_vox = {i: sum(w.phantom.mu_at(X + a * dx, Y + b * dx, zs[i] + c * csw)
                        for a in off for b in off for c in off) / 27.0 for i in sl}
    HU = S.MU_WATER / 1000.0

    def hu_rmse(x, t=truth):
        xs = x.detach().cpu().numpy()
        d = np.concatenate([((xs[i] - t[i]) / HU)[fov] for i in sl])
        return float(np.sqrt(np.mean(d * d)))

    out_modes = {}
    mask = AG.inside_rays(plan)
    log("data term on", int(mask.sum()), "of", mask.numel(), "(view, row) rays")
    for op, modes in (("ray", ("store", "recompute")), ("matched", ("store",)),
                      ("adjoint", ("store",))):
        lam = AG.power_iteration(plan, x0.shape, operator=op)
        for mode in modes:
            hist, hist_v = [], []

            def cb(it, xx):
                hist.append(hu_rmse(xx))
                hist_v.append(hu_rmse(xx, truth_vox))
            x, st = AG.refine(plan, raw, x0, steps=10, mode=mode, lam=lam, operator=op,
                              mask=mask if op != "adjoint" else None, callback=cb)
            key = f"{op}/{mode}"
            out_modes[key] = {"hu_rmse_per_step": hist, "hu_rmse_voxel_avg_per_step": hist_v,
                              "loss": st["loss"], "scale": st["scale"], "lambda_max": lam,
                              "ms_per_step": float(np.median(st["ms_per_step"])),
                              "peak_mb": float(max(st["peak_mb"]))}
            log(key, out_modes[key])
    out = {
        "metric": "differentiable WFBP: 10 data-consistency steps, HU RMSE vs noiseless phantom",
        "value": out_modes["ray/store"]["hu_rmse_per_step"][-1], "unit": "HU RMSE",
        "n_gpus": 1, "steps": 10, "warmup": 0,
        "ms_per_step": out_modes["ray/store"]["ms_per_step"], "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantom, noiseless, generated on the GPU)",
        "config": {"workload": "adjoint: 736x16 at 984 views/rot, 2952 views (3 rotations), "
                               "512^2 x 64 slices",
                   "objective": "0.5 ||m (k P x - p)||^2; ray: P = Joseph line-integral "
                                "projector (k = 1); matched: P = R^T M^T (voxel-driven, k least "
                                "squares at x0); adjoint: P = A^T (k = 1, m = 1); m the rays "
                                "inside the volume's z-range",
                   "hu_mask": "native slices n/4..3n/4 step 4, r < 0.45 FOV"},
        "hu_rmse_wfbp": hu_rmse(x0), "hu_rmse_wfbp_voxel_avg": hu_rmse(x0, truth_vox),
        "modes": out_modes,
    }
    print(json.dumps(out), flush=True)


def bench_ffs(args, torch):
    """Flying-focal-spot scan (Siemens-like, phi + z FFS, 2304 projections per
    rotation interleaved to 1472 x 64): plain FP32 and mixed precision, HU
    error of both against the FP64 restatement on the central native slice
    (computed in the CPU-baseline leg, which also times it)."""
    import numpy as np
    from freect_amd import scans as S
    from freect_amd import wfbp
    w = S.ffs()
    sc, rc = w.scan, w.recon
    ctx = wfbp.Context(0)
    raw = ctx.phantom_project(sc, w.phantom, w.i0, w.seed)
    torch.cuda.synchronize()
    res, vols = {}, {}
    n_vv = None
    for prec in ("fp32", args.precision):
        plan = wfbp.Plan(ctx, sc, rc, precision=prec)
        if n_vv is None:
            n_vv = plan.count_voxel_views()
            info = plan.info
        ms, bp_ms, vol = time_path(plan, raw, args.steps, args.warmup, torch)
        res[prec] = {"ms_per_step": ms, "bp_ms": bp_ms, "voxel_views_per_s": n_vv / (ms * 1e-3)}
        vols[prec] = (plan, vol)
        log("ffs", prec, res[prec])
    cpu, hu = None, {}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        s0 = info.n_slices_native // 2
        t = time.perf_counter()
        ref, n_sel = pyoracle.reconstruct_native(sc, rc, raw.cpu().numpy(), s0, s0 + 1)
        t_cpu = time.perf_counter() - t
        cpu = {"value": n_sel / t_cpu, "unit": "voxel-views/s", "cores": _cores(),
               "kind": "port", "sample": f"FP64 restatement, rebin+filter+BP of native slice {s0}"}
        xy = (np.arange(rc.nx) - 0.5 * (rc.nx - 1)) * rc.recon_fov / rc.nx
        X, Y = np.meshgrid(xy, xy)
        mask = X * X + Y * Y <= (0.5 * info.acquisition_fov) ** 2
        for prec, (plan, _) in vols.items():
            nat = plan.alloc_native()
            filt, scale = plan.alloc_filtered()
            plan.rebin_filter(raw, filt, scale)
            plan.backproject(filt, scale, nat)
            plan.synchronize()
            d = (nat[s0].cpu().numpy().astype(np.float64) - ref[0]) / (S.MU_WATER / 1000.0)
            hu[prec] = {"rmse": float(np.sqrt(np.mean(d[mask] ** 2))),
                        "max": float(np.abs(d[mask]).max()), "native_slice": int(s0)}
        log("ffs hu", hu)
    chosen = res[args.precision]
    bpv = BYTES_PER_VV[args.precision]
    out = {
        "metric": "FFS scan: voxel-views/s, mixed vs plain FP32, HU vs FP64",
        "value": chosen["voxel_views_per_s"], "unit": "voxel-views/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": chosen["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "storage": args.precision,
        "data": "synthetic (seeded ellipsoid phantom through phi+z FFS rays, Poisson I0=2.5e4)",
        "config": {"workload": "ffs: Siemens-like 736x32, 2304 proj/rot phi+z FFS -> 1472x64, "
                               "pitch 0.8, 8 rotations, 512^2 x 200 slices of 1.0 mm at 1.0 mm "
                               "over [-100, 100) mm",
                   "voxel_views": n_vv, "n_parallel_views": info.n_parallel_views},
        "plain_fp32": res["fp32"], "speedup_vs_plain_fp32":
            res["fp32"]["ms_per_step"] / chosen["ms_per_step"],
        "gather_rate": {"kernel": "bp_v1_kernel",
                        "GBps": n_vv * bpv / (chosen["bp_ms"] * 1e-3) / 1e9,
                        "bytes_per_voxel_view": bpv, "bp_ms": chosen["bp_ms"],
                        "note": "algorithmic tap bytes (served from L1/L2) per backprojection "
                                "time; not an HBM roofline"},
        "cpu_baseline": cpu, "hu_error_vs_fp64": hu,
    }
    print(json.dumps(out), flush=True)


def bench_e2e(args, torch):
    """The full workload end to end from host memory only (the host path of
    fct_wfbp_gpu_reconstruct: pinned scan in, pinned volume out, streamed),
    FP32 and the chosen precision, for tracing the streaming schedule."""
    from freect_amd import scans as S
    from freect_amd import wfbp
    w = S.full()
    ctx = wfbp.Context(0)
    raw = ctx.phantom_project(w.scan, w.phantom, w.i0, w.seed)
    torch.cuda.synchronize()
    raw_host = raw.cpu().pin_memory()
    del raw
    res = {}
    for prec in sorted({"fp32", args.precision}):
        plan = wfbp.Plan(ctx, w.scan, w.recon, precision=prec)
        info = plan.info
        vol = torch.empty((info.n_slices_requested, w.recon.ny, w.recon.nx),
                          dtype=torch.float32).pin_memory()
        for _ in range(max(1, args.warmup)):
            plan.reconstruct_host(raw_host, out=vol.numpy())
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.reconstruct_host(raw_host, out=vol.numpy())
        res[prec] = (time.perf_counter() - t0) / args.steps
        log("end to end", prec, res[prec], "s/volume")
        del plan, vol
    out = {
        "metric": "s/volume end to end from pinned host memory", "value": res[args.precision],
        "unit": "s/volume", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "higher_is_better": False, "vs_baseline": None, "storage": args.precision,
        "data": "synthetic (seeded ellipsoid phantom, Poisson noise I0=5e4, generated on the GPU)",
        "config": {"workload": "full, streamed: views [%d, %d) of %d uploaded"
                               % (info.read_views_first, info.read_views_end, w.scan.n_views)},
        "s_per_volume_e2e": res,
    }
    print(json.dumps(out), flush=True)


def bench_valu():
    """VALU issue rates of this GPU from the probe built beside the library
    (freect_amd/csrc/probe/valu_probe.hip): independent chains of one wave64
    instruction, 8 waves per SIMD.  The backprojection's VALU issue
    (valu_issue_per_simd_cycle in the full workload's roofline) is read
    against these: a wave64 FP32 instruction holds a 16-lane SIMD for 4
    cycles (0.25 per SIMD-cycle)."""
    import subprocess
    from freect_amd import build as fbuild
    if not os.path.exists(fbuild.PROBE_PATH):
        raise SystemExit(f"{fbuild.PROBE_PATH} is missing: run __graft_entry__.build() first")
    r = subprocess.run([fbuild.PROBE_PATH], capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        raise SystemExit(f"valu probe failed ({r.returncode}): {r.stderr.strip()}")
    probe = json.loads(r.stdout)
    fma = probe["rates"]["v_fma_f32"]["per_simd_cycle"]
    print(json.dumps({"metric": "VALU wave64 instructions per SIMD-cycle (v_fma_f32 chains)",
                      "value": fma, "unit": "instructions/SIMD-cycle", "n_gpus": 1,
                      "higher_is_better": True, "data": "synthetic", "probe": probe}))


def bench_toeplitz(args, torch):
    """The ramp filter of the full workload as one dense Toeplitz product on
    the matrix cores (hipBLASLt through torch.matmul; contiguous operands) in
    FP32 / FP16 / BF16, against the register-FFT rebin+filter kernel: time,
    and the relative error against an FP64 direct convolution of 4096 lines.
    Lines: the 64 rows of the views the slices read; taps: Ram-Lak at the
    parallel channel spacing R dgamma."""
    import numpy as np
    from freect_amd import scans as S
    from freect_amd import wfbp
    w = S.full()
    sc = w.scan
    n = sc.n_channels
    ds = sc.r * sc.dgamma
    k = np.arange(-(n - 1), n)
    h = np.zeros(2 * n - 1)
    h[k == 0] = 1.0 / (4.0 * ds * ds)
    odd = (k % 2) != 0
    h[odd] = -1.0 / (np.pi * k[odd] * ds) ** 2
    T64 = h[np.arange(n)[None, :] - np.arange(n)[:, None] + (n - 1)]   # T[i][c] = h[c - i]
    ctx = wfbp.Context(0)
    raw = ctx.phantom_project(sc, w.phantom, w.i0, w.seed)
    torch.cuda.synchronize()
    p32 = wfbp.Plan(ctx, sc, w.recon, precision="fp32")
    lines = (p32.info.read_views_end - p32.info.read_views_first) * sc.n_rows
    X = raw.reshape(-1, n)[:lines].contiguous()
    sample = slice(0, 4096)
    ref = X[sample].double().cpu().numpy() @ T64
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    reps = max(args.steps, 1)

    def timed(fn):
        for _ in range(max(args.warmup, 1)):
            fn()
        torch.cuda.synchronize()
        ev[0].record()
        for _ in range(reps):
            out = fn()
        ev[1].record()
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]) / reps, out

    gemm = {}
    for name, dt in (("fp32", torch.float32), ("fp16", torch.float16), ("bf16", torch.bfloat16)):
        Tt = torch.from_numpy(T64).to("cuda", dt).contiguous()
        Xt = X.to(dt).contiguous()
        ms, Y = timed(lambda: torch.matmul(Xt, Tt))
        got = Y[sample].double().cpu().numpy()
        gemm[name] = {"ms": ms, "tflops": 2.0 * lines * n * n / (ms * 1e-3) / 1e12,
                      "rel_rms_err": float(np.sqrt(np.mean((got - ref) ** 2) / np.mean(ref ** 2))),
                      "rel_max_err": float(np.abs(got - ref).max() / np.abs(ref).max())}
        log("toeplitz", name, gemm[name])
        del Tt, Xt, Y
    fft = {}
    for prec in ("fp32", "fp16"):
        p = p32 if prec == "fp32" else wfbp.Plan(ctx, sc, w.recon, precision=prec)
        filt, scale = p.alloc_filtered()
        ms, _ = timed(lambda: p.rebin_filter(raw, filt, scale))
        fft[prec] = {"ms": ms, "lines": sc.n_views * sc.n_rows}
        log("rebin+filter", prec, fft[prec])
        del filt, scale
    out = {
        "metric": "ramp filter of the full workload: dense Toeplitz GEMM vs register-FFT kernel",
        "value": gemm["fp32"]["ms"], "unit": "ms (FP32 GEMM)", "n_gpus": 1, "steps": reps,
        "warmup": args.warmup, "higher_is_better": False, "vs_baseline": None,
        "data": "synthetic (the full workload's noisy scan rows as filter input)",
        "config": {"workload": f"toeplitz: {lines} lines x {n} channels, {n} x {n} Toeplitz "
                               f"Ram-Lak at ds = {ds:.4f} mm"},
        "gemm": gemm,
        "rebin_filter_kernel": fft,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="fp16", choices=["fp16", "bf16", "fp32"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--tile-z", type=int, default=0)
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a kernel of known "
                         "traffic, to calibrate rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--workload", default="full",
                    choices=["full", "batch", "adjoint", "ffs", "toeplitz", "e2e", "valu"],
                    help="full: the headline scan (default); batch: 24 uneven jobs; adjoint: "
                         "the refinement report; ffs: the phi+z FFS scan; toeplitz: the ramp "
                         "filter as a dense GEMM against the FFT kernel; e2e: the full "
                         "workload from host memory only; valu: the VALU issue rates of this "
                         "GPU (the backprojection's roofline)")
    ap.add_argument("--no-counters", action="store_true",
                    help="skip the rocprofv3 counter passes (HBM traffic, VALU activity)")
    ap.add_argument("--probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--path", default="batch", choices=["batch", "plan"],
                    help="chosen path: hipGraph-replayed Batch with the persistent "
                         "backprojection (default) or the single-plan stage calls")
    args = ap.parse_args()
    # The CPU baseline's OpenMP threads stay on neighbouring cores (set before
    # the restatement's OpenMP runtime starts).
    os.environ.setdefault("OMP_PROC_BIND", "close")
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus > 1 is not supported "
                                   "(multi-GPU z-slab sharding is design-only)"}))
        sys.exit(2)

    if args.workload == "valu":
        bench_valu()
        return

    import numpy as np
    import torch

    from freect_amd import scans as S
    from freect_amd import wfbp

    torch.cuda.set_device(0)
    if args.probe:
        traffic_probe(args, torch)
        return
    if args.copy_mib:
        src = torch.ones(args.copy_mib * 2 ** 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.copy_mib} MiB read + {args.copy_mib} MiB written")
        del src, dst
    if args.workload == "batch":
        bench_batch(args, torch)
        return
    if args.workload == "adjoint":
        bench_adjoint(args, torch)
        return
    if args.workload == "ffs":
        bench_ffs(args, torch)
        return
    if args.workload == "toeplitz":
        bench_toeplitz(args, torch)
        return
    if args.workload == "e2e":
        bench_e2e(args, torch)
        return
    w = S.full()
    ctx = wfbp.Context(0)
    log("generating synthetic scan", w.scan.n_views, w.scan.n_rows, w.scan.n_channels)
    raw = ctx.phantom_project(w.scan, w.phantom, w.i0, w.seed)
    torch.cuda.synchronize()

    plans = {}
    for prec in sorted({"fp32", args.precision}):
        plans[prec] = wfbp.Plan(ctx, w.scan, w.recon, precision=prec, tile_z=args.tile_z)
    p0 = plans[args.precision]
    n_vv = p0.count_voxel_views()
    log("voxel-views", n_vv)

    res = {}
    vols = {}
    for prec, plan in plans.items():
        log("timing", prec)
        ms, bp_ms, vol = time_path(plan, raw, args.steps, args.warmup, torch)
        res[prec] = {"ms_per_step": ms, "bp_ms": bp_ms,
                     "voxel_views_per_s": n_vv / (ms * 1e-3),
                     "bp_voxel_views_per_s": n_vv / (bp_ms * 1e-3)}
        vols[prec] = vol
        log(prec, res[prec])

    kernel = "bp_v1_kernel"
    tuned = None
    if args.path == "batch":
        log("timing batched", args.precision)
        ms, bp_ms, tuned, b, bvols = batch_path(ctx, [(w.scan, w.recon)], [raw], args.precision,
                                                args.steps, args.warmup, torch)
        res["batch"] = {"ms_per_step": ms, "bp_ms": bp_ms,
                        "voxel_views_per_s": n_vv / (ms * 1e-3),
                        "bp_voxel_views_per_s": n_vv / (bp_ms * 1e-3),
                        "tile_z": b.info()["tile_z"], "autotune_ms": tuned}
        vols["batch"] = bvols[0]
        log("batch", res["batch"])
        chosen = res["batch"]
        kernel = "bp_v1_kernel"   # a batch of one job launches the grid kernel
    else:
        chosen = res[args.precision]
    info = p0.info

    # End to end from host memory: the scan in pinned host memory, streamed
    # to the device in chunks of views overlapped with rebin+filter and the
    # backprojection of the slabs whose views have arrived, the volume copied
    # back (fct_wfbp_gpu_reconstruct with host buffers).
    e2e = {}
    raw_host = raw.cpu().pin_memory()
    vol_host = torch.empty((info.n_slices_requested, w.recon.ny, w.recon.nx),
                           dtype=torch.float32).pin_memory()
    for prec in sorted(plans):
        plan = plans[prec]
        for _ in range(max(1, args.warmup)):
            plan.reconstruct_host(raw_host, out=vol_host.numpy())
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.reconstruct_host(raw_host, out=vol_host.numpy())
        e2e[prec] = (time.perf_counter() - t0) / args.steps
        log("end to end", prec, e2e[prec], "s/volume")
    # PCIe rates of the same buffers (one copy each way, pinned memory).
    dev = torch.empty_like(raw)
    vdev = torch.empty(vol_host.shape, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev.copy_(raw_host)
    torch.cuda.synchronize()
    t_h2d = time.perf_counter() - t0
    t0 = time.perf_counter()
    vol_host.copy_(vdev)
    torch.cuda.synchronize()
    t_d2h = time.perf_counter() - t0
    pcie = {"h2d_GBps": raw.numel() * 4 / t_h2d / 1e9, "d2h_GBps": vdev.numel() * 4 / t_d2h / 1e9,
            "h2d_s": t_h2d, "d2h_s": t_d2h}
    log("pcie", pcie)
    del raw_host, vol_host, dev, vdev

    cpu = None
    hu = {}
    if not args.no_cpu:
        log("cpu baseline")
        raw_np = raw.cpu().numpy()
        cpu, refs = cpu_baseline(w, raw_np, n_vv, info.n_slices_native)
        log("cpu", cpu)
        # HU error of the GPU paths against the FP64 restatement on those
        # slices (native == requested grid for this workload).
        x = (np.arange(w.recon.nx) - 0.5 * (w.recon.nx - 1)) * w.recon.recon_fov / w.recon.nx
        X, Y = np.meshgrid(x, x)
        mask = X * X + Y * Y <= (0.5 * info.acquisition_fov) ** 2
        for prec, vol in vols.items():
            ds = []
            for s0, ref in refs:
                n = ref.shape[0]
                d = (vol[s0:s0 + n].cpu().numpy().astype(np.float64) - ref) / (S.MU_WATER / 1000.0)
                ds.append(d[:, mask].ravel())
            d = np.concatenate(ds)
            hu[prec] = {"rmse": float(np.sqrt(np.mean(d * d))), "max": float(np.abs(d).max()),
                        "slices": [[int(s0), int(s0 + r.shape[0] - 1)] for s0, r in refs]}
        log("hu", hu)

    bpv = BYTES_PER_VV[args.precision]
    gather = n_vv * bpv / (chosen["bp_ms"] * 1e-3)
    counters = None
    if not args.no_counters:
        del plans, p0, vols
        torch.cuda.empty_cache()
        counters = measure_counters(args, chosen.get("tile_z", info.tile_z))
    key = "bp_" + args.precision
    roof = {"bound": "hbm", "kernel": f"bp_v1_kernel<{chosen.get('tile_z', info.tile_z)}, "
                                      f"{args.precision.upper()}>",
            "achieved": None, "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": None,
            "traffic": None, "bp_ms": chosen["bp_ms"],
            "gather_rate": {"GBps": gather / 1e9, "bytes_per_voxel_view": bpv,
                            "note": "4 bilinear taps per voxel-view read from L1/L2 (the "
                                    "algorithmic gather rate, not HBM traffic)"},
            "lds_bytes": 0}
    if counters and key in counters:
        c = counters[key]
        hbm = c["fetch_bytes"] + c["write_bytes"]
        roof.update({"achieved": hbm / (chosen["bp_ms"] * 1e-3) / 1e9,
                     "frac": hbm / (chosen["bp_ms"] * 1e-3) / HBM_PEAK, "traffic": hbm,
                     "fetch_bytes": c["fetch_bytes"], "write_bytes": c["write_bytes"],
                     "valu_quad_cycle_load": c.get("valu_quad_cycle_load"),
                     "valu_issue_per_simd_cycle": c.get("valu_insts", 0) / (
                         1024 * 2.4e9 * chosen["bp_ms"] * 1e-3),
                     "valu_issue_peak_fp32_per_simd_cycle": 0.25,
                     "valu_insts_per_voxel_view_wave": c.get("valu_insts", 0) * 64.0 / n_vv,
                     "limiter": "VALU issue (a wave64 FP32 instruction holds a 16-lane SIMD "
                                "for 4 cycles; the kernel issues above 0.25 per SIMD-cycle "
                                "only through its faster integer instructions), then the tap "
                                "gathers' latency; HBM at frac of peak",
                     "counters": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE / SQ_ACTIVE_INST_VALU "
                                 "on a child run of this build (bench.py --probe), calibrated "
                                 "by the abs-max pass (reads the raw scan) and the thickening "
                                 "(writes the volume)"})
    out = {
        "metric": "voxel-views/s + s/volume at 512^2, mixed+batched vs plain FP32, <=1 HU RMSE",
        "value": chosen["voxel_views_per_s"],
        "unit": "voxel-views/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": chosen["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "storage": args.precision,
        "data": "synthetic (seeded ellipsoid phantom, Poisson noise I0=5e4, generated on the GPU)",
        "config": {"workload": "full: 888ch x 64 rows, 9840 views (984/rot, pitch 0.984), "
                               "512x512x320 slices at 0.625 mm, FOV 500 mm, ramp",
                   "voxel_views": n_vv, "voxel_views_nominal": nominal_voxel_views(w, info),
                   "path": ("batch of one job: the whole chain recorded in a hipGraph and "
                            "replayed, tile depth from the autotuner (the persistent multi-job "
                            "kernel is measured by --workload batch)")
                           if args.path == "batch" else "single plan, stage calls",
                   "tile_z": chosen.get("tile_z", info.tile_z), "autotune_ms": tuned},
        "s_per_volume": chosen["ms_per_step"] * 1e-3,
        "s_per_volume_e2e": e2e[args.precision],
        "e2e_note": "scan in pinned host memory; the views some slice reads (%d of %d, views "
                    "[%d, %d)) uploaded in chunks of 256 on a copy stream, overlapped with "
                    "rebin+filter and the backprojection of the slab groups whose views have "
                    "arrived (two streams); slice ranges thickened and copied back to pinned "
                    "host memory as their groups finish"
                    % (info.read_views_end - info.read_views_first, w.scan.n_views,
                       info.read_views_first, info.read_views_end),
        "pcie": pcie,
        "plain_fp32": {"ms_per_step": res["fp32"]["ms_per_step"],
                       "s_per_volume": res["fp32"]["ms_per_step"] * 1e-3,
                       "s_per_volume_e2e": e2e["fp32"],
                       "voxel_views_per_s": res["fp32"]["voxel_views_per_s"],
                       "bp_ms": res["fp32"]["bp_ms"]},
        "speedup_vs_plain_fp32": res["fp32"]["ms_per_step"] / chosen["ms_per_step"],
        "paths": res,
        "roofline": roof,
        "counters": counters,
        "cpu_baseline": cpu,
        "hu_error_vs_fp64": hu,
    }
    print(json.dumps(out), flush=True)


def traffic_probe(args, torch):
    """Child of bench.py's counter runs (rocprofv3 --pmc): one FP16 and one
    FP32 pass of rebin+filter -> backprojection -> thickening over the full
    workload, nothing else on the GPU but the scan generation and the plans'
    set-up kernels."""
    from freect_amd import scans as S
    from freect_amd import wfbp
    w = S.full()
    ctx = wfbp.Context(0)
    raw = ctx.phantom_project(w.scan, w.phantom, w.i0, w.seed)
    torch.cuda.synchronize()
    for prec in ("fp16", "fp32"):
        plan = wfbp.Plan(ctx, w.scan, w.recon, precision=prec, tile_z=args.tile_z)
        filt, scale = plan.alloc_filtered()
        nat, vol = plan.alloc_native(), plan.alloc_volume()
        plan.rebin_filter(raw, filt, scale)
        plan.backproject(filt, scale, nat)
        plan.thicken(nat, vol)
        plan.synchronize()
        info = plan.info
        del filt, scale, nat, vol, plan
        torch.cuda.empty_cache()
    print(json.dumps({"raw_bytes": info.raw_elems * 4, "volume_bytes": info.volume_elems * 4}),
          flush=True)


def _read_counter_csv(d):
    """{(dispatch, kernel): {counter: value}} from a rocprofv3 csv output tree."""
    import csv
    import glob
    rows = {}
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                k = (int(r.get("Dispatch_Id") or 0), r.get("Kernel_Name", ""))
                rows.setdefault(k, {})
                name = r.get("Counter_Name", "")
                rows[k][name] = rows[k].get(name, 0.0) + float(r.get("Counter_Value") or 0.0)
    return rows


def measure_counters(args, tile_z, simds=1024):
    """HBM bytes and VALU activity per backprojection dispatch of this
    build, from three rocprofv3 --pmc passes over a child run (--probe):
    FETCH_SIZE, WRITE_SIZE, and SQ_INSTS_VALU + SQ_ACTIVE_INST_VALU +
    GRBM_GUI_ACTIVE.  Calibrated against dispatches of known traffic: the
    abs-max pass reads the raw scan once, the thickening writes the volume
    once.  valu_quad_cycle_load = 4 SQ_ACTIVE_INST_VALU / (SIMDs x
    GRBM_GUI_ACTIVE / XCDs): SQ_ACTIVE_INST_VALU counts each VALU instruction
    at its issue cost in quad-cycles (1 for a wave64 FP32 instruction, 2 for a
    square root) and GRBM_GUI_ACTIVE reads the sum of the 8 XCDs' cycles, so
    a dispatch saturated with FP32 instructions reads 1 (0.98 measured on a
    chain of v_fma_f32); integer additions issue in about 2.5 cycles, so an
    integer-rich mix can read above 1.  None if rocprofv3 is unavailable or a
    pass fails."""
    import shutil
    import subprocess
    import tempfile
    rp = shutil.which("rocprofv3")
    if rp is None:
        log("counters: rocprofv3 not found")
        return None
    passes = [["FETCH_SIZE"], ["WRITE_SIZE"],
              ["SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "GRBM_GUI_ACTIVE"]]
    merged = {}
    meta = None
    for counters in passes:
        d = tempfile.mkdtemp(prefix="fct_pmc_")
        cmd = ["timeout", "-s", "KILL", "180", rp, "--pmc", *counters, "--output-format", "csv",
               "-d", d, "-o", "pmc", "--", sys.executable, os.path.abspath(__file__),
               "--probe", "--tile-z", str(tile_z)]
        log("counters:", " ".join(counters))
        r = subprocess.run(cmd, capture_output=True, text=True)
        for line in r.stdout.splitlines():
            if line.startswith("{"):
                meta = json.loads(line)
        rows = _read_counter_csv(d)
        shutil.rmtree(d, ignore_errors=True)
        if not rows or meta is None:
            log("counters: pass", counters, "failed (rc", r.returncode, ")", r.stderr[-400:])
            return None
        for k, v in rows.items():
            merged.setdefault(k, {}).update(v)
    order = sorted(merged)

    def first(pred):
        for k in order:
            if pred(k[1]):
                return merged[k]
        return None

    absmax = first(lambda n: "absmax_kernel" in n)
    thk = first(lambda n: "thicken_kernel" in n)
    bps = [merged[k] for k in order if "bp_v1_kernel" in k[1]]
    if absmax is None or thk is None or len(bps) != 2:
        log("counters: unexpected dispatches", [k[1][:60] for k in order])
        return None
    fscale = meta["raw_bytes"] / absmax["FETCH_SIZE"]
    wscale = meta["volume_bytes"] / thk["WRITE_SIZE"]
    out = {"calibration": {"fetch_bytes_per_unit": fscale, "write_bytes_per_unit": wscale}}
    for name, c in zip(("bp_fp16", "bp_fp32"), bps):
        out[name] = {"fetch_bytes": c["FETCH_SIZE"] * fscale,
                     "write_bytes": c["WRITE_SIZE"] * wscale,
                     "valu_insts": c.get("SQ_INSTS_VALU"),
                     "valu_quad_cycle_load": 4.0 * c["SQ_ACTIVE_INST_VALU"] /
                     (simds * c["GRBM_GUI_ACTIVE"] / XCDS) if c.get("GRBM_GUI_ACTIVE") else None,
                     "raw": {k: c.get(k) for k in ("FETCH_SIZE", "WRITE_SIZE", "SQ_INSTS_VALU",
                                                   "SQ_ACTIVE_INST_VALU", "GRBM_GUI_ACTIVE")}}
    rf = [merged[k] for k in order if "rebin_filter" in k[1]]
    if len(rf) == 2:
        for name, c in zip(("rebin_filter_fp16", "rebin_filter_fp32"), rf):
            out[name] = {"fetch_bytes": c["FETCH_SIZE"] * fscale,
                         "write_bytes": c["WRITE_SIZE"] * wscale}
    log("counters", out)
    return out


if __name__ == "__main__":
    main()
