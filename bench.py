#!/usr/bin/env python3
"""Benchmark: all-pairs best-fit RMSD (rms2d) on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Headline workload: the rms2d matrix of 50,000 frames x 1,200 atoms (seeded
random-walk chain, AR(1) drift rho=0.99 sigma=0.3 A, uniform random rigid
motion), packed upper triangle of 1,249,975,000 float RMSDs.  One step = one
complete rmsd_pairwise call (per-frame stage + contraction + FP64 QCP) with the
frames already resident in HBM.

value = pairs/s of the split-BF16 MFMA kernel (hi*hi + hi*lo + lo*hi, FP32
accumulation; FP64 centroid / E0 / QCP).  Reported beside it: the plain FP64
kernel and the FP32-MFMA kernel on the same matrix, the max / RMS error of the
split-BF16 matrix against the FP64 one over every pair, the FP64 kernel against
the CPU oracle on sampled rows, the contraction kernel's roofline, the
reference's own per-pair code timed on the host cores (cpu_baseline), and the
streaming fit-to-reference workload (1,000,000 frames x 3,000 atoms) with its
HBM bandwidth.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

# The OpenMP CPU baselines run with their threads bound to consecutive cores;
# the runtime reads this when it starts, so it is set before anything loads it.
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RMSD pairs/s: MFMA split-BF16 vs plain FP64 kernel at ≤1e-4 Å max error"
PEAK_BF16_TFLOPS = 2500.0   # MI355X dense bf16 MFMA
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed_steps(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        fn()
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps  # ms per step


def profiled_main_ms(ctx, fn, reps, want_post=False):
    """Average (per-frame stage, main kernel[, last stage]) ms over `reps`
    profiled calls."""
    ctx.set_profiling(True)
    pre, main, post = [], [], []
    for _ in range(reps):
        fn()
        s = ctx.stage_ms()
        pre.append(s[0])
        main.append(s[1])
        post.append(s[2])
    ctx.set_profiling(False)
    if want_post:
        return float(np.mean(pre)), float(np.mean(main)), float(np.mean(post))
    return float(np.mean(pre)), float(np.mean(main))


def host_cpu_info():
    """CPU model, the cores this process may run on (affinity) and the cgroup
    CPU quota in cores (None when unlimited)."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    quota = None
    for path in ("/sys/fs/cgroup/cpu.max",):
        try:
            q, per = open(path).read().split()[:2]
            if q != "max":
                quota = int(q) / int(per)
        except (OSError, ValueError):
            pass
    return {"model": model, "affinity_cores": len(os.sched_getaffinity(0)),
            "cgroup_quota_cores": quota}


def baseline_threads():
    """Threads for the OpenMP CPU baselines: every core this process may use,
    capped by the cgroup quota (a GPU box grants a share of its host)."""
    info = host_cpu_info()
    n = info["affinity_cores"]
    if info["cgroup_quota_cores"]:
        n = min(n, max(1, int(info["cgroup_quota_cores"])))
    return n


def profiled_value(tag, kernel, workload):
    """A figure of `kernel` on `workload` from the committed rocprofv3 counter
    summaries (profiles/*.md lines '<!-- <tag> <kernel> <workload> <value> -->'),
    or None.  Tags: 'traffic' (HBM bytes per launch: memory reads by request
    size, plus WRITE_SIZE), 'mfma_busy' (SQ_VALU_MFMA_BUSY_CYCLES over
    4 SIMDs x CUs x GRBM_GUI_ACTIVE)."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(path) as fh:
            for m in re.finditer(r"<!-- (\S+) (\S+) (\S+) ([0-9.eE+-]+) -->", fh.read()):
                if m.group(1) == tag and m.group(2) == kernel and m.group(3) == workload:
                    return float(m.group(4))
    return None


def chunked_err(a, b, chunk=1 << 27):
    import torch
    mx = 0.0
    ss = 0.0
    n = a.numel()
    for s in range(0, n, chunk):
        d = (a[s:s + chunk].double() - b[s:s + chunk].double()).abs()
        mx = max(mx, float(d.max()))
        ss += float((d * d).sum())
    return mx, (ss / max(n, 1)) ** 0.5


def cpu_baseline(X, n_threads, target_s=12.0):
    """The reference's Frame::RMSD_CenteredRef (compiled from its sources) in the
    rms2d loop nest, on evenly spaced reference rows; pairs/s extrapolates exactly
    because every pair of a row is computed."""
    F = X.shape[0]
    try:
        from oracle.oracle import RefLib, ref_available
        if not ref_available():
            raise FileNotFoundError
        lib, kind = RefLib(), "reference"
        run = lambda rows: lib.rms2d_rows(X, rows, nthreads=n_threads, want_output=False)
    except Exception:
        from oracle.oracle import Oracle
        lib, kind = Oracle(), "port"
        run = lambda rows: len(lib.rms2d_rows(X, rows, nthreads=n_threads))
    probe = np.linspace(0, F - 2, max(2, n_threads), dtype=np.int32)
    t0 = time.perf_counter()
    n = run(probe)
    rate = n / (time.perf_counter() - t0)
    avg_row = (F - 1) / 2.0
    nrows = int(max(n_threads, min(F - 1, target_s * rate / avg_row)))
    rows = np.unique(np.linspace(0, F - 2, nrows).astype(np.int32))
    t0 = time.perf_counter()
    npairs = run(rows)
    dt = time.perf_counter() - t0
    return {"value": npairs / dt, "unit": "pairs/s", "cores": n_threads, "kind": kind,
            "sample": f"{len(rows)} evenly spaced reference rows of the 50000-frame matrix "
                      f"({npairs} pairs, {dt:.1f} s), OpenMP dynamic over rows, "
                      f"FP64 Kabsch per pair as in Analysis_Rms2d::Calculate_2D"}


def fit_stream(ctx, F, n, steps, warmup):
    """Streaming fit of F frames x n atoms (FP32 SoA) to frame 0, unweighted and
    mass-weighted, nomod.  Returns throughput and HBM bandwidth of the moments
    kernel (algorithmic bytes: 12 n per frame read + 8 B RMSD written)."""
    import torch
    import cpptraj_amd as ca
    from cpptraj_amd import synth

    X, ref = synth.noisy_soa_torch(1, F, n, 1.0, ld=n)
    m = torch.as_tensor(synth.masses(np.random.default_rng(1), n), dtype=torch.float64).cuda()
    d = ca.frames_desc(X, None, m, soa=True)
    rmsd = torch.empty(F, dtype=torch.float64, device="cuda")
    out = {}
    for use_mass in (False, True):
        fn = lambda: ctx.fit_desc(d, d, 0, True, use_mass, 0, rmsd)
        ms = timed_steps(fn, steps, warmup)
        pre, main = profiled_main_ms(ctx, fn, 3)
        algo_bytes = F * (12.0 * n + 8.0)
        key = "mass" if use_mass else "nomass"
        out[key] = {"frames_per_s": F / (ms / 1e3), "ms_per_step": ms,
                    "moments_kernel_ms": main, "hbm_gbs": algo_bytes / (main / 1e3) / 1e9}
    # parity on a sample of frames: FP64 oracle on the host
    from oracle.oracle import Oracle
    o = Oracle()
    idx = np.linspace(0, F - 1, 64).astype(np.int64)
    Xs = X[idx].cpu().numpy()  # (64, 3, n)
    aos = np.transpose(Xs, (0, 2, 1)).reshape(64, -1)
    ref0 = np.transpose(X[0].cpu().numpy(), (1, 0)).astype(np.float64)
    rc, rt = o.center_on_origin(ref0)
    r_o, _, _, _ = o.fit_frames(aos, rc, rt)
    ctx.fit_desc(d, d, 0, True, False, 0, rmsd)
    torch.cuda.synchronize()
    got = rmsd[idx].cpu().numpy()
    err = np.abs(got - r_o)
    err[0] = 0.0 if abs(got[0]) < 1e-5 else err[0]  # frame 0 fitted to itself: RMSD ~ 0
    out["max_err_vs_oracle_64_frames"] = float(err.max())
    out["workload"] = f"{F} frames x {n} atoms FP32 SoA, fit to frame 0, FP64 QCP, nomod"
    out["cpu_baseline"] = fit_cpu_baseline(X, rc, rt)
    out["perres"] = perres_bench(ctx, X, d, m, steps, warmup)
    del X
    torch.cuda.empty_cache()
    out["reftraj"] = reftraj_bench(ctx, 100000, n, steps, warmup)
    return out


def fit_cpu_baseline(X, ref_centered, ref_trans, target_s=6.0):
    """The reference's own per-frame body of Action_Rmsd::DoAction (compiled
    Frame::RMSD_CenteredRef, unweighted, nomod) over a sample of the fit
    workload's frames: serially, as the action runs, and with OpenMP over
    frames (a stronger baseline the reference does not have)."""
    from oracle.oracle import RefLib, ref_available
    if not ref_available():
        return None
    lib = RefLib()
    F = X.shape[0]

    def sample(k):
        idx = np.linspace(0, F - 1, k).astype(np.int64)
        return np.transpose(X[idx].cpu().numpy(), (0, 2, 1)).reshape(k, -1)

    probe = sample(256)
    t0 = time.perf_counter()
    lib.fit_frames(probe, ref_centered, ref_trans, nthreads=1)
    rate = 256 / (time.perf_counter() - t0)
    k = int(min(F, max(256, target_s * rate)))
    S = sample(k)
    t0 = time.perf_counter()
    lib.fit_frames(S, ref_centered, ref_trans, nthreads=1)
    serial = k / (time.perf_counter() - t0)
    nt = baseline_threads()
    t0 = time.perf_counter()
    lib.fit_frames(S, ref_centered, ref_trans, nthreads=nt)
    dt = time.perf_counter() - t0
    info = host_cpu_info()
    return {"value": serial, "unit": "frames/s", "cores": 1, "kind": "reference",
            "sample": f"{k} evenly spaced frames of the workload, serial per frame as "
                      "Action_Rmsd::DoAction runs",
            "openmp_over_frames": {"value": k / dt, "unit": "frames/s", "cores": nt,
                                   "omp_proc_bind": os.environ.get("OMP_PROC_BIND")},
            "model": info["model"]}


def reftraj_bench(ctx, F, n, steps, warmup):
    """'rmsd ... reftraj': frame f of F frames x n atoms (FP32 SoA, device
    resident) fitted to frame f of a second set, unweighted, nomod, one
    rmsd_fit_pairs call.  Algorithmic bytes: both frames read once (24 n per
    pair) + 8 B RMSD written.  CPU baseline: the compiled reference's per-pair
    body (centre the reference frame, RMSD_CenteredRef) on evenly spaced pairs,
    serial as the action runs."""
    import torch
    import cpptraj_amd as ca
    from cpptraj_amd import synth

    X, _ = synth.noisy_soa_torch(7, F, n, 1.0, ld=n)
    Y, _ = synth.noisy_soa_torch(8, F, n, 1.0, ld=n)
    dx = ca.frames_desc(X, soa=True)
    dy = ca.frames_desc(Y, soa=True)
    rmsd = torch.empty(F, dtype=torch.float64, device="cuda")
    L = ctx._L
    fn = lambda: ctx.check(L.rmsd_fit_pairs(ctx.handle, C.byref(dx), C.byref(dy), 1, 0, 0,
                                            ca._ptr(rmsd), None, None))
    ms = timed_steps(fn, steps, warmup)
    _, main = profiled_main_ms(ctx, fn, 3)
    algo = F * (24.0 * n + 8.0)
    out = {"workload": f"{F} frames x {n} atoms FP32 SoA, frame f against frame f of a "
                       "second set, fit, unweighted, nomod",
           "frames_per_s": F / (ms / 1e3), "ms_per_step": ms, "kernel_ms": main,
           "hbm_gbs": algo / (main / 1e3) / 1e9}
    # parity on 64 pairs: the FP64 oracle on the host
    from oracle.oracle import Oracle, RefLib, ref_available
    o = Oracle()
    idx = np.linspace(0, F - 1, 64).astype(np.int64)
    aos = lambda T, ids: np.transpose(T[ids].cpu().numpy(), (0, 2, 1)).reshape(len(ids), -1)
    Xs, Ys = aos(X, idx), aos(Y, idx)
    fn()
    torch.cuda.synchronize()
    got = rmsd[idx].cpu().numpy()
    exp = []
    for k in range(len(idx)):
        rc, rt = o.center_on_origin(Ys[k].reshape(n, 3).astype(np.float64))
        r, _, _, _ = o.fit_frames(Xs[k:k + 1], rc, rt)
        exp.append(r[0])
    out["max_err_vs_oracle_64_pairs"] = float(np.abs(got - np.array(exp)).max())
    if ref_available():
        lib = RefLib()
        k = 256
        ids = np.linspace(0, F - 1, k).astype(np.int64)
        t0 = time.perf_counter()
        lib.fit_pairs(aos(X, ids), aos(Y, ids), nthreads=1)
        rate = k / (time.perf_counter() - t0)
        k = int(min(F, max(256, 5.0 * rate)))
        ids = np.linspace(0, F - 1, k).astype(np.int64)
        Xs, Ys = aos(X, ids), aos(Y, ids)
        t0 = time.perf_counter()
        lib.fit_pairs(Xs, Ys, nthreads=1)
        out["cpu_baseline"] = {"value": k / (time.perf_counter() - t0), "unit": "frames/s",
                               "cores": 1, "kind": "reference",
                               "sample": f"{k} evenly spaced pairs, serial per frame as "
                                         "Action_Rmsd::DoAction runs in 'reftraj' mode",
                               "model": host_cpu_info()["model"]}
    del X, Y
    torch.cuda.empty_cache()
    return out


def perres_bench(ctx, X, d, m, steps, warmup, res_atoms=15):
    """Per-residue no-fit RMSD ('rmsd ... perres mass') over the fit workload:
    residues of `res_atoms` consecutive atoms, the fit's U and translations
    applied to the target in FP64 (the rotate + translate update), one value
    per frame and residue.  Algorithmic bytes: 12 per residue atom and frame
    read + 8 per result written."""
    import ctypes as C
    import torch
    F, n = d.nframes, d.natom
    nres = n // res_atoms
    off = np.arange(0, nres * res_atoms + 1, res_atoms, dtype=np.int32)
    atoms = np.arange(nres * res_atoms, dtype=np.int32)
    mh = m.cpu().numpy()[atoms].copy()
    U = torch.empty((F, 9), dtype=torch.float64, device="cuda")
    tv = torch.empty((F, 6), dtype=torch.float64, device="cuda")
    rmsd = torch.empty(F, dtype=torch.float64, device="cuda")
    ctx.fit_desc(d, d, 0, True, True, 0, rmsd, U, tv)
    res = torch.empty((F, nres), dtype=torch.float64, device="cuda")
    IP, DP = C.POINTER(C.c_int32), C.POINTER(C.c_double)
    L = ctx._L

    def fn():
        ctx.check(L.rmsd_perres_nofit(ctx.handle, C.byref(d), C.byref(d), 0, nres,
                                      off.ctypes.data_as(IP), atoms.ctypes.data_as(IP),
                                      atoms.ctypes.data_as(IP), mh.ctypes.data_as(DP),
                                      mh.ctypes.data_as(DP), 1, 0, 1,
                                      C.c_void_p(U.data_ptr()), C.c_void_p(tv.data_ptr()),
                                      C.c_void_p(res.data_ptr())))
    ms = timed_steps(fn, steps, warmup)
    _, main = profiled_main_ms(ctx, fn, 3)
    algo = F * (12.0 * nres * res_atoms + 8.0 * nres)
    # parity on sampled frames against the oracle (AoS copies of the frames)
    from oracle.oracle import Oracle
    idx = np.linspace(0, F - 1, 32).astype(np.int64)
    aos = np.transpose(X[idx].cpu().numpy(), (0, 2, 1)).reshape(len(idx), -1)
    ref0 = np.transpose(X[0].cpu().numpy(), (1, 0)).reshape(-1)
    exp = Oracle().perres(aos, ref0, off, atoms, tmass=mh, use_mass=True,
                          U=U[idx].cpu().numpy(), tvec=tv[idx].cpu().numpy(), modify=1)
    err = float(np.abs(res[idx].cpu().numpy() - exp).max())
    out = {"workload": f"{F} frames x {n} atoms FP32 SoA, {nres} residues of {res_atoms} atoms, "
                       "mass-weighted, fit update applied in FP64",
           "frames_per_s": F / (ms / 1e3), "ms_per_step": ms, "kernel_ms": main,
           "hbm_gbs": algo / (main / 1e3) / 1e9, "max_err_vs_oracle_32_frames": err}
    # The reference's own per-frame loop (compiled Frame::Trans_Rot_Trans,
    # CenterOnOrigin, RMSD_NoFit; serial, as the action runs) on a sample.
    from oracle.oracle import RefLib, ref_available
    if ref_available():
        ns = min(F, 40000)
        sample = np.transpose(X[:ns].cpu().numpy(), (0, 2, 1)).reshape(ns, -1)
        Us, ts = U[:ns].cpu().numpy(), tv[:ns].cpu().numpy()
        t0 = time.perf_counter()
        RefLib().perres(sample, ref0, off, atoms, tmass=mh, use_mass=True, U=Us, tvec=ts,
                        modify=1)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": ns / dt, "unit": "frames/s", "cores": 1,
                               "kind": "reference", "sample": f"first {ns} frames ({dt:.1f} s)"}
    return out


def avgcorr_bench(ctx, F=10000, n=300, steps=3, warmup=1):
    """RMS average correlation ('rmsavgcorr') over F frames x n atoms (AoS
    FP32, a CA-sized selection) against a fixed reference, every window size
    1 .. F-1: sum_w (F - w + 1) averaged frames, each an FP64 fit and a QCP
    solve.  Timed by both methods: the prefix form (the default for a fixed
    reference) and the running sum.  Parity: the first 20 window sizes
    against the oracle (stop 20), and the two methods against each other on
    every window."""
    import torch
    import cpptraj_amd as ca
    from cpptraj_amd import synth
    X = synth.noisy(5, F, n, 0.6).reshape(F, -1).astype(np.float32)
    rng = np.random.default_rng(5)
    ref = (X[0].reshape(-1, 3) + rng.normal(size=(n, 3)) * 0.3).astype(np.float32)
    Xd = torch.as_tensor(X).cuda()
    R = torch.as_tensor(ref.reshape(1, -1)).cuda()
    w = ca.avgcorr_windows(F)
    fits = float(np.sum(F - w + 1))
    out, ms, stages = {}, {}, {}
    for method in ("prefix", "running"):
        def fn():
            out[method] = ca.rms_avgcorr(Xd, ref_frames=R, method=method, ctx=ctx)
        ms[method] = timed_steps(fn, steps if method == "prefix" else 1, warmup)
        ctx.set_profiling(True)
        fn()
        stages[method] = ctx.stage_ms()
        ctx.set_profiling(False)
    # 'first': every window against its own first averaged frame.
    for method in ("prefix", "running"):
        def fn():
            out["first_" + method] = ca.rms_avgcorr(Xd, method=method, ctx=ctx)
        ms["first_" + method] = timed_steps(fn, steps if method == "prefix" else 1, warmup)
    ca.rms_avgcorr(Xd[:2], ref_frames=R, method="auto", ctx=ctx)  # back to the default
    _, ct, sd = out["prefix"]
    from oracle.oracle import Oracle
    o = Oracle()
    rc, _ = o.center_on_origin(ref.astype(np.float64))
    w20, ct20, sd20 = o.avgcorr(X, ref_sel_centered=rc, maxwindow=20)
    err = max(float(np.abs(ct[:len(w20)].cpu().numpy() - ct20).max()),
              float(np.abs(sd[:len(w20)].cpu().numpy() - sd20).max()))
    _, ct_r, sd_r = out["running"]
    err_methods = max(float((ct - ct_r).abs().max()), float((sd - sd_r).abs().max()))
    # Tile kernel: per (row, column) prefix-row pair 3 k FP64 flops (difference
    # and fused multiply-add per coordinate) over 64 x 64 pairs per tile.
    T = 64
    nt = (F + 1 + T - 1) // T
    tile_pairs = T * T * nt * (nt + 1) / 2
    tile_flops = 3.0 * 3 * n * tile_pairs
    res = {"workload": f"{F} frames x {n} atoms, fixed reference, window sizes 1..{F - 1}",
           "windows": int(len(w)), "averaged_frames": fits, "ms_per_step": ms["prefix"],
           "averaged_frames_per_s": fits / (ms["prefix"] / 1e3),
           "prefix_stages_ms": {"prefix_and_row_moments": stages["prefix"][0],
                                "tiles": stages["prefix"][1], "final": stages["prefix"][2]},
           "tiles_fp64_tflops": tile_flops / (stages["prefix"][1] / 1e3) / 1e12,
           "running_ms": ms["running"],
           "running_averaged_frames_per_s": fits / (ms["running"] / 1e3),
           "max_err_vs_oracle_first_20_windows": err,
           "max_diff_prefix_vs_running": err_methods,
           "first_prefix_ms": ms["first_prefix"], "first_running_ms": ms["first_running"],
           "first_max_diff_prefix_vs_running": max(
               float((out["first_prefix"][1] - out["first_running"][1]).abs().max()),
               float((out["first_prefix"][2] - out["first_running"][2]).abs().max()))}
    # The reference's window loop (compiled Frame operations, serial) on the
    # first 64 window sizes.
    from oracle.oracle import RefLib, ref_available
    if ref_available():
        # Evenly spaced window sizes, so that the sample's cost per window
        # matches the whole run's; OpenMP over windows as the reference builds it.
        nt = baseline_threads()
        ws = w[np.linspace(0, len(w) - 1, 16 * nt).astype(np.int64)]
        t0 = time.perf_counter()
        RefLib().avgcorr(X, ws, ref_sel_centered=rc, nthreads=nt)
        dt = time.perf_counter() - t0
        nfit = float(np.sum(F - ws + 1))
        res["cpu_baseline"] = {"value": nfit / dt, "unit": "averaged frames/s", "cores": nt,
                               "kind": "reference",
                               "sample": f"{len(ws)} evenly spaced window sizes, OpenMP dynamic "
                                         f"over windows ({dt:.1f} s)"}
    return res


def symm_bench(ctx, F=50000, n=2000, ngroups=250, steps=3, warmup=1):
    """Symmetry-corrected RMSD ('symmrmsd', fit) of F frames x n atoms (AoS
    FP32) against frame 0, with `ngroups` compact groups of 2-3 symmetric
    atoms (a protein-sized all-atom selection: methyl and amine hydrogens,
    carboxylate oxygens, ring pairs), each frame's groups permuted at random.
    One step = initial fit, per-(frame, group) assignment, remapped fit.
    Parity: 256 sampled frames against the oracle (maps and RMSD)."""
    import torch
    import cpptraj_amd as ca
    from cpptraj_amd import synth
    _, ref, groups, m, _ = synth.symm_case(9, 1, n, sizes=(3, 2, 3, 2, 2), ngroups=ngroups)
    rng = np.random.default_rng(9)
    base = (ref[None] + rng.normal(scale=0.3, size=(F, n, 3))).astype(np.float32)
    for g in groups:  # per-frame permutations inside every group
        order = np.argsort(rng.random((F, len(g))), axis=1)
        base[:, g] = base[np.arange(F)[:, None], np.asarray(g)[order]]
    Xd = torch.as_tensor(base.reshape(F, -1)).cuda()
    del base
    R = Xd[:1].clone()
    md = torch.as_tensor(m).cuda()
    out = {}

    def fn():
        out["r"] = ca.rmsd_symm(Xd, groups, ref_frames=R, mass=md, want_map=True, ctx=ctx)
    ms = timed_steps(fn, steps, warmup)
    ctx.set_profiling(True)
    fn()
    st = ctx.stage_ms()
    ctx.set_profiling(False)
    rm, amap, _, _ = out["r"]
    # parity on a sample
    from oracle.oracle import Oracle
    o = Oracle()
    idx = np.linspace(1, F - 1, 256).astype(int)
    Xs = Xd[idx].cpu().numpy()
    rc, _ = o.center_on_origin(R.cpu().numpy().reshape(-1, 3).astype(np.float64))
    eo, ao, _, _ = o.symm(Xs, None, m, rc, False, True, groups)
    err = float(np.abs(rm.cpu().numpy()[idx] - eo).max())
    maps_equal = bool(np.array_equal(amap.cpu().numpy()[idx], ao))
    S = sum(len(g) for g in groups)
    algo = F * (2 * 12.0 * n + 12.0 * S + 4.0 * n + 8)  # two reads of the frame, group atoms, map, RMSD
    res = {"workload": f"{F} frames x {n} atoms, {len(groups)} groups of 2-3 atoms, fit, "
                       "reference frame 0",
           "ms_per_step": ms, "frames_per_s": F / (ms / 1e3),
           "stages_ms": {"ref_prep_and_initial_fit": st[0], "assign_and_remapped_fit": st[1]},
           "algorithmic_GB_per_s": algo / (ms / 1e3) / 1e9,
           "max_err_vs_oracle_256_frames": err, "maps_equal_256_frames": maps_equal}
    from oracle.oracle import RefLib, ref_available
    if ref_available():
        ns = 5000
        Xc = Xd[:ns].cpu().numpy()
        t0 = time.perf_counter()
        RefLib().symm(Xc, None, m, rc, False, True, groups)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": ns / dt, "unit": "frames/s", "cores": 1,
                               "kind": "reference",
                               "sample": f"first {ns} frames ({dt:.1f} s)"}
    return res


def batch_bench(ctx, seed=3, nsets=64, precision="bf16x3"):
    """64 trajectories of uneven size (frames log-uniform in 500-20000, atoms
    log-uniform in 50-5000, each its own chain, sigma 0.5 A, rigid motion):
    one rms2d triangle each.  Times the tuned, graph-captured batch plan
    against one rmsd_pairwise call per trajectory, and checks the batch
    against the FP64 batch over every pair."""
    import torch
    import cpptraj_amd as ca
    from cpptraj_amd import synth

    rng = np.random.default_rng(seed)
    Fs = np.exp(rng.uniform(np.log(500), np.log(20000), nsets)).astype(int)
    Ns = np.exp(rng.uniform(np.log(50), np.log(5000), nsets)).astype(int)
    sets, descs, outs = [], [], []
    for k in range(nsets):
        X, _ = synth.noisy_soa_torch(seed * 1000 + k, int(Fs[k]), int(Ns[k]), 0.5, ld=int(Ns[k]))
        sets.append(X)
        descs.append(ca.frames_desc(X, soa=True))
        outs.append(torch.empty(int(Fs[k]) * (int(Fs[k]) - 1) // 2, dtype=torch.float32,
                                device="cuda"))
    pairs = Fs * (Fs - 1) // 2
    flops = float(np.sum(18.0 * Ns * pairs))
    t0 = time.perf_counter()
    plan = ca.Batch(ctx, descs, outs, "fit", False, precision, graph=True, autotune=True)
    tune_s = time.perf_counter() - t0
    info = plan.info()
    ms = timed_steps(plan.run, 5, 1)
    # one call per trajectory, same precision
    sep = [torch.empty_like(o) for o in outs]

    def separate():
        for d, o in zip(descs, sep):
            ctx.pairwise_desc(d, None, 0, False, ca.PRECISIONS[precision], o)
    ms_sep = timed_steps(separate, 3, 1)
    # FP64 batch (plain kernel) and the error of the batch against it
    ref = [torch.empty_like(o) for o in outs]
    p64 = ca.Batch(ctx, descs, ref, "fit", False, "fp64", graph=False, autotune=False)
    ms64 = timed_steps(p64.run, 1, 0)
    plan.run()
    torch.cuda.synchronize()
    mx, ss, cnt = 0.0, 0.0, 0
    for a, b in zip(outs, ref):
        d = (a.double() - b.double()).abs()
        mx = max(mx, float(d.max()))
        ss += float((d * d).sum())
        cnt += d.numel()
    # separate calls must agree bit for bit with the batch
    same = all(torch.equal(a, b) for a, b in zip(outs, sep))
    peak_exec = PEAK_BF16_TFLOPS if precision == "bf16x3" else (
        PEAK_FP32_TFLOPS if precision == "fp32" else PEAK_FP64_TFLOPS)
    terms = 3 if precision == "bf16x3" else 1
    floor_ms = flops * terms / (peak_exec * 1e12) * 1e3
    plan.close()
    p64.close()
    out = {"workload": f"{nsets} trajectories, frames logU[500,20000], atoms logU[50,5000], "
                       f"seed {seed}, FP32 SoA, fit, unweighted, {precision}",
           "pairs": int(pairs.sum()), "algorithmic_flops": flops,
           "batch_ms": ms, "pairs_per_s": float(pairs.sum()) / (ms / 1e3),
           "algo_tflops": flops / (ms / 1e3) / 1e12,
           "separate_calls_ms": ms_sep, "speedup_vs_separate": ms_sep / ms,
           "sum_of_floors_ms": floor_ms, "frac_of_floor": floor_ms / ms,
           "fp64_batch_ms": ms64, "max_err_vs_fp64": mx,
           "rms_err_vs_fp64": (ss / max(cnt, 1)) ** 0.5,
           "bit_identical_to_separate_calls": bool(same),
           "plan": info, "plan_create_s": tune_s}
    del sets, outs, sep, ref
    torch.cuda.empty_cache()
    return out


def grad_bench(ctx, steps=5, warmup=2):
    """4096 frames x 2000 atoms, masses: loss = sum_f RMSD(frame f, frame 0);
    forward (fit, nomod) and backward through rmsd_fit_backward; and the
    512 x 512 pairwise block gradient with rotations recomputed."""
    import torch
    import cpptraj_amd as ca
    from cpptraj_amd import synth
    from cpptraj_amd.autograd import fit_backward, pairwise_backward

    F, n = 4096, 2000
    X = torch.as_tensor(synth.noisy(4, F, n, 1.0)).cuda()
    m = torch.as_tensor(synth.masses(np.random.default_rng(4), n), dtype=torch.float64).cuda()
    d = ca.frames_desc(X, None, m)
    r = ca.frames_desc(X[:1], None, m)
    rmsd = torch.empty(F, dtype=torch.float64, device="cuda")
    g = torch.ones(F, dtype=torch.float64, device="cuda")
    fwd = timed_steps(lambda: ctx.fit_desc(d, r, 0, True, True, 0, rmsd), steps, warmup)
    bwd = timed_steps(lambda: fit_backward(X, X[:1], g, m, True, ctx=ctx), steps, warmup)
    A, B = X[:512], X[512:1024]
    G = torch.ones((512, 512), dtype=torch.float64, device="cuda")
    blk = timed_steps(lambda: pairwise_backward(A, B, G, m, True, ctx=ctx), steps, warmup)
    # bytes: forward reads 12 n per frame; backward re-reads 12 n and writes 24 n (FP64)
    out = {"workload": f"{F} frames x {n} atoms AoS, masses, loss = sum RMSD to frame 0; "
                       "512 x 512 block gradient",
           "forward_ms": fwd, "backward_ms": bwd,
           "backward_gbs": F * (12.0 * n + 24.0 * n) / (bwd / 1e3) / 1e9,
           "block_grad_ms": blk,
           "block_grad_pairs_per_s": 512 * 512 / (blk / 1e3)}
    del X
    torch.cuda.empty_cache()
    return out


def host_stream_bench(ctx, F=200000, n=3000):
    """The C++ host layer's 'rms' over a host-resident COORDS set (pinned FP32
    AoS, as cpptraj keeps frames): frames stream to the device in 256 MB
    chunks on two transfer slots, overlapped with the fit.  PCIe-inclusive
    rate, beside the same fit on device-resident frames."""
    import torch
    import cpptraj_amd as ca
    from cpptraj_amd import synth
    from cpptraj_amd.host import HostSession

    X, _ = synth.noisy_soa_torch(5, F, n, 1.0, ld=n)
    aos = X.permute(0, 2, 1).contiguous()  # (F, n, 3)
    del X
    d = ca.frames_desc(aos)
    rmsd = torch.empty(F, dtype=torch.float64, device="cuda")
    dev_ms = timed_steps(lambda: ctx.fit_desc(d, d, 0, True, False, 0, rmsd), 3, 1)
    host = aos.cpu().numpy().reshape(F, -1)
    del aos
    torch.cuda.empty_cache()
    hs = HostSession(0)
    hs.add_coords("big", host, n)
    del host
    hs.run("rms W first * nomod crdset big")  # warm-up: buffers, contexts
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        hs.run("rms W first * nomod crdset big")
    ms = (time.perf_counter() - t0) / reps * 1e3
    got = hs.series("W")[-F:]
    hs.close()
    ref = rmsd.cpu().numpy()
    return {"workload": f"{F} frames x {n} atoms, host COORDS set (pinned FP32 AoS), "
                        "rms first nomod through the C++ host layer",
            "host_ms": ms, "host_frames_per_s": F / (ms / 1e3),
            "pcie_inclusive_gbs": F * 12.0 * n / (ms / 1e3) / 1e9,
            "device_resident_ms": dev_ms,
            "max_diff_vs_device_resident": float(np.abs(got - ref).max())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--frames", type=int, default=50000)
    ap.add_argument("--atoms", type=int, default=1200)
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-fit", action="store_true")
    ap.add_argument("--only", choices=["bf16x3", "fp32", "fp64"], default=None,
                    help="time one precision only (profiling runs)")
    ap.add_argument("--fit-frames", type=int, default=1000000)
    ap.add_argument("--fit-atoms", type=int, default=3000)
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--skip-grad", action="store_true")
    ap.add_argument("--skip-host", action="store_true")
    ap.add_argument("--skip-avgcorr", action="store_true")
    ap.add_argument("--skip-symm", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (the pair tiles are "
              "independent; split across GPUs they would need no reduction).", file=sys.stderr)
        sys.exit(2)

    import torch
    import cpptraj_amd as ca
    from cpptraj_amd import synth

    assert torch.cuda.is_available(), "bench needs a GPU"
    F, n = args.frames, args.atoms
    npairs = F * (F - 1) // 2
    log(f"generating drift workload {F} x {n}")
    t0 = time.time()
    X = torch.as_tensor(synth.drift(2, F, n).reshape(F, -1)).cuda()
    log(f"generated in {time.time() - t0:.1f} s")
    ctx = ca.Context(0)
    d = ca.frames_desc(X)
    outs = {}
    res = {}

    def run(prec):
        out = outs.get(prec)
        if out is None:
            out = torch.empty(npairs, dtype=torch.float32, device="cuda")
            outs[prec] = out
        fn = lambda: ctx.pairwise_desc(d, None, 0, False, ca.PRECISIONS[prec], out)
        return fn

    precs = [args.only] if args.only else ["bf16x3", "fp32", "fp64"]
    for prec in precs:
        fn = run(prec)
        steps = args.fp64_steps if prec == "fp64" else args.steps
        warm = 1 if prec == "fp64" else args.warmup
        ms = timed_steps(fn, steps, warm)
        pre, main, post = profiled_main_ms(ctx, fn, 2 if prec == "fp64" else 3, want_post=True)
        bound, nfp64 = ctx.error_report()
        res[prec] = {"ms_per_step": ms, "pairs_per_s": npairs / (ms / 1e3),
                     "prep_ms": pre, "contraction_ms": main, "fp64_recompute_ms": post,
                     "stated_max_bound": bound, "pairs_recomputed_fp64": nfp64}
        log(f"{prec}: {ms:.2f} ms/step, {npairs / (ms / 1e3):.4g} pairs/s "
            f"(prep {pre:.2f} ms, contraction {main:.2f} ms, FP64 recompute {post:.2f} ms of "
            f"{nfp64} pairs; largest stated bound {bound:.3e} A)")

    head = args.only or "bf16x3"
    flops_pair = 18.0 * n
    algo_tflops = npairs * flops_pair / (res[head]["contraction_ms"] / 1e3) / 1e12
    peak = {"bf16x3": PEAK_BF16_TFLOPS, "fp32": PEAK_FP32_TFLOPS, "fp64": PEAK_FP64_TFLOPS}[head]
    # the contraction kernel of each precision (split-BF16: operands shared through LDS)
    kname = f"k_pairwise_shared<{head}>" if head == "bf16x3" else f"k_pairwise<{head}>"
    terms = 3 if head == "bf16x3" else 1
    line = {
        "metric": METRIC,
        "value": res[head]["pairs_per_s"],
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res[head]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"bf16x3": "bf16", "fp32": "f32", "fp64": "f64"}[head],
        "data": "synthetic (seeded random-walk chain + AR(1) drift + rigid motion)",
        "config": {"workload": f"rms2d all-pairs best-fit RMSD, {F} frames x {n} atoms, "
                               "unweighted, fit, packed upper-triangle float output",
                   "frames": F, "atoms": n, "pairs": npairs,
                   "precision": {"bf16x3": "split-BF16 3-term MFMA, FP32 accumulate, "
                                           "FP64 centroid/E0/QCP",
                                 "fp32": "FP32 MFMA, FP64 centroid/E0/QCP",
                                 "fp64": "FP64 MFMA, FP64 QCP"}[head]},
        "kernels": res,
        "roofline": {"bound": "mfma", "achieved": algo_tflops, "peak": peak,
                     "unit": "TFLOP/s", "frac": algo_tflops / peak,
                     "traffic": profiled_value("traffic", kname, f"{F}x{n}"),
                     "mfma_busy": profiled_value("mfma_busy", kname, f"{F}x{n}"),
                     "kernel": kname,
                     "algorithmic_flops_per_launch": npairs * flops_pair,
                     "executed_mfma_tflops": algo_tflops * terms},
    }
    if "fp64" in res and head != "fp64":
        line["fp64_pairs_per_s"] = res["fp64"]["pairs_per_s"]
        line["speedup_vs_fp64"] = res[head]["pairs_per_s"] / res["fp64"]["pairs_per_s"]
        for prec in ("bf16x3", "fp32"):
            if prec in outs:
                mx, rms = chunked_err(outs[prec], outs["fp64"])
                line[f"{prec}_max_err_vs_fp64"] = mx
                line[f"{prec}_rms_err_vs_fp64"] = rms
                log(f"{prec} vs fp64: max {mx:.3e} A, rms {rms:.3e} A over {npairs} pairs")
        # FP64 kernel vs CPU oracle on sampled rows
        from oracle.oracle import Oracle
        o = Oracle()
        rows = np.linspace(0, F - 2, 8).astype(np.int32)
        Xh = X.cpu().numpy()
        exp = o.rms2d_rows(Xh, rows, nthreads=baseline_threads())
        starts = [int(r) * F - int(r) * (int(r) + 1) // 2 for r in rows]
        got = np.concatenate([outs["fp64"][s:s + F - 1 - int(r)].cpu().numpy()
                              for s, r in zip(starts, rows)])
        line["fp64_max_err_vs_oracle_sampled_rows"] = float(np.abs(got.astype(np.float64) - exp).max())
        line["fp64_oracle_sample"] = f"rows {rows.tolist()} ({len(exp)} pairs)"
    if not args.skip_cpu:
        n_threads = baseline_threads()
        log(f"cpu baseline on {n_threads} threads")
        line["cpu_baseline"] = cpu_baseline(X.cpu().numpy(), n_threads)
        line["cpu_baseline"].update(host_cpu_info())
        line["cpu_baseline"]["omp_proc_bind"] = os.environ.get("OMP_PROC_BIND")
        log(f"cpu baseline {line['cpu_baseline']['value']:.4g} pairs/s")
    for o_ in list(outs):
        del outs[o_]
    del X
    torch.cuda.empty_cache()
    if not args.skip_fit:
        log(f"fit stream {args.fit_frames} x {args.fit_atoms}")
        line["fit_stream"] = fit_stream(ctx, args.fit_frames, args.fit_atoms, args.steps,
                                        args.warmup)
    if not args.skip_batch:
        log("batch of 64 trajectories")
        line["batch"] = batch_bench(ctx)
        log(f"batch {line['batch']['batch_ms']:.2f} ms vs separate "
            f"{line['batch']['separate_calls_ms']:.2f} ms")
    if not args.skip_grad:
        log("gradients")
        line["grad"] = grad_bench(ctx, args.steps, args.warmup)
    if not args.skip_host:
        log("host-resident fit stream")
        line["host_stream"] = host_stream_bench(ctx)
    if not args.skip_avgcorr:
        log("rms average correlation")
        line["avgcorr"] = avgcorr_bench(ctx)
        log(f"avgcorr {line['avgcorr']['ms_per_step']:.1f} ms prefix form "
            f"({line['avgcorr']['averaged_frames_per_s']:.3g} averaged frames/s), "
            f"{line['avgcorr']['running_ms']:.1f} ms running sum")
    if not args.skip_symm:
        log("symmetry-corrected RMSD")
        line["symm"] = symm_bench(ctx)
        log(f"symm {line['symm']['ms_per_step']:.1f} ms, "
            f"{line['symm']['frames_per_s']:.3g} frames/s")
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
