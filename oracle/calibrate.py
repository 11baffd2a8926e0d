#!/usr/bin/env python3
"""Times the compiled reference (oracle/_ref) against the C restatement
(oracle/_build) on the small all-pairs case (1,000 frames x 300 atoms) and on a
sample of the streaming-fit case (frames x 3,000 atoms), serially and with all
cores, and writes oracle/calibration.md.

TEST INFRASTRUCTURE ONLY: it shows that the restatement used as a CPU baseline
runs at the reference's speed (the ratio), and gives the serial and OpenMP
rates of the reference's per-frame fit (the rmsd action is serial per frame).
Run in the build container after `make -C oracle oracle ref`:

    python oracle/calibrate.py [--fit-frames 20000]
"""
import argparse
import os
import platform
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# run as a script, this folder would shadow the package of the same name
sys.path = [p for p in sys.path if os.path.abspath(p or ".") != HERE]
sys.path.insert(0, os.path.dirname(HERE))

from cpptraj_amd import synth  # noqa: E402
from oracle.oracle import Oracle, RefLib  # noqa: E402


def best(fn, reps=3):
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        t.append(time.perf_counter() - t0)
    return min(t)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--fit-frames", type=int, default=20000)
    args = ap.parse_args()
    o, r = Oracle(), RefLib()
    ncpu = os.cpu_count() or 1
    model = "unknown"
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True).stdout.splitlines():
            if line.startswith("Model name"):
                model = line.split(":", 1)[1].strip()
    except OSError:
        pass
    rows = []
    # all-pairs, 1,000 x 300 (499,500 pairs), fit
    X = synth.noisy(0, 1000, 300, 0.5).reshape(1000, -1).astype(np.float32)
    pairs = 1000 * 999 // 2
    for nt in (1, ncpu):
        tr = best(lambda: r.rms2d(X, nthreads=nt))
        to = best(lambda: o.rms2d(X, nthreads=nt))
        rows.append(("rms2d 1,000 x 300 (fit)", nt, pairs / tr, pairs / to, "pairs/s"))
    # frame-to-reference fit sample, frames x 3,000, unweighted, nomod
    F = args.fit_frames
    Y = synth.noisy(1, F, 3000, 1.0).reshape(F, -1).astype(np.float32)
    ref = Y[0].reshape(-1, 3).astype(np.float64)
    rc, rt = o.center_on_origin(ref)
    for nt in (1, ncpu):
        tr = best(lambda: r.fit_frames(Y, rc, rt, nthreads=nt), reps=2)
        to = best(lambda: o.fit_frames(Y, rc, rt, nthreads=nt), reps=2)
        rows.append((f"fit {F:,} x 3,000 (nomod)", nt, F / tr, F / to, "frames/s"))
    lines = [
        "# CPU calibration: compiled reference vs the C restatement",
        "",
        "Written by `python oracle/calibrate.py` (test infrastructure) in the build",
        f"container: {model}, {ncpu} cores, {platform.python_version()} / gcc -O3 -fopenmp.",
        "Reference = the reference's own translation units compiled by",
        "`oracle/Makefile` (oracle/_ref); restatement = oracle/rmsd_oracle.c.  Best of",
        "several runs.  The rmsd action itself is serial per frame; the threaded fit",
        "rows are an OpenMP-over-frames restatement of the same per-frame work.",
        "",
        "| workload | threads | reference | restatement | unit | restatement / reference |",
        "|---|---|---|---|---|---|",
    ]
    for name, nt, vr, vo, unit in rows:
        lines.append(f"| {name} | {nt} | {vr:.4g} | {vo:.4g} | {unit} | {vo / vr:.2f} |")
    lines.append("")
    out = os.path.join(HERE, "calibration.md")
    with open(out, "w") as fh:
        fh.write("\n".join(lines))
    print("\n".join(lines))


if __name__ == "__main__":
    main()
