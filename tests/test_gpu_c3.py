"""The 64-trajectory batch workload under pytest.

64 trajectories with frames log-uniform in 500-20,000 and atoms log-uniform in
50-5,000 (synth.batch_sizes(3), the same seeded sets as bench.py's batch leg),
each a random-walk chain with sigma 0.5 A noise and random rigid motion, one
graph-captured batched plan per precision:
  * the split-BF16 and FP32 batches against the FP64 batch on every pair of
    every set: <= 1e-4 A (the stated tolerance of the mixed-precision forms);
  * the FP64 batch against the CPU oracle on sampled rows of every set:
    <= 1e-6 A;
  * the largest stated bound of the mixed-precision batch is within 1e-4 A.
"""
import os

import numpy as np
import pytest

from cpptraj_amd import synth

pytestmark = pytest.mark.gpu

SEED = 3


@pytest.fixture(scope="module")
def c3_sets():
    import cpptraj_amd as ca
    import torch
    Fs, Ns = synth.batch_sizes(SEED)
    sets, descs = [], []
    for k in range(len(Fs)):
        X, _ = synth.noisy_soa_torch(SEED * 1000 + k, int(Fs[k]), int(Ns[k]), 0.5, ld=int(Ns[k]))
        sets.append(X)
        descs.append(ca.frames_desc(X, soa=True))
    yield Fs, Ns, sets, descs
    del sets
    torch.cuda.empty_cache()


@pytest.fixture(scope="module")
def c3_fp64(ctx, c3_sets):
    Fs, Ns, sets, descs = c3_sets
    ref, rep = run_batch(ctx, descs, Fs, "fp64")
    assert rep == (0.0, 0)
    yield ref
    del ref


def run_batch(ctx, descs, Fs, precision):
    import cpptraj_amd as ca
    import torch
    outs = [torch.full((int(F) * (int(F) - 1) // 2,), -7.0, dtype=torch.float32, device="cuda")
            for F in Fs]
    plan = ca.Batch(ctx, descs, outs, "fit", False, precision, graph=True, autotune=False)
    plan.run()
    report = ctx.error_report()
    torch.cuda.synchronize()
    plan.close()
    return outs, report


def test_c3_batch_sizes_are_the_workload():
    Fs, Ns = synth.batch_sizes(SEED)
    assert len(Fs) == 64
    assert Fs.min() >= 500 and Fs.max() <= 20000 and Ns.min() >= 50 and Ns.max() <= 5000
    # the draw spans the range: both ends are exercised (502..18134 frames,
    # 54..4366 atoms)
    assert Fs.min() < 700 and Fs.max() > 15000 and Ns.min() < 80 and Ns.max() > 3500


@pytest.mark.parametrize("prec", ["bf16x3", "fp32"])
def test_c3_mixed_precision_batches_against_fp64(ctx, c3_sets, c3_fp64, prec):
    Fs, Ns, sets, descs = c3_sets
    ref = c3_fp64
    ctx.set_tolerance(1e-4)
    outs, (bound, nfp64) = run_batch(ctx, descs, Fs, prec)
    mx, worst = 0.0, -1
    for k, (a, b) in enumerate(zip(outs, ref)):
        d = float((a.double() - b.double()).abs().max())
        if d > mx:
            mx, worst = d, k
    print(f"C3 {prec}: max |batch - FP64 batch| = {mx:.3e} A (set {worst}: "
          f"{Fs[worst]} x {Ns[worst]}), stated max bound {bound:.3e}, {nfp64} pairs "
          f"recomputed in FP64")
    assert mx <= 1e-4
    assert 0.0 < bound <= 1e-4


def test_c3_fp64_batch_against_oracle(oracle, c3_sets, c3_fp64):
    ref = c3_fp64
    Fs, Ns, sets, descs = c3_sets
    nthreads = min(16, len(os.sched_getaffinity(0)))
    worst = 0.0
    for k, X in enumerate(sets):
        F, n = int(Fs[k]), int(Ns[k])
        # two rows per set (the first and one in the middle); one for the
        # largest sets so that the oracle stays within seconds
        rows = [0] if F * n > 2e7 else [0, F // 2]
        rows = np.asarray(rows, np.int32)
        aos = X.transpose(1, 2).reshape(F, 3 * n).cpu().numpy()
        exp = oracle.rms2d_rows(aos, rows, nthreads=nthreads)
        starts = [int(r) * F - int(r) * (int(r) + 1) // 2 for r in rows]
        got = np.concatenate([ref[k][s:s + F - 1 - int(r)].cpu().numpy()
                              for s, r in zip(starts, rows)])
        err = float(np.abs(got.astype(np.float64) - exp).max())
        worst = max(worst, err)
        assert err <= 1e-6, (k, F, n, err)
    print(f"C3 FP64 batch vs oracle on sampled rows of all 64 sets: max {worst:.3e} A")
