"""Batched matrices (rmsd_batch_*): every set's triangle equals the
single-matrix call on the same set (bit for bit: the same per-pair arithmetic),
the FP64 form matches the oracle, graph replays track new coordinates written
into the bound buffers, and the tuner keeps its fastest variant.

Sizes are a reduced form of the 64-trajectory workload (uneven frame and atom
counts, per-set selections and masses) so that the oracle check runs in
seconds; bench.py runs the full 64-set batch."""
import numpy as np
import pytest

import cpptraj_amd as ca
from cpptraj_amd import synth

pytestmark = pytest.mark.gpu

# (frames, atoms, selected) per set: ragged sizes, 0/1-frame sets, one atom;
# (90, 128) and (41, 64) allow 16-byte loads of four atoms (AoS rows of a
# multiple of 4 floats, no selection), so the plan's prep runs both its quad and
# its atom-by-atom launch.
SIZES = [(130, 257, None), (2, 5, None), (1, 40, None), (70, 33, 20), (200, 100, 100),
         (90, 128, None), (65, 1, None), (0, 12, None), (97, 300, 151), (41, 64, None)]


def make_sets(seed, sizes=SIZES, masses=True):
    import torch
    rng = np.random.default_rng(seed)
    sets = []
    for k, (F, n, nsel) in enumerate(sizes):
        X = synth.noisy(seed * 100 + k, max(F, 1), n, 0.5)[:F]
        sel = None
        if nsel is not None:
            sel = np.sort(rng.choice(n, nsel, replace=False)).astype(np.int32)
        m = synth.masses(rng, nsel if nsel is not None else n) if masses else None
        Xt = torch.as_tensor(X.reshape(F, -1).copy() if F else np.zeros((0, 3 * n), np.float32)).cuda()
        st = None if sel is None else torch.as_tensor(sel).cuda()
        mt = None if m is None else torch.as_tensor(m).cuda()
        sets.append((X, sel, m, Xt, st, mt))
    return sets


def single(ctx, Xt, st, mt, mode, use_mass, prec):
    import torch
    F = Xt.shape[0]
    out = torch.empty(max(F * (F - 1) // 2, 1), dtype=torch.float32, device="cuda")
    if F:
        d = ca.frames_desc(Xt, st, mt)
        ctx.pairwise_desc(d, None, ca.MODES[mode], use_mass, ca.PRECISIONS[prec], out)
    return out[: F * (F - 1) // 2]


def plan(ctx, sets, mode, use_mass, prec, graph, autotune=False):
    import torch
    descs, outs = [], []
    for X, sel, m, Xt, st, mt in sets:
        F = Xt.shape[0]
        descs.append(ca.frames_desc(Xt, st, mt))
        outs.append(torch.full((max(F * (F - 1) // 2, 1),), -7.0, dtype=torch.float32,
                               device="cuda"))
    return ca.Batch(ctx, descs, outs, mode, use_mass, prec, graph=graph, autotune=autotune), outs


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_batch_equals_single_calls(ctx, prec, variant):
    import torch
    sets = make_sets(3)
    b, outs = plan(ctx, sets, "fit", True, prec, graph=False)
    b.set_variant(variant)
    b.run()
    for k, (X, sel, m, Xt, st, mt) in enumerate(sets):
        F = Xt.shape[0]
        exp = single(ctx, Xt, st, mt, "fit", True, prec)
        got = outs[k][: F * (F - 1) // 2]
        assert torch.equal(got, exp), (prec, variant, k)
    b.close()


def maxerr(got, exp):
    """Largest |got - exp|; NaN matches NaN (the reference's QCP gives NaN for a
    single atom), NaN against a number counts as infinite."""
    got = np.asarray(got, np.float64)
    exp = np.asarray(exp, np.float64)
    d = np.abs(got - exp)
    d[np.isnan(got) & np.isnan(exp)] = 0.0
    d[np.isnan(d)] = np.inf
    return d.max() if d.size else 0.0


@pytest.mark.parametrize("mode", ["fit", "qrmsd", "nofit"])
def test_batch_fp64_matches_oracle(oracle, ctx, mode):
    sets = make_sets(5)
    b, outs = plan(ctx, sets, mode, True, "fp64", graph=True)
    try:
        b.run()
        ctx.synchronize()
        for k, (X, sel, m, Xt, st, mt) in enumerate(sets):
            F = Xt.shape[0]
            if F < 2:
                continue
            exp = oracle.rms2d(X.reshape(F, -1), sel, m,
                               mode={"fit": 0, "nofit": 1, "qrmsd": 2}[mode], use_mass=True)
            got = outs[k][: F * (F - 1) // 2].cpu().numpy()
            nsel = len(sel) if sel is not None else X.shape[1]
            # one or two atoms: no unique rotation, the distances agree to sqrt's
            # accuracy near 0 (as for the single-matrix call)
            tol = 1e-6 if (nsel >= 3 or mode == "nofit") else 1e-4
            assert maxerr(got, exp) <= tol, (mode, k)
    finally:
        b.close()


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
def test_graph_replays_follow_new_coordinates(ctx, prec):
    import torch
    sets = make_sets(7, masses=False)
    b, outs = plan(ctx, sets, "fit", False, prec, graph=True)
    assert b.info()["graph"]
    gen = torch.Generator(device="cuda")
    gen.manual_seed(11)
    for rep in range(3):
        for X, sel, m, Xt, st, mt in sets:
            if Xt.numel():
                Xt.add_(0.05 * torch.randn(Xt.shape, generator=gen, device="cuda"))
        b.run()
        torch.cuda.synchronize()
        for k, (X, sel, m, Xt, st, mt) in enumerate(sets):
            F = Xt.shape[0]
            exp = single(ctx, Xt, st, mt, "fit", False, prec)
            ctx.synchronize()
            assert torch.equal(outs[k][: F * (F - 1) // 2], exp), (rep, k)
    b.close()


def test_autotune_keeps_fastest(ctx):
    sets = make_sets(9, masses=False)
    b, outs = plan(ctx, sets, "fit", False, "bf16x3", graph=True, autotune=True)
    inf = b.info()
    t = inf["tune_ms"]
    assert all(x > 0 for x in t)
    assert inf["variant"] == int(np.argmin(t))
    assert inf["tile"] == (64 if inf["variant"] < 2 else 32)
    b.run()
    ctx.synchronize()
    b.close()


def test_one_shot_and_empty(ctx):
    import torch
    sets = make_sets(13, sizes=[(40, 30, None), (33, 64, None)], masses=False)
    got = ca.rms2d_batch([s[3] for s in sets], precision="fp64", ctx=ctx)
    for k, s in enumerate(sets):
        assert torch.equal(got[k], single(ctx, s[3], None, None, "fit", False, "fp64"))
    assert ca.rms2d_batch([], ctx=ctx) == []
