"""RMS average correlation (rmsd_avgcorr) against the oracle.

The library fits every averaged frame by single-pass FP64 moments and QCP,
the oracle by the reference's two-pass Kabsch; the running sums are the same
FP64 sequence.  Bar: 1e-9 A on the mean and the SD of every window with a
fixed reference.  With 'first', every window fits its first averaged frame
to itself: the true RMSD is 0, where the reference's Kabsch eigenvalue
formula returns rounding noise up to ~sqrt(eps E0 / M) ~ 1e-7 A and QCP
~1e-8 A, so that one sample moves the window's mean and SD by up to ~1e-7 A:
the bar there is 5e-7 A.  Cases:
fixed reference and 'first', mass weighting, 'stop' and 'offset', a
selection, SoA frames, and the argument errors.

Both methods of rmsd_avgcorr_set_method run every case: the running sum (one
wavefront per window) and the prefix form (window sums as differences of
prefix sums of the mean-subtracted coordinates, every averaged frame fitted
independently in tiles of prefix rows: 64 x 64 with a fixed reference,
32 x 32 with 'first', whose window references are staged per diagonal).
Both meet the same bars; the prefix form also past the running method's LDS
atom limit, across many tiles, on a drifting trajectory, and against the
running method on 3,000 frames.
"""
import numpy as np
import pytest

from cpptraj_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-9


def _avgcorr(ca, ctx, *args, method="auto", **kw):
    """rms_avgcorr with the method set for this call only (the setting stays
    with the shared context otherwise)."""
    try:
        return ca.rms_avgcorr(*args, method=method, ctx=ctx, **kw)
    finally:
        ctx.check(ctx._L.rmsd_avgcorr_set_method(ctx.handle, 0))


@pytest.mark.parametrize("use_mass", [False, True])
@pytest.mark.parametrize("first,method", [(False, "running"), (False, "prefix"), (True, "running"),
                                          (True, "prefix")])
@pytest.mark.parametrize("stop_offset", [(-1, 1), (40, 3), (-1, 7)])
def test_avgcorr_matches_oracle(ctx, oracle, use_mass, first, method, stop_offset):
    import torch

    import cpptraj_amd as ca
    stop, offset = stop_offset
    rng = np.random.default_rng(41 + use_mass + 2 * first)
    F, n = 97, 150
    X = synth.noisy(9, F, n, 0.5).reshape(F, -1).astype(np.float32)
    sel = np.sort(rng.choice(n, 101, replace=False)).astype(np.int32)
    m = synth.masses(rng, n)
    Xd = torch.as_tensor(X).cuda()
    seld = torch.as_tensor(sel).cuda()
    md = torch.as_tensor(m[sel]).cuda()
    R = None
    rc = None
    if not first:
        ref = (X[3].reshape(-1, 3) + rng.normal(size=(n, 3)) * 0.3).astype(np.float32)
        R = torch.as_tensor(ref.reshape(1, -1)).cuda()
        rc, _ = oracle.center_on_origin(ref[sel].astype(np.float64), m[sel], use_mass)
    w, ct, sd = _avgcorr(ca, ctx, Xd, seld, md, ref_frames=R, use_mass=use_mass, maxwindow=stop,
                         offset=offset, method=method)
    w_o, ct_o, sd_o = oracle.avgcorr(X, sel, m[sel], rc, use_mass, stop, offset)
    assert np.array_equal(w, w_o)
    tol = 5e-7 if first else TOL
    assert np.abs(ct.cpu().numpy() - ct_o).max() <= tol
    assert np.abs(sd.cpu().numpy() - sd_o).max() <= tol


@pytest.mark.parametrize("method", ["running", "prefix"])
def test_avgcorr_soa_frames(ctx, oracle, method):
    import torch

    import cpptraj_amd as ca
    F, n, ld = 60, 77, 96
    X = synth.noisy(10, F, n, 0.5).reshape(F, n, 3).astype(np.float32)
    S = np.zeros((F, 3, ld), np.float32)
    S[:, :, :n] = X.transpose(0, 2, 1)
    Sd = torch.as_tensor(S).cuda()
    # reference: frame 2 of a separate SoA set (perturbed copies, so that no
    # window compares a frame with itself)
    rng = np.random.default_rng(10)
    RX = np.stack([X[k] + rng.normal(size=(n, 3)).astype(np.float32) * 0.4 for k in (1, 4, 7)])
    RS = np.zeros((3, 3, ld), np.float32)
    RS[:, :, :n] = RX.transpose(0, 2, 1)
    w, ct, sd = _avgcorr(ca, ctx, Sd, ref_frames=torch.as_tensor(RS).cuda(), ref_index=2,
                         soa=True, natom=n, offset=2, method=method)
    rc, _ = oracle.center_on_origin(RX[2].astype(np.float64))
    _, ct_o, sd_o = oracle.avgcorr(X.reshape(F, -1), ref_sel_centered=rc, offset=2)
    assert np.abs(ct.cpu().numpy() - ct_o).max() <= TOL
    assert np.abs(sd.cpu().numpy() - sd_o).max() <= TOL


@pytest.mark.parametrize("use_mass", [False, True])
def test_avgcorr_prefix_many_tiles_with_drift(ctx, oracle, use_mass):
    # 300 frames (5 x 5 row tiles, windows on every diagonal class) of a
    # trajectory that drifts 8 A, so that the prefix sums are large.
    import torch

    import cpptraj_amd as ca
    rng = np.random.default_rng(77)
    F, n = 300, 60
    X = synth.noisy(11, F, n, 0.4).reshape(F, n, 3).astype(np.float64)
    X += np.linspace(0.0, 8.0, F)[:, None, None] * rng.normal(size=(1, n, 3)) + 40.0
    X = X.reshape(F, -1).astype(np.float32)
    m = synth.masses(rng, n)
    ref = (X[5].reshape(n, 3) + rng.normal(size=(n, 3)) * 0.2).astype(np.float32)
    rc, _ = oracle.center_on_origin(ref.astype(np.float64), m, use_mass)
    for stop, offset in ((-1, 1), (200, 64), (130, 5)):
        w, ct, sd = _avgcorr(ca, ctx, torch.as_tensor(X).cuda(), mass=torch.as_tensor(m).cuda(),
                             ref_frames=torch.as_tensor(ref.reshape(1, -1)).cuda(),
                             use_mass=use_mass, maxwindow=stop, offset=offset, method="prefix")
        w_o, ct_o, sd_o = oracle.avgcorr(X, None, m, rc, use_mass, stop, offset)
        assert np.array_equal(w, w_o)
        assert np.abs(ct.cpu().numpy() - ct_o).max() <= TOL
        assert np.abs(sd.cpu().numpy() - sd_o).max() <= TOL


def test_avgcorr_prefix_beyond_running_atom_limit(ctx, oracle):
    # 7,000 selected atoms: over the running method's LDS limit; the default
    # (auto) method takes the prefix form.
    import torch

    import cpptraj_amd as ca
    rng = np.random.default_rng(78)
    F, n = 24, 7000
    X = synth.noisy(12, F, n, 0.5).reshape(F, -1).astype(np.float32)
    ref = (X[2].reshape(n, 3) + rng.normal(size=(n, 3)) * 0.3).astype(np.float32)
    rc, _ = oracle.center_on_origin(ref.astype(np.float64))
    w, ct, sd = _avgcorr(ca, ctx, torch.as_tensor(X).cuda(),
                         ref_frames=torch.as_tensor(ref.reshape(1, -1)).cuda())
    _, ct_o, sd_o = oracle.avgcorr(X, ref_sel_centered=rc)
    assert np.abs(ct.cpu().numpy() - ct_o).max() <= TOL
    assert np.abs(sd.cpu().numpy() - sd_o).max() <= TOL


def test_avgcorr_prefix_matches_running_3000_frames(ctx):
    # Every window of 3,000 frames x 200 atoms: the two methods agree.
    import torch

    import cpptraj_amd as ca
    rng = np.random.default_rng(79)
    F, n = 3000, 200
    X = synth.noisy(13, F, n, 0.6).reshape(F, -1).astype(np.float32)
    m = synth.masses(rng, n)
    Xd, md = torch.as_tensor(X).cuda(), torch.as_tensor(m).cuda()
    R = torch.as_tensor(X[7:8] + rng.normal(size=(1, 3 * n)).astype(np.float32) * 0.3).cuda()
    out = {}
    for method in ("running", "prefix"):
        w, ct, sd = _avgcorr(ca, ctx, Xd, mass=md, ref_frames=R, use_mass=True, method=method)
        out[method] = (ct.cpu().numpy(), sd.cpu().numpy())
    assert len(w) == F - 1
    assert np.abs(out["running"][0] - out["prefix"][0]).max() <= TOL
    assert np.abs(out["running"][1] - out["prefix"][1]).max() <= TOL


@pytest.mark.parametrize("use_mass", [False, True])
def test_avgcorr_first_prefix_matches_running_3000_frames(ctx, use_mass):
    # 'first': every window's reference is its own first averaged frame.
    import torch

    import cpptraj_amd as ca
    rng = np.random.default_rng(80)
    F, n = 3000, 150
    X = synth.noisy(14, F, n, 0.6).reshape(F, -1).astype(np.float32)
    m = synth.masses(rng, n)
    Xd, md = torch.as_tensor(X).cuda(), torch.as_tensor(m).cuda()
    out = {}
    for method in ("running", "prefix"):
        w, ct, sd = _avgcorr(ca, ctx, Xd, mass=md, use_mass=use_mass, method=method)
        out[method] = (ct.cpu().numpy(), sd.cpu().numpy())
    assert np.abs(out["running"][0] - out["prefix"][0]).max() <= 5e-7
    assert np.abs(out["running"][1] - out["prefix"][1]).max() <= 5e-7


def test_avgcorr_errors(ctx):
    import torch

    import cpptraj_amd as ca
    X = torch.zeros((1, 30), dtype=torch.float32, device="cuda")
    with pytest.raises(ca.RmsdError, match="Not enough frames"):
        ca.rms_avgcorr(X, ctx=ctx)
    Y = torch.zeros((4, 3 * 4000), dtype=torch.float32, device="cuda")
    with pytest.raises(ca.RmsdError, match="selected atoms with 'first'"):
        _avgcorr(ca, ctx, Y, method="running")
    Z = torch.zeros((4, 3 * 7000), dtype=torch.float32, device="cuda")
    with pytest.raises(ca.RmsdError, match="selected atoms"):
        _avgcorr(ca, ctx, Z, ref_frames=Z[:1], method="running")
    with pytest.raises(ca.RmsdError, match="unknown rmsavgcorr method"):
        ctx.check(ctx._L.rmsd_avgcorr_set_method(ctx.handle, 7))
