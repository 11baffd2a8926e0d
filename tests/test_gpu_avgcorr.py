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
"""
import numpy as np
import pytest

from cpptraj_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-9


@pytest.mark.parametrize("use_mass", [False, True])
@pytest.mark.parametrize("first", [False, True])
@pytest.mark.parametrize("stop_offset", [(-1, 1), (40, 3), (-1, 7)])
def test_avgcorr_matches_oracle(ctx, oracle, use_mass, first, stop_offset):
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
    w, ct, sd = ca.rms_avgcorr(Xd, seld, md, ref_frames=R, use_mass=use_mass, maxwindow=stop,
                               offset=offset, ctx=ctx)
    w_o, ct_o, sd_o = oracle.avgcorr(X, sel, m[sel], rc, use_mass, stop, offset)
    assert np.array_equal(w, w_o)
    tol = 5e-7 if first else TOL
    assert np.abs(ct.cpu().numpy() - ct_o).max() <= tol
    assert np.abs(sd.cpu().numpy() - sd_o).max() <= tol


def test_avgcorr_soa_frames(ctx, oracle):
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
    w, ct, sd = ca.rms_avgcorr(Sd, ref_frames=torch.as_tensor(RS).cuda(), ref_index=2, soa=True,
                               natom=n, offset=2, ctx=ctx)
    rc, _ = oracle.center_on_origin(RX[2].astype(np.float64))
    _, ct_o, sd_o = oracle.avgcorr(X.reshape(F, -1), ref_sel_centered=rc, offset=2)
    assert np.abs(ct.cpu().numpy() - ct_o).max() <= TOL
    assert np.abs(sd.cpu().numpy() - sd_o).max() <= TOL


def test_avgcorr_errors(ctx):
    import torch

    import cpptraj_amd as ca
    X = torch.zeros((1, 30), dtype=torch.float32, device="cuda")
    with pytest.raises(ca.RmsdError, match="Not enough frames"):
        ca.rms_avgcorr(X, ctx=ctx)
    Y = torch.zeros((4, 3 * 4000), dtype=torch.float32, device="cuda")
    with pytest.raises(ca.RmsdError, match="selected atoms with 'first'"):
        ca.rms_avgcorr(Y, ctx=ctx)
    Z = torch.zeros((4, 3 * 7000), dtype=torch.float32, device="cuda")
    with pytest.raises(ca.RmsdError, match="selected atoms"):
        ca.rms_avgcorr(Z, ref_frames=Z[:1], ctx=ctx)
