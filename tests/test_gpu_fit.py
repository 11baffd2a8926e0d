"""GPU parity of the streaming fit-to-reference path (rmsd action) with the oracle.

Tolerances: RMSD <= 1e-6 A, except where RMSD ~ 0 (the reference frame fitted
to itself), where sqrt magnifies FP64 rounding of E0 - lambda and RMSD^2 is
compared instead, to a few FP64 ulps of 2 E0 / M; U <= 1e-9 (QCP eigenvector
vs the reference's Jacobi-Kabsch); Trans / refTrans <= 1e-9 A; rewritten FP32
coordinates within 2 float ulps of the oracle's; nofit <= 1e-12 relative (the
per-frame sum is a parallel reduction, not the reference's sequential order).
"""
import numpy as np
import pytest

from cpptraj_amd import synth

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    a = np.ascontiguousarray(a)
    dt = {np.float32: torch.float32, np.float64: torch.float64, np.int32: torch.int32}[a.dtype.type]
    return torch.as_tensor(a, dtype=dt).cuda()


def assert_rmsd_close(got, exp, X, tol=1e-6):
    """|got - exp| <= tol, or |got^2 - exp^2| <= 64 ulp(2 E0 / M) near zero."""
    F = X.shape[0]
    x = X.reshape(F, -1, 3).astype(np.float64)
    scale = ((x - x.mean(1, keepdims=True)) ** 2).sum(2).mean()  # ~ E0 / M
    tol_sq = 64 * np.finfo(np.float64).eps * 2 * scale
    d = np.abs(got - exp)
    d2 = np.abs(got ** 2 - exp ** 2)
    bad = (d > tol) & (d2 > tol_sq)
    assert not bad.any(), (d.max(), d2[bad][:5], np.nonzero(bad)[0][:5])


def oracle_fit(oracle, X, sel, m, use_mass, fit=True, modify=0):
    F = X.shape[0]
    ref = X[0].reshape(-1, 3).astype(np.float64)
    if sel is not None:
        ref = ref[sel]
    if fit:
        rc, rt = oracle.center_on_origin(ref, m, use_mass)
    else:
        rc, rt = ref, np.zeros(3)
    return oracle.fit_frames(X, rc, rt, sel=sel, mass=m, fit=fit, use_mass=use_mass,
                             modify=modify, want_frames=True)


@pytest.mark.parametrize("layout", ["aos", "aos_sel", "soa"])
@pytest.mark.parametrize("use_mass", [False, True])
def test_fit_matches_oracle(oracle, ctx, layout, use_mass):
    import cpptraj_amd as ca
    F, n = 600, 3000
    X3 = synth.noisy(1, F, n, 1.0)
    m_all = synth.masses(np.random.default_rng(1), n)
    sel = None
    if layout == "aos_sel":
        sel = np.sort(np.random.default_rng(3).choice(n, 2000, replace=False)).astype(np.int32)
    m = m_all if sel is None else m_all[sel]
    X = X3.reshape(F, -1)
    r_o, U_o, tv_o, _ = oracle_fit(oracle, X, sel, m, use_mass)
    if layout == "soa":
        Xd = dev(synth.to_soa(X3, ld=3008))
    else:
        Xd = dev(X)
    r, U, tv = ca.rmsd_fit(Xd, None, 0, mask=None if sel is None else dev(sel), mass=dev(m),
                           use_mass=use_mass, modify="nomod", save_matrices=True,
                           save_vectors=True, soa=(layout == "soa"), natom=n, ctx=ctx)
    r, U, tv = r.cpu().numpy(), U.cpu().numpy(), tv.cpu().numpy()
    assert_rmsd_close(r, r_o, X)
    assert np.abs(U - U_o).max() <= 1e-9
    assert np.abs(tv - tv_o).max() <= 1e-9


@pytest.mark.parametrize("modify", [1, 2])
def test_coordinate_update(oracle, ctx, modify):
    import cpptraj_amd as ca
    F, n = 200, 500
    X = synth.noisy(2, F, n, 1.0).reshape(F, -1)
    sel = np.arange(10, 400, 2, dtype=np.int32)
    _, _, _, fo = oracle_fit(oracle, X, sel, None, False, modify=modify)
    Xd = dev(X)
    ca.rmsd_fit(Xd, None, 0, mask=dev(sel), modify={1: "rottrans", 2: "norotate"}[modify], ctx=ctx)
    got = Xd.cpu().numpy()
    ulp = np.spacing(np.abs(fo).astype(np.float32))
    assert np.all(np.abs(got - fo) <= 2 * ulp + 1e-30)


@pytest.mark.parametrize("use_mass", [False, True])
def test_nofit_matches_oracle(oracle, ctx, use_mass):
    import cpptraj_amd as ca
    F, n = 300, 777
    X = synth.noisy(3, F, n, 1.0).reshape(F, -1)
    m = synth.masses(np.random.default_rng(3), n)
    r_o, _, _, _ = oracle_fit(oracle, X, None, m, use_mass, fit=False)
    r, _, _ = ca.rmsd_fit(dev(X), None, 0, mass=dev(m), fit=False, use_mass=use_mass,
                          modify="nomod", ctx=ctx)
    np.testing.assert_allclose(r.cpu().numpy(), r_o, rtol=1e-12, atol=0)


def test_zero_mass_sentinel(ctx):
    import cpptraj_amd as ca
    X = synth.noisy(4, 10, 30, 1.0).reshape(10, -1)
    r, _, _ = ca.rmsd_fit(dev(X), None, 0, mass=dev(np.zeros(30)), use_mass=True, modify="nomod",
                          ctx=ctx)
    assert np.all(r.cpu().numpy() == -1.0)


def test_large_offsets(oracle, ctx):
    """Frames far from the origin: the single-pass moments stay exact to 1e-6 A."""
    import cpptraj_amd as ca
    X3 = synth.noisy(5, 100, 400, 0.3) + np.float32(900.0)
    X = X3.reshape(100, -1)
    r_o, _, _, _ = oracle_fit(oracle, X, None, None, False)
    r, _, _ = ca.rmsd_fit(dev(X), None, 0, modify="nomod", ctx=ctx)
    assert_rmsd_close(r.cpu().numpy(), r_o, X)


@pytest.mark.parametrize("F", [1, 7, 601])
@pytest.mark.parametrize("use_mass", [False, True])
def test_fit_soa_ragged(oracle, ctx, F, use_mass):
    """SoA rows with an atom count that is not a multiple of four (the vector
    path's scalar tail) and frame counts that leave a wave with one frame."""
    import cpptraj_amd as ca
    n = 2999
    X3 = synth.noisy(5, F, n, 1.0)
    m = synth.masses(np.random.default_rng(5), n)
    X = X3.reshape(F, -1)
    r_o, U_o, tv_o, _ = oracle_fit(oracle, X, None, m, use_mass)
    Xd = dev(synth.to_soa(X3, ld=3008))
    r, U, tv = ca.rmsd_fit(Xd, None, 0, mass=dev(m), use_mass=use_mass, modify="nomod",
                           save_matrices=True, save_vectors=True, soa=True, natom=n, ctx=ctx)
    assert_rmsd_close(r.cpu().numpy(), r_o, X)
    assert np.abs(U.cpu().numpy() - U_o).max() <= 1e-9
    assert np.abs(tv.cpu().numpy() - tv_o).max() <= 1e-9


@pytest.mark.parametrize("use_mass", [False, True])
def test_fit_aos_padded_stride_tail(oracle, ctx, use_mass):
    # AoS frames of 301 atoms stored 908 floats apart (a 16-byte-aligned stride
    # with trailing values, as box data would be): the vector AoS path takes
    # 300 atoms four at a time and the last one alone.
    import cpptraj_amd as ca
    F, n, stride = 200, 301, 908
    X3 = synth.noisy(5, F, n, 0.8)
    m = synth.masses(np.random.default_rng(5), n)
    Xp = np.full((F, stride), 7.5, np.float32)
    Xp[:, :3 * n] = X3.reshape(F, -1)
    r_o, U_o, tv_o, _ = oracle_fit(oracle, X3.reshape(F, -1), None, m, use_mass)
    r, U, tv = ca.rmsd_fit(dev(Xp), None, 0, mass=dev(m), use_mass=use_mass, modify="nomod",
                           save_matrices=True, save_vectors=True, natom=n, ctx=ctx)
    r, U, tv = r.cpu().numpy(), U.cpu().numpy(), tv.cpu().numpy()
    assert_rmsd_close(r, r_o, X3.reshape(F, -1))
    assert np.abs(U - U_o).max() <= 1e-9
    assert np.abs(tv - tv_o).max() <= 1e-9
