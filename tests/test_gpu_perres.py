"""Per-residue no-fit RMSD (rmsd_perres_nofit) against the oracle.

The kernel sums atom by atom in the reference's order with FP64 contraction
off, so it agrees with the oracle to rounding of the final square root at
most; the bar is 1e-12 A.  Cases: ragged and scattered residues, residues
larger than a wavefront, a residue that needs a workgroup piece of its own,
tables cut into many pieces, masses, centring, the fit's coordinate update
(rotate + translate, translate only), SoA frames, frame lists, a reference
frame taken from another set, zero total mass, and argument errors.
"""
import numpy as np
import pytest

from cpptraj_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _table(rng, natom, sizes, scattered=0):
    """Residues of the given sizes over consecutive atoms, plus one residue of
    `scattered` atoms picked across the whole frame."""
    off = [0]
    atoms = []
    a = 0
    for s in sizes:
        s = int(min(s, natom - a))
        if s <= 0:
            break
        atoms.append(np.arange(a, a + s))
        a += s
        off.append(off[-1] + s)
    if scattered:
        atoms.append(np.sort(rng.choice(natom, size=scattered, replace=False)))
        off.append(off[-1] + scattered)
    return np.array(off, np.int32), np.concatenate(atoms).astype(np.int32)


def _case(seed, F, natom, sizes, scattered=7):
    rng = np.random.default_rng(seed)
    X = synth.noisy(seed, F, natom, 0.6).reshape(F, -1).astype(np.float32)
    ref = (X[0] + rng.normal(size=3 * natom)).astype(np.float32)
    off, atoms = _table(rng, natom, sizes, scattered)
    m = synth.masses(rng, natom)
    return rng, X, ref, off, atoms, m


def _gpu(X, ref, off, atoms, **kw):
    import torch

    import cpptraj_amd as ca
    Xd = torch.as_tensor(X).cuda()
    R = torch.as_tensor(ref.reshape(1, -1)).cuda()
    return ca.rmsd_perres(Xd, off, atoms, ref_frames=R, **kw).cpu().numpy()


@pytest.mark.parametrize("use_mass", [False, True])
@pytest.mark.parametrize("center", [False, True])
def test_perres_ragged(ctx, oracle, use_mass, center):
    rng, X, ref, off, atoms, m = _case(3, 37, 700, [1, 2, 5, 13, 24, 64, 65, 130, 7, 300])
    ra = rng.permutation(len(ref) // 3)[atoms].astype(np.int32)
    tm, rm = m[atoms], synth.masses(rng, len(ref) // 3)[ra]
    kw = dict(ref_atoms=ra, tmass=tm, rmass=rm, use_mass=use_mass, center=center)
    exp = oracle.perres(X, ref, off, atoms, **kw)
    got = _gpu(X, ref, off, atoms, ctx=ctx, **kw)
    assert got.shape == exp.shape == (37, len(off) - 1)
    assert np.abs(got - exp).max() <= TOL


@pytest.mark.parametrize("modify", ["rottrans", "norotate"])
def test_perres_after_fit_update(ctx, oracle, modify):
    import torch
    rng, X, ref, off, atoms, m = _case(4, 50, 400, [10, 20, 15, 9, 31, 70, 22])
    F = X.shape[0]
    U = synth.rotations(rng, F).reshape(F, 3, 3)
    tv = rng.uniform(-15, 15, size=(F, 6))
    mod = {"rottrans": 1, "norotate": 2}[modify]
    exp = oracle.perres(X, ref, off, atoms, tmass=m[atoms], use_mass=True, U=U, tvec=tv,
                        modify=mod)
    got = _gpu(X, ref, off, atoms, tmass=m[atoms], use_mass=True, modify=modify,
               U=torch.as_tensor(U).cuda(), tvec=torch.as_tensor(tv).cuda(), ctx=ctx)
    assert np.abs(got - exp).max() <= TOL


def test_perres_many_pieces_and_frames(ctx, oracle):
    # 20,000 atoms in 1,300 residues: the table is cut into many workgroup
    # pieces; 300 frames walk through several frame batches per workgroup.
    sizes = np.random.default_rng(5).integers(5, 26, size=1300)
    rng, X, ref, off, atoms, m = _case(5, 300, 20000, sizes, scattered=40)
    exp = oracle.perres(X, ref, off, atoms, tmass=m[atoms], use_mass=True, center=True)
    got = _gpu(X, ref, off, atoms, tmass=m[atoms], use_mass=True, center=True, ctx=ctx)
    assert np.abs(got - exp).max() <= TOL


def test_perres_oversized_residue(ctx, oracle):
    # One 3,000-atom residue (its LDS image exceeds a shared piece) between
    # small ones; a 5,000-atom residue is beyond one workgroup and rejected.
    # The scattered residue of the other tests is gathered, not spanned.
    import torch

    import cpptraj_amd as ca
    rng, X, ref, off, atoms, m = _case(6, 9, 3100, [12, 3000, 40], scattered=0)
    exp = oracle.perres(X, ref, off, atoms, center=True)
    got = _gpu(X, ref, off, atoms, center=True, ctx=ctx)
    assert np.abs(got - exp).max() <= TOL
    _, X2, ref2, off2, atoms2, _ = _case(7, 2, 5000, [5000], scattered=0)
    with pytest.raises(ca.RmsdError, match="too many atoms"):
        ca.rmsd_perres(torch.as_tensor(X2).cuda(), off2, atoms2, ctx=ctx)


def test_perres_soa_fidx_and_reference_set(ctx, oracle):
    import torch

    import cpptraj_amd as ca
    rng, X, ref, off, atoms, m = _case(8, 40, 333, [11, 23, 17, 64, 2, 90])
    F, n = X.shape[0], X.shape[1] // 3
    ld = 384
    S = np.zeros((F, 3, ld), np.float32)
    S[:, :, :n] = X.reshape(F, n, 3).transpose(0, 2, 1)
    Sd = torch.as_tensor(S).cuda()
    # frame list: every third stored frame, reversed
    fl = np.arange(F)[::3][::-1].copy()
    # The descriptor holds raw pointers: the frame list tensor must stay alive.
    fl_d = torch.as_tensor(fl.astype(np.int32)).cuda()
    d = ca.frames_desc(Sd, soa=True, natom=n, fidx=fl_d)
    # reference: stored frame 5 of a second (AoS) set
    Rset = np.stack([X[(k * 7) % F] for k in range(8)])
    Rd = torch.as_tensor(Rset).cuda()
    r = ca.frames_desc(Rd)
    out = torch.empty((len(fl), len(off) - 1), dtype=torch.float64, device="cuda")
    import ctypes as C
    IP, DP = C.POINTER(C.c_int32), C.POINTER(C.c_double)
    L = ctx._L
    ctx.check(L.rmsd_perres_nofit(ctx.handle, C.byref(d), C.byref(r), 5, len(off) - 1,
                                  off.ctypes.data_as(IP), atoms.ctypes.data_as(IP),
                                  atoms.ctypes.data_as(IP), None, None, 0, 1, 0, None, None,
                                  C.c_void_p(out.data_ptr())))
    got = out.cpu().numpy()
    del fl_d
    exp = oracle.perres(X[fl], Rset[5], off, atoms, center=True)
    assert np.abs(got - exp).max() <= TOL


def test_perres_table_changes_and_sentinels(ctx, oracle):
    # Two calls with different tables on one context (the cached device copy
    # must follow), a zero-mass residue (-1) and an empty residue (-1).
    rng, X, ref, off, atoms, m = _case(9, 20, 200, [10, 20, 30])
    tm = m[atoms].copy()
    tm[off[1]:off[2]] = 0.0
    exp = oracle.perres(X, ref, off, atoms, tmass=tm, use_mass=True)
    got = _gpu(X, ref, off, atoms, tmass=tm, use_mass=True, ctx=ctx)
    assert np.abs(got - exp).max() <= TOL and np.all(got[:, 1] == -1.0)
    off2 = np.array([0, 5, 5, 12], np.int32)
    atoms2 = np.arange(100, 112, dtype=np.int32)
    exp2 = oracle.perres(X, ref, off2, atoms2)
    got2 = _gpu(X, ref, off2, atoms2, ctx=ctx)
    assert np.abs(got2 - exp2).max() <= TOL and np.all(got2[:, 1] == -1.0)


def test_perres_argument_errors(ctx):
    import torch

    import cpptraj_amd as ca
    X = torch.zeros((3, 30), dtype=torch.float32, device="cuda")
    off = np.array([0, 3, 6], np.int32)
    with pytest.raises(ca.RmsdError, match="out of range"):
        ca.rmsd_perres(X, off, np.array([0, 1, 2, 3, 4, 10], np.int32), ctx=ctx)
    with pytest.raises(ca.RmsdError, match="decreasing"):
        ca.rmsd_perres(X, np.array([0, 4, 2], np.int32), np.arange(4, dtype=np.int32), ctx=ctx)
    with pytest.raises(ca.RmsdError, match="masses"):
        ca.rmsd_perres(X, off, np.arange(6, dtype=np.int32), use_mass=True, ctx=ctx)
    with pytest.raises(ca.RmsdError, match="needs tvec"):
        ca.rmsd_perres(X, off, np.arange(6, dtype=np.int32), modify="rottrans", ctx=ctx)
    # no frames / no residues: nothing to do
    assert ca.rmsd_perres(X[:0], off, np.arange(6, dtype=np.int32), ref_frames=X,
                          ctx=ctx).shape == (0, 2)
