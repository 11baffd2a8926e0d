"""Pins the FP64 oracle (oracle/rmsd_oracle.c).

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails:
  * bit-for-bit agreement with the reference's own compiled sources
    (oracle/_ref) on random and degenerate inputs;
  * agreement with an independent textbook SVD-Kabsch (numpy.linalg.svd);
  * closed forms: exact rigid copies give RMSD 0 and recover the rotation;
    mirror images exercise the reflection branch;
  * symmetry and rigid-motion invariance;
  * the analytic gradient against central finite differences.
"""
import numpy as np
import pytest

from cpptraj_amd import synth


def svd_kabsch(tgt, ref, w=None):
    """Textbook weighted Kabsch by SVD: returns (rmsd, U) with U mapping the
    centred target onto the centred reference."""
    w = np.ones(len(tgt)) if w is None else np.asarray(w, float)
    W = w.sum()
    t = tgt - (w[:, None] * tgt).sum(0) / W
    r = ref - (w[:, None] * ref).sum(0) / W
    H = (w[:, None] * t).T @ r
    V, S, Wt = np.linalg.svd(H)
    d = np.sign(np.linalg.det(V @ Wt))
    D = np.diag([1.0, 1.0, d])
    U = (V @ D @ Wt).T
    resid = (w * ((t @ U.T - r) ** 2).sum(1)).sum() / W
    return np.sqrt(max(resid, 0.0)), U


def same(a, b):
    """Bitwise-equal values, NaN matching NaN (degenerate 1-atom fits give NaN
    in the reference too)."""
    return np.array_equal(np.asarray(a, float), np.asarray(b, float), equal_nan=True)


def rand_case(rng, n, spread=5.0, noise=0.5):
    ref = rng.normal(size=(n, 3)) * spread
    Q = synth.rotations(rng, 1)[0]
    tgt = (ref + rng.normal(size=(n, 3)) * noise) @ Q.T + rng.uniform(-20, 20, 3)
    m = synth.masses(rng, n)
    return ref, tgt, m


# -- bit-exact agreement with the compiled reference ------------------------------
@pytest.mark.parametrize("use_mass", [False, True])
def test_rmsd_centered_ref_bitexact(oracle, reflib, use_mass):
    rng = np.random.default_rng(11)
    for _ in range(300):
        n = int(rng.integers(1, 80))
        ref, tgt, m = rand_case(rng, n, noise=float(rng.choice([0.0, 1e-3, 0.5, 5.0])))
        rc_o, c_o = oracle.center_on_origin(ref, m, use_mass)
        rc_r, c_r = reflib.center_on_origin(ref, m, use_mass)
        assert np.array_equal(rc_o, rc_r) and np.array_equal(c_o, c_r)
        a = oracle.rmsd_centered_ref(tgt, rc_o, m, use_mass)
        b = reflib.rmsd_centered_ref(tgt, rc_o, m, m, use_mass)
        assert same(a[0], b[0])
        if a[0] > 0:  # U is left unset on the Jacobi-failure / sentinel paths
            assert np.array_equal(a[1], b[1], equal_nan=True)
        assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])


@pytest.mark.parametrize("use_mass", [False, True])
def test_qcp_and_nofit_bitexact(oracle, reflib, use_mass):
    rng = np.random.default_rng(12)
    for _ in range(300):
        n = int(rng.integers(1, 80))
        ref, tgt, m = rand_case(rng, n, noise=float(rng.choice([0.0, 1e-4, 0.5, 3.0])))
        rc, _ = oracle.center_on_origin(ref, m, use_mass)
        a = oracle.qrmsd_centered_ref(rc, tgt, m, use_mass, want_rot=True)
        b = reflib.qrmsd_centered_ref(rc, tgt, m, use_mass, want_rot=True)
        assert same(a[0], b[0]) and same(a[1], b[1]) and same(a[2], b[2])
        assert same(oracle.rmsd_nofit(tgt, ref, m, use_mass), reflib.rmsd_nofit(tgt, ref, m, use_mass))


def test_jacobi_bitexact_and_eigh(oracle, reflib):
    rng = np.random.default_rng(13)
    for k in range(300):
        A = rng.normal(size=(3, 3))
        S = A @ A.T
        if k % 5 == 0:  # repeated eigenvalues / rank deficiency
            S = np.diag(rng.choice([0.0, 1.0, 2.0], 3))
        e1, E1, ev1 = oracle.jacobi3_sorted(S)
        e2, E2, ev2 = reflib.diagonalize_sort(S)
        assert e1 == e2 and same(E1, E2) and same(ev1, ev2)
        w = np.linalg.eigh(S)[0][::-1]
        # Ties are not reordered by the reference's strict '>' comparisons,
        # so compare the sets; distinct eigenvalues come out descending.
        np.testing.assert_allclose(np.sort(ev1), np.sort(w), atol=1e-10 * max(1.0, abs(w).max()))
        if len(set(np.round(w, 6))) == 3:
            np.testing.assert_allclose(ev1, w, atol=1e-10 * max(1.0, abs(w).max()))
        # eigenvectors are the rows
        for r in range(3):
            np.testing.assert_allclose(S @ E1[r], ev1[r] * E1[r], atol=1e-9 * max(1, abs(w).max()))


def test_jacobi_tie_order_quirk(oracle, reflib):
    # Matrix_3x3::Diagonalize_Sort (src/Matrix_3x3.cpp:229-245) uses strict
    # '>' tests, so a tie for the largest eigenvalue leaves them ascending.
    for diag, order in (((2.0, 2.0, 1.0), (1.0, 2.0, 2.0)), ((3.0, 1.0, 2.0), (3.0, 2.0, 1.0))):
        e, E, ev = oracle.jacobi3_sorted(np.diag(diag))
        assert tuple(ev) == order
        assert same(ev, reflib.diagonalize_sort(np.diag(diag))[2])


def test_rms2d_loop_bitexact(oracle, reflib):
    rng = np.random.default_rng(14)
    X = synth.noisy(5, 40, 50, 0.7).reshape(40, -1)
    sel = np.array(sorted(rng.choice(50, 33, replace=False)), dtype=np.int32)
    m = synth.masses(rng, 33)
    for mode in (0, 1, 2):
        for um in (False, True):
            assert same(oracle.rms2d(X, sel, m, mode, um), reflib.rms2d(X, sel, m, mode, um))
    sel2 = np.arange(17, 50, dtype=np.int32)
    m2 = synth.masses(rng, 33)
    a = oracle.rms2d(X, sel, m, 0, True, ref_sel=sel2, ref_mass=m2)
    b = reflib.rms2d(X, sel, m, 0, True, ref_sel=sel2, ref_mass=m2)
    assert a.shape == (1600,) and same(a, b)


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("use_mass", [False, True])
def test_rms2d_rows_bitexact(oracle, reflib, mode, use_mass):
    """orc_rms2d_rows against the compiled row loop of Calculate_2D
    (src/Analysis_Rms2d.cpp:244-287; cref_rms2d_rows) on random rows, a random
    selection and masses; and against the same rows of the full triangle."""
    rng = np.random.default_rng(40 + mode + 3 * use_mass)
    F, n = 37, 45
    X = synth.noisy(11 + mode, F, n, 0.8).reshape(F, -1)
    sel = np.array(sorted(rng.choice(n, 29, replace=False)), dtype=np.int32)
    m = synth.masses(rng, 29)
    rows = np.array(sorted(rng.choice(F - 1, 9, replace=False)) + [F - 1], dtype=np.int32)
    rng.shuffle(rows)
    a = oracle.rms2d_rows(X, rows, sel, m, mode, use_mass)
    b = reflib.rms2d_rows(X, rows, sel, m, mode, use_mass)
    assert a.shape == b.shape == (sum(F - 1 - int(r) for r in rows),)
    assert same(a, b)
    tri = oracle.rms2d(X, sel, m, mode, use_mass)
    pieces = [tri[oracle.tri_index(F, int(r), int(r) + 1):
                  oracle.tri_index(F, int(r), int(r) + 1) + F - 1 - int(r)]
              if r < F - 1 else np.zeros(0, np.float32) for r in rows]
    assert same(a, np.concatenate(pieces))


@pytest.mark.parametrize("nofit", [False, True])
@pytest.mark.parametrize("use_mass", [False, True])
def test_cluster_cache_bitexact(oracle, reflib, nofit, use_mass):
    """orc_cluster_cache against a cache built pair by pair from the compiled
    Frame::RMSD / Frame::RMSD_NoFit the way MetricArray::calcFrameDistances
    (src/Cluster/MetricArray.cpp:766-801) calls Metric_RMS::FrameDist
    (src/Cluster/Metric_RMS.cpp:44-63): frame framesToCache[f1] is the target,
    framesToCache[f2] the re-centred reference, times the metric weight
    (Dist_Manhattan, :539-547), stored as float (Matrix<float>::setElement)."""
    rng = np.random.default_rng(50 + 2 * nofit + use_mass)
    F, n = 30, 40
    X = synth.noisy(21, F, n, 0.9).reshape(F, -1)
    sel = np.array(sorted(rng.choice(n, 31, replace=False)), dtype=np.int32)
    m = synth.masses(rng, 31)
    fidx = np.arange(1, F, 3, dtype=np.int32)         # a regular sieve, offset 1
    weight = 0.75
    K = len(fidx)
    exp = np.zeros(K * (K - 1) // 2, np.float32)
    sel_xyz = X.reshape(F, n, 3)[:, sel].astype(np.float64)
    for f1 in range(K - 1):
        for f2 in range(f1 + 1, K):
            A, B = sel_xyz[fidx[f1]], sel_xyz[fidx[f2]]
            if nofit:
                v = reflib.rmsd_nofit(A, B, m, use_mass)
            else:
                v = reflib.rmsd(A, B, m, m, use_mass)[0]
            exp[oracle.tri_index(K, f1, f2)] = np.float32(weight * v)
    got = oracle.cluster_cache(X, fidx, sel, m, nofit, use_mass, weight)
    assert same(got, exp)


def test_cluster_cache_identity_list_is_rms2d_triangle(oracle):
    """With every frame cached and weight 1 the cache is the rms2d triangle
    with target and reference swapped: the same RMSD to FP64 rounding, so the
    stored floats differ by at most one unit in the last place."""
    F, n = 33, 50
    X = synth.noisy(23, F, n, 0.6).reshape(F, -1)
    for nofit in (False, True):
        c = oracle.cluster_cache(X, None, nofit=nofit).astype(np.float64)
        t = oracle.rms2d(X, mode=1 if nofit else 0).astype(np.float64)
        ulp = np.spacing(np.maximum(np.abs(c), np.abs(t)).astype(np.float32))
        assert np.all(np.abs(c - t) <= ulp)
        assert np.abs(c - t).max() < 1e-6


def test_rotate_ucell_bitexact(oracle, reflib):
    """orc_rotate_ucell against the compiled Box::SetupFromUcell + RotateUcell
    (src/Box.cpp:571-586, :732-745) on orthorhombic, triclinic and truncated
    octahedral cells and random rotations; the fractional cell the reference
    recomputes is the inverse of the rotated cell."""
    rng = np.random.default_rng(61)
    cells = [[35.26, 41.85, 36.17, 90, 90, 90], [40.0, 40.0, 40.0, 109.4712206, 109.4712206,
                                                 109.4712206], [30.0, 33.0, 37.0, 80, 95, 110]]
    for box in cells:
        ucell, _ = reflib.ucell_from_xyzabg(box)
        for U in synth.rotations(rng, 8):
            a, frac = reflib.rotate_ucell(U, ucell)
            assert same(oracle.rotate_ucell(U, ucell), a)
            assert np.abs(frac @ a.T - np.eye(3)).max() < 1e-12 or \
                np.abs(a.T @ frac - np.eye(3)).max() < 1e-12


def test_min_image_matches_reference(reflib):
    """The tests' minimum-image distance against the compiled
    DIST2_ImageNonOrtho in rotated orthorhombic and in triclinic cells."""
    from conftest import min_image_distance
    rng = np.random.default_rng(62)
    for box in ([35.26, 41.85, 36.17, 90, 90, 90], [30.0, 33.0, 37.0, 80, 95, 110]):
        ucell, _ = reflib.ucell_from_xyzabg(box)
        for U in synth.rotations(rng, 6):
            cell, frac = reflib.rotate_ucell(U, ucell)
            for _ in range(20):
                a, b = rng.uniform(-60, 60, 3), rng.uniform(-60, 60, 3)
                assert abs(min_image_distance(a, b, cell) - reflib.dist_image(a, b, cell, frac)) \
                    < 1e-9


def test_fit_frames_bitexact(oracle, reflib):
    X = synth.noisy(6, 20, 60, 1.0).reshape(20, -1)
    rng = np.random.default_rng(15)
    m = synth.masses(rng, 60)
    for um in (False, True):
        rc, rt = oracle.center_on_origin(X[0].reshape(-1, 3).astype(np.float64), m, um)
        for mod in (0, 1, 2):
            a = oracle.fit_frames(X, rc, rt, mass=m, use_mass=um, modify=mod, want_frames=True)
            b = reflib.fit_frames(X, rc, rt, mass=m, use_mass=um, modify=mod, want_frames=True)
            for x, y in zip(a, b):
                assert same(x, y)


# -- independent textbook and closed forms -----------------------------------------
@pytest.mark.parametrize("use_mass", [False, True])
def test_matches_svd_kabsch(oracle, use_mass):
    rng = np.random.default_rng(16)
    for _ in range(100):
        n = int(rng.integers(5, 200))
        ref, tgt, m = rand_case(rng, n)
        w = m if use_mass else None
        r_svd, U_svd = svd_kabsch(tgt, ref, w)
        rc, _ = oracle.center_on_origin(ref, m, use_mass)
        r, U, _, _ = oracle.rmsd_centered_ref(tgt, rc, m, use_mass)
        assert abs(r - r_svd) < 1e-10
        np.testing.assert_allclose(U, U_svd, atol=1e-9)
        rq, Uq, _ = oracle.qrmsd_centered_ref(rc, tgt, m, use_mass, want_rot=True)
        assert abs(rq - r_svd) < 1e-8
        np.testing.assert_allclose(Uq, U_svd.T, atol=1e-8)  # QCP returns U^T


def test_rigid_copy_is_zero_and_recovers_rotation(oracle):
    rng = np.random.default_rng(17)
    for _ in range(50):
        ref = rng.normal(size=(40, 3)) * 8
        Q = synth.rotations(rng, 1)[0]
        tgt = ref @ Q.T + rng.uniform(-20, 20, 3)
        rc, _ = oracle.center_on_origin(ref)
        r, U, _, _ = oracle.rmsd_centered_ref(tgt, rc)
        assert r < 1e-6
        np.testing.assert_allclose(U, Q.T, atol=1e-9)  # U undoes Q


def test_mirror_image_uses_reflection_branch(oracle):
    rng = np.random.default_rng(18)
    ref = rng.normal(size=(30, 3)) * 5
    tgt = ref * np.array([1.0, 1.0, -1.0])  # improper copy
    rc, _ = oracle.center_on_origin(ref)
    r, U, _, _ = oracle.rmsd_centered_ref(tgt, rc)
    r_svd, U_svd = svd_kabsch(tgt, ref)
    assert r > 0.1 and abs(r - r_svd) < 1e-9
    assert abs(np.linalg.det(U) - 1.0) < 1e-9


def test_symmetry_and_invariance(oracle):
    rng = np.random.default_rng(19)
    ref, tgt, _ = rand_case(rng, 50)
    a = oracle.rmsd_centered_ref(tgt, oracle.center_on_origin(ref)[0])[0]
    b = oracle.rmsd_centered_ref(ref, oracle.center_on_origin(tgt)[0])[0]
    Q = synth.rotations(rng, 1)[0]
    c = oracle.rmsd_centered_ref(tgt @ Q.T + 3.0, oracle.center_on_origin(ref)[0])[0]
    assert abs(a - b) < 1e-12 and abs(a - c) < 1e-12


def test_sentinels(oracle):
    rng = np.random.default_rng(20)
    ref, tgt, _ = rand_case(rng, 10)
    z = np.zeros(10)
    assert oracle.rmsd_centered_ref(tgt, ref, z, True)[0] == -1.0
    assert oracle.rmsd_nofit(tgt, ref, z, True) == -1.0
    # collinear atoms: RMSD is still defined
    line = np.outer(np.arange(8.0), [1.0, 2.0, 0.5])
    rc, _ = oracle.center_on_origin(line)
    Q = synth.rotations(rng, 1)[0]
    assert oracle.rmsd_centered_ref(line @ Q.T, rc)[0] < 1e-6


def test_nofit_closed_form(oracle):
    rng = np.random.default_rng(21)
    ref, tgt, m = rand_case(rng, 33)
    exp = np.sqrt((m * ((ref - tgt) ** 2).sum(1)).sum() / m.sum())
    assert abs(oracle.rmsd_nofit(tgt, ref, m, True) - exp) < 1e-12


def test_tri_index_exhaustive(oracle):
    for n in (2, 3, 7, 64, 65):
        iu = np.triu_indices(n, 1)
        got = [oracle.tri_index(n, i, j) for i, j in zip(*iu)]
        assert got == list(range(n * (n - 1) // 2))
        assert all(oracle.tri_index(n, j, i) == oracle.tri_index(n, i, j) for i, j in zip(*iu))
        assert oracle.tri_index(n, 1, 1) == -1


@pytest.mark.parametrize("use_mass", [False, True])
def test_gradient_central_differences(oracle, use_mass):
    rng = np.random.default_rng(22)
    ref, tgt, m = rand_case(rng, 25)
    rc, _ = oracle.center_on_origin(ref, m, use_mass)
    r, g = oracle.rmsd_grad_target(tgt, rc, m, use_mass)
    h = 1e-4
    fd = np.zeros_like(tgt)
    for a in range(len(tgt)):
        for c in range(3):
            p = tgt.copy(); p[a, c] += h
            q = tgt.copy(); q[a, c] -= h
            fd[a, c] = (oracle.rmsd_centered_ref(p, rc, m, use_mass)[0] -
                        oracle.rmsd_centered_ref(q, rc, m, use_mass)[0]) / (2 * h)
    np.testing.assert_allclose(g, fd, atol=1e-7, rtol=1e-6)
    # translation invariance: the gradient sums to zero
    assert np.abs(g.sum(0)).max() < 1e-12


# -- per-residue no-fit RMSD ('rmsd ... perres') ----------------------------------
def _ragged_residues(rng, natom, nres, maxsize=30):
    """Random residue table over atoms 0..natom-1: contiguous runs plus one
    residue with scattered atoms (a perresmask-style selection)."""
    sizes = rng.integers(1, maxsize, size=nres)
    cuts = np.minimum(np.cumsum(sizes), natom)
    first = np.concatenate([[0], cuts[:-1]])
    atoms = [np.arange(a, b) for a, b in zip(first, cuts) if b > a]
    atoms.append(np.sort(rng.choice(natom, size=7, replace=False)))
    off = np.concatenate([[0], np.cumsum([len(a) for a in atoms])]).astype(np.int32)
    return off, np.concatenate(atoms).astype(np.int32)


@pytest.mark.parametrize("use_mass", [False, True])
@pytest.mark.parametrize("center", [False, True])
@pytest.mark.parametrize("modify", [0, 1, 2])
def test_perres_bitexact(oracle, reflib, use_mass, center, modify):
    rng = np.random.default_rng(13 + 4 * modify + 2 * center + use_mass)
    natom, F = 150, 6
    X = synth.noisy(int(rng.integers(1000)), F, natom, 0.6).reshape(F, -1).astype(np.float32)
    ref = X[0] + rng.normal(size=3 * natom).astype(np.float32)
    off, atoms = _ragged_residues(rng, natom, 12)
    ref_atoms = rng.permutation(natom)[atoms].astype(np.int32)  # a different reference order
    m = synth.masses(rng, natom)
    tm, rm = m[atoms], synth.masses(rng, natom)[ref_atoms]
    U = synth.rotations(rng, F).reshape(F, 9) if modify else None
    tv = rng.uniform(-10, 10, size=(F, 6)) if modify else None
    kw = dict(ref_atoms=ref_atoms, tmass=tm, rmass=rm, use_mass=use_mass, center=center, U=U,
              tvec=tv, modify=modify)
    a = oracle.perres(X, ref, off, atoms, **kw)
    b = reflib.perres(X, ref, off, atoms, **kw)
    assert same(a, b)


@pytest.mark.parametrize("use_mass", [False, True])
def test_perres_closed_forms(oracle, use_mass):
    """One residue that is the fit mask itself: after the fit's update the
    no-fit RMSD against the uncentred reference is the best-fit RMSD (the fit
    superposes the target onto the reference where it lies); with centring
    too.  Translate-only ('norotate') equals the centred no-fit value."""
    rng = np.random.default_rng(21)
    n, F = 60, 8
    X = synth.noisy(5, F, n, 0.7).reshape(F, -1).astype(np.float32)
    m = synth.masses(rng, n)
    ref = X[0].reshape(-1, 3).astype(np.float64)
    rc, rt = oracle.center_on_origin(ref, m, use_mass)
    r, U, tv, _ = oracle.fit_frames(X, rc, rt, mass=m, use_mass=use_mass)
    off = np.array([0, n], np.int32)
    atoms = np.arange(n, dtype=np.int32)
    kw = dict(tmass=m, rmass=m, use_mass=use_mass)
    # Frame 0 is the reference itself: the direct sum gives ~0 where the
    # Kabsch eigenvalue formula bottoms out near sqrt(eps * E0 / M) ~ 1e-7.
    P = oracle.perres(X, X[0], off, atoms, U=U, tvec=tv, modify=1, **kw)[:, 0]
    assert np.abs(P[1:] - r[1:]).max() < 1e-10 and P[0] < 1e-12 and r[0] < 1e-6
    Pc = oracle.perres(X, X[0], off, atoms, U=U, tvec=tv, modify=1, center=True, **kw)[:, 0]
    assert np.abs(Pc[1:] - r[1:]).max() < 1e-10 and Pc[0] < 1e-12
    Pt = oracle.perres(X, X[0], off, atoms, U=U, tvec=tv, modify=2, **kw)[:, 0]
    P0c = oracle.perres(X, X[0], off, atoms, center=True, **kw)[:, 0]
    assert np.abs(Pt - P0c).max() < 1e-10
    # a rigid translation of a residue: |d| without centring, 0 with it
    d = np.array([1.5, -2.0, 0.5])
    Y = (X[:1].reshape(-1, 3) + d).reshape(1, -1).astype(np.float32)
    Q = oracle.perres(Y, X[0], off, atoms, **kw)[0, 0]
    dd = Y[0].reshape(-1, 3).astype(np.float64) - X[0].reshape(-1, 3)
    w = m if use_mass else np.ones(n)
    assert abs(Q - np.sqrt((w * (dd ** 2).sum(1)).sum() / w.sum())) < 1e-12
    assert oracle.perres(Y, X[0], off, atoms, center=True, **kw)[0, 0] < 1e-5
    # zero total mass -> -1 (RMSD_NoFit's divide-by-zero sentinel)
    z = oracle.perres(X, X[0], off, atoms, tmass=np.zeros(n), rmass=m, use_mass=True)
    assert np.all(z == -1.0)


# -- RMS average correlation (rmsavgcorr) -------------------------------------------
@pytest.mark.parametrize("use_mass", [False, True])
@pytest.mark.parametrize("first", [False, True])
@pytest.mark.parametrize("stop_offset", [(-1, 1), (-1, 4), (17, 3), (60, 1)])
def test_avgcorr_bitexact(oracle, reflib, use_mass, first, stop_offset):
    stop, offset = stop_offset
    rng = np.random.default_rng(31 + 2 * use_mass + first)
    F, n = 41, 23
    X = synth.noisy(7, F, n, 0.8).reshape(F, -1).astype(np.float32)
    sel = np.sort(rng.choice(n, 17, replace=False)).astype(np.int32)
    m = synth.masses(rng, len(sel))
    rc = None
    if not first:
        ref = X[5].reshape(-1, 3)[sel].astype(np.float64) + rng.normal(size=(len(sel), 3))
        rc, _ = oracle.center_on_origin(ref, m, use_mass)
    w, ct, sd = oracle.avgcorr(X, sel, m, rc, use_mass, stop, offset)
    ct2, sd2 = reflib.avgcorr(X, w, sel, m, rc, use_mass)
    assert same(ct, ct2) and same(sd, sd2)


def test_avgcorr_windows_and_closed_forms(oracle):
    # window sizes: 1, then 1 + offset ... below min(stop + 1, nframes)
    assert list(oracle.avgcorr_windows(101, -1, 10)) == [1, 11, 21, 31, 41, 51, 61, 71, 81, 91]
    assert list(oracle.avgcorr_windows(10, 3, 1)) == [1, 2, 3]
    assert list(oracle.avgcorr_windows(5, -1, 1)) == [1, 2, 3, 4]
    assert len(oracle.avgcorr_windows(1, -1, 1)) == 0  # not enough frames
    # Frames that are translated copies of one structure: every running
    # average is a translated copy too, so every window's mean and SD are 0.
    rng = np.random.default_rng(3)
    base = rng.normal(size=(12, 3)) * 4
    X = np.stack([(base + rng.integers(-8, 8, size=3)).reshape(-1) for _ in range(20)])
    X = X.astype(np.float32)
    w, ct, sd = oracle.avgcorr(X, ref_sel_centered=oracle.center_on_origin(base)[0])
    assert len(w) == 19 and np.abs(ct).max() < 1e-6 and np.abs(sd).max() < 1e-6
    # 'first' with a constant trajectory: every averaged frame is the first
    w, ct, sd = oracle.avgcorr(np.repeat(X[:1], 9, axis=0))
    assert np.abs(ct).max() < 1e-6


# -- symmetry-corrected RMSD ----------------------------------------------------
def test_hungarian_matches_reference(oracle, reflib):
    # Random costs, and small integers / one-decimal values full of ties that
    # steer the heuristic's tie rules: the same map as the compiled reference.
    rng = np.random.default_rng(61)
    for t in range(6000):
        n = int(rng.integers(1, 11))
        if t % 3 == 0:
            M = rng.uniform(0, 10, (n, n))
        elif t % 3 == 1:
            M = rng.integers(0, 4, (n, n)).astype(float)
        else:
            M = np.round(rng.uniform(0, 3, (n, n)), 1)
        assert np.array_equal(oracle.hungarian(M), reflib.hungarian(M)), (t, M)


def test_hungarian_recovers_planted_assignment(oracle):
    # One zero per row and column on a permutation, positive costs elsewhere:
    # the assignment is that permutation (map[col] = row).
    rng = np.random.default_rng(62)
    for n in range(1, 17):
        perm = rng.permutation(n)
        M = rng.uniform(1.0, 9.0, (n, n))
        M[perm, np.arange(n)] = 0.0
        assert np.array_equal(oracle.hungarian(M), perm)


@pytest.mark.parametrize("fit", [True, False])
@pytest.mark.parametrize("use_mass", [False, True])
def test_symm_matches_reference(oracle, reflib, fit, use_mass):
    X, ref, groups, m, _ = synth.symm_case(63, 30, 120, rigid=fit)
    r = ref.copy()
    if fit:
        r, _ = oracle.center_on_origin(r, m, use_mass)
    a = oracle.symm(X, None, m, r, use_mass, fit, groups)
    b = reflib.symm(X, None, m, r, use_mass, fit, groups)
    for x, y in zip(a, b):
        assert same(x, y)


def test_symm_undoes_planted_permutations(oracle):
    # Frames = rigid motions of the reference with atoms permuted inside their
    # groups: the map is the inverse permutation and the RMSD that of the
    # un-permuted frame (0 without noise, to the Kabsch floor).
    for sigma in (0.0, 0.05):
        X, ref, groups, m, perms = synth.symm_case(64, 12, 90, sigma=sigma)
        rc, _ = oracle.center_on_origin(ref)
        rm, amap, _, _ = oracle.symm(X, None, None, rc, False, True, groups)
        inv = np.argsort(perms, axis=1)
        assert np.array_equal(amap, inv)
        fixed = np.stack([X[f].reshape(-1, 3)[inv[f]] for f in range(len(X))])
        r0, _, _, _ = oracle.symm(fixed.reshape(len(X), -1), None, None, rc, False, True, [])
        assert np.abs(rm - r0).max() <= 1e-12
        if sigma == 0.0:
            # Kabsch's eigenvalue formula floor, ~sqrt(eps E0 / M)
            assert rm.max() < 5e-6


def test_symm_without_groups_is_the_plain_fit(oracle):
    X, ref, _, m, _ = synth.symm_case(65, 8, 50)
    rc, rt = oracle.center_on_origin(ref, m, True)
    rm, amap, U, tr = oracle.symm(X, None, m, rc, True, True, [])
    r2, U2, tv2, _ = oracle.fit_frames(X, rc, rt, mass=m, use_mass=True, modify=0,
                                       want_frames=True)
    assert same(rm, r2) and same(U, U2.reshape(-1, 9)) and same(tr, tv2[:, :3])
    assert np.array_equal(amap, np.tile(np.arange(50), (8, 1)))


@pytest.mark.parametrize("use_mass", [False, True])
def test_fit_pairs_shim_is_the_per_pair_fit(reflib, use_mass):
    """cref_fit_pairs (the 'reftraj' CPU baseline) equals, bit for bit, the
    compiled fit of each pair on its own: the reference frame centred with its
    masses (Frame::CenterOnOrigin), then RMSD_CenteredRef."""
    from cpptraj_amd import synth
    F, n = 12, 60
    X = synth.noisy(5, F, n, 1.0).reshape(F, -1).astype(np.float32)
    Y = synth.noisy(6, F, n, 1.0).reshape(F, -1).astype(np.float32)
    sel = np.arange(0, n, 2, dtype=np.int32)
    m = synth.masses(np.random.default_rng(1), len(sel))
    got = reflib.fit_pairs(X, Y, sel, m, use_mass)
    exp = []
    for f in range(F):
        rc, rt = reflib.center_on_origin(Y[f].reshape(n, 3).astype(np.float64)[sel], m, use_mass)
        r, _, _, _ = reflib.fit_frames(X[f:f + 1], rc, rt, sel=sel, mass=m, use_mass=use_mass)
        exp.append(r[0])
    assert np.array_equal(got, np.array(exp))
