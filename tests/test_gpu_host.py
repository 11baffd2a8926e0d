"""The C++ host layer (Analysis_Rms2d / Action_Rmsd / cluster cache classes)
driven with the commands of the reference's own tests on the same frames:
test/Test_2DRMS/RunTest.sh and test/Test_RMSD/RunTest.sh.  Their DoTest
compares outputs with plain diff, so the text files written here must equal
the expected .save files byte for byte (vecs.dat: |diff| <= 1e-6, the
tolerance its DoTest passes).  Fitted coordinates are compared with the
.crd.save values (written %8.3f) to half a unit of the last digit."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, min_image_distance, tri_to_square, ucell_from_box

pytestmark = pytest.mark.gpu
TXT = os.path.join(GOLDEN, "text")


def expected(name):
    with open(os.path.join(TXT, name)) as fh:
        return fh.read()


def text(path):
    with open(path) as fh:
        return fh.read()


@pytest.fixture
def hs():
    from cpptraj_amd.host import HostSession
    s = HostSession(0)
    yield s
    s.close()


def load_tz2(hs, golden):
    g = golden("tz2_2drms")
    hs.add_coords("tz2", g["frames"], 223, g["mass"])
    for expr, key in ((":3-7", "sel_3_7"), (":2", "sel_2"), (":11", "sel_11"),
                      (":2-12@CA", "sel_2_12_ca")):
        hs.add_mask(expr, g[key])
    return g


# Test_2DRMS tests 1, 2, 5 and 8.
@pytest.mark.parametrize("cmd,save", [
    ("2drms crd1 :3-7 rmsout {out}", "Test_2DRMS__rmsd.dat.save"),
    ("2drms crd1 :3-7 rmsout {out} mass", "Test_2DRMS__rmsd.mass.dat.save"),
    ("2drms :2 :11 out {out}", "Test_2DRMS__trp.dat.save"),
    ("2drms crd1 :3-7 rmsout {out} qrmsd", "Test_2DRMS__rmsd.dat.save"),
])
def test_2drms_outputs(hs, golden, tmp_path, cmd, save):
    load_tz2(hs, golden)
    out = str(tmp_path / "out.dat")
    hs.run(cmd.format(out=out))
    assert text(out) == expected(save)


# Test_2DRMS test 3: the same frames as a reference trajectory give the full
# matrix, printed like the triangle.
def test_2drms_reftraj(hs, golden, tmp_path):
    g = load_tz2(hs, golden)
    hs.add_coords("ref", g["frames"].copy(), 223, g["mass"])
    out = str(tmp_path / "rmsd1.dat")
    hs.run(f"2drms crdset tz2 :3-7 rmsout {out} reftraj ref")
    assert text(out) == expected("Test_2DRMS__rmsd.dat.save")


# Test_2DRMS test 6: 'rms first :2-12@CA' rewrites the frames, then a no-fit matrix.
def test_2drms_nofit_after_fit(hs, golden, tmp_path):
    load_tz2(hs, golden)
    out = str(tmp_path / "nofit.dat")
    hs.run("rms first :2-12@CA")
    hs.run(f"2drms crd1 :2 nofit out {out}")
    assert text(out) == expected("Test_2DRMS__nofit.dat.save")


# Test_2DRMS test 7: nine identical frames, XYZ ('nosquare2d') output.
def test_2drms_nosquare2d(hs, golden, tmp_path):
    g = golden("rgacc")
    fr = np.tile(g["frame"][None], (9, 1))
    hs.add_coords("rgacc", fr, fr.shape[1] // 3)
    out = str(tmp_path / "rgacc.dat")
    hs.run(f"2drms crd1 * out {out} nosquare2d")
    assert text(out) == expected("Test_2DRMS__rgacc.dat.save")


# Test_RMSD 'Basic RMSD tests': three actions over the same frames in order.
def test_rms_basic(hs, golden, tmp_path):
    g = golden("tz2_rmsd")
    n = g["frames"].shape[1] // 3
    hs.add_coords("trunc", g["frames"], n, g["mass"])
    hs.add_coords("raw", g["frames"].copy(), n, g["mass"])
    hs.add_mask(":2-11", np.arange(n))
    d = str(tmp_path)
    hs.run(f"rms Res2-11 first :2-11 out {d}/rmsd.dat crdset trunc")
    hs.run(f"rms Res2-11_mass first :2-11 out {d}/rmsd.mass.dat mass savevectors combined "
           f"vecsout {d}/vecs.dat crdset trunc")
    hs.run(f"rms Res2_11_traj reftraj raw :2-11 out {d}/rmsd.reftraj.dat crdset trunc")
    assert text(f"{d}/rmsd.dat") == expected("Test_RMSD__rmsd.dat.save")
    assert text(f"{d}/rmsd.mass.dat") == expected("Test_RMSD__rmsd.mass.dat.save")
    assert text(f"{d}/rmsd.reftraj.dat") == expected("Test_RMSD__rmsd.reftraj.dat.save")
    got, exp = text(f"{d}/vecs.dat").splitlines(), expected("Test_RMSD__vecs.dat.save").splitlines()
    assert got[0] == exp[0] and len(got) == len(exp)
    a = np.array([[float(t) for t in l.split()] for l in got[1:]])
    b = np.array([[float(t) for t in l.split()] for l in exp[1:]])
    assert np.abs(a - b).max() <= 1e-6


# Test_RMSD 'RMS coordinate rotation/rotation matrices test'.
def test_rms_rotation(hs, golden, tmp_path):
    g = golden("tz2_rotate")
    hs.add_reference("[first]", g["ref_frame"], 223)
    hs.add_coords("WAT", g["frames"], 220)
    hs.add_mask("@CA", g["sel_ca_ref"], set_name="[first]")
    hs.add_mask("@CA", g["sel_ca"], set_name="WAT")
    d = str(tmp_path)
    hs.run("rms NOROT ref [first] norotate @CA crdset WAT")
    nr = hs.coords("WAT").reshape(10, 220, 3)
    assert np.abs(nr - g["norotate"]).max() <= 5e-4 + 1e-6
    # The reference keeps frames in FP64 between actions, so its ROT fits the
    # NOROT output before any float rounding; the same fit of the original
    # frames differs from it only by that translation.  A COORDS set stores
    # FP32, so ROT runs on a fresh copy of the original frames here.
    hs.add_coords("WAT2", g["frames"], 220)
    hs.add_mask("@CA", g["sel_ca"], set_name="WAT2")
    hs.run(f"rms ROT ref [first] out {d}/rms.dat @CA savematrices matricesout {d}/rmatrices.dat "
           "crdset WAT2")
    # The GPU solves the fit with QCP, the reference with Jacobi: U agrees to
    # 5e-10, so a 9th decimal may round the other way.
    got, exp = text(f"{d}/rmatrices.dat").splitlines(), \
        expected("Test_RMSD__rmatrices.dat.save").splitlines()
    assert got[0] == exp[0] and len(got) == len(exp)
    a = np.array([[float(t) for t in l.split()] for l in got[1:]])
    b = np.array([[float(t) for t in l.split()] for l in exp[1:]])
    assert np.array_equal(a[:, 0], b[:, 0])
    assert np.abs(a - b).max() <= 1.5e-9
    rot = hs.coords("WAT2").reshape(10, 220, 3)
    assert np.abs(rot - g["rotate"]).max() <= 5e-4 + 1e-6


def ortho_distances(X, box, g):
    """(no image, image) distances between the centres of mass of the fixture's
    two residues, per frame, in the given cells."""
    m = g["mass"]
    Y = np.asarray(X, np.float64).reshape(len(X), -1, 3)
    ra, rb = g["res_a"], g["res_b"]
    ca = (Y[:, ra] * m[ra, None]).sum(1) / m[ra].sum()
    cb = (Y[:, rb] * m[rb, None]).sum(1) / m[rb].sum()
    return (np.linalg.norm(cb - ca, axis=1),
            np.array([min_image_distance(ca[f], cb[f], box[f]) for f in range(len(X))]))


@pytest.mark.parametrize("cmd", ["rms :1-13 first", "align :1-13", "rms :1-13 first chunk 3"])
def test_rms_rotates_the_unit_cell(hs, golden, oracle, cmd):
    """Test_RMSD 'imaging after unit cell rotation': 'rms :1-13 first' on a
    COORDS set with a box rotates coordinates and unit cells (Box::RotateUcell,
    src/Action_Rmsd.cpp:381; align: src/Action_Align.cpp:98); the distances of
    distances.agr.save then follow, and the cells equal the oracle's whole
    chain (its Jacobi U differs from the GPU's QCP U by ~5e-10) to 1e-7."""
    g = golden("tz2_ortho")
    X, nfit = g["frames"], int(g["nfit"])
    F, n = X.shape[0], X.shape[1] // 3
    ucell = np.stack([ucell_from_box(b) for b in g["box"]])
    hs.add_coords("ortho", X, n, g["mass"])
    hs.set_box("ortho", ucell)
    hs.add_mask(":1-13", np.arange(nfit))
    hs.run(cmd)
    moved = hs.coords("ortho").reshape(F, n, 3)
    box = hs.box("ortho")
    noimage, image = ortho_distances(moved, box, g)
    HALF4 = 0.00005 + 1e-7
    assert np.abs(noimage - g["dist_noimage"]).max() <= HALF4
    assert np.abs(image - g["dist_image"]).max() <= HALF4
    ref = X[0].reshape(-1, 3)[:nfit].astype(np.float64)
    rc, rt = oracle.center_on_origin(ref)
    _, U_o, _, _ = oracle.fit_frames(X, rc, rt, sel=np.arange(nfit, dtype=np.int32))
    exp = np.stack([oracle.rotate_ucell(U_o[f], ucell[f]) for f in range(F)])
    assert np.abs(box - exp).max() <= 1e-7
    assert not np.array_equal(box, ucell)


def test_rotate_ucell_matches_oracle_on_the_fit_rotations(ctx, golden, oracle):
    """rmsd_rotate_ucell against orc_rotate_ucell applied to the same U
    (the fit's own rotations): the reference's operation order, <= 1e-12."""
    import torch
    import cpptraj_amd as ca
    g = golden("tz2_ortho")
    X, nfit = g["frames"], int(g["nfit"])
    F = X.shape[0]
    Xd = torch.as_tensor(X).cuda()
    sel = torch.arange(nfit, dtype=torch.int32, device="cuda")
    _, U, _ = ca.rmsd_fit(Xd, mask=sel, modify="rottrans", save_matrices=True, ctx=ctx)
    ucell = np.stack([ucell_from_box(b) for b in g["box"]])
    u = torch.as_tensor(ucell).cuda()
    ca.rotate_ucell(U, u, ctx=ctx)
    Uh = U.cpu().numpy()
    exp = np.stack([oracle.rotate_ucell(Uh[f], ucell[f]) for f in range(F)])
    assert np.abs(u.cpu().numpy() - exp).max() <= 1e-12


# Test_RMSD 'nomod' and 'previous'.
@pytest.mark.parametrize("cmd,save", [
    ("rms First_CA :2-12@CA out {out} nomod", "Test_RMSD__NoMod.dat.save"),
    ("rms ToPrevious :2-12@CA previous out {out}", "Test_RMSD__Previous.dat.save"),
])
def test_rms_nomod_previous(hs, golden, tmp_path, cmd, save):
    g = golden("tz2_camask")
    n = g["frames"].shape[1] // 3
    hs.add_coords("tz2", g["frames"], n, g["mass"])
    hs.add_mask(":2-12@CA", np.arange(n))
    out = str(tmp_path / "o.dat")
    hs.run(cmd.format(out=out))
    assert text(out) == expected(save)


def test_rms2d_corr(hs, golden, tmp_path):
    load_tz2(hs, golden)
    out, corr = str(tmp_path / "m.dat"), str(tmp_path / "corr.dat")
    hs.run(f"2drms :3-7 out {out} corr {corr} M")
    tri = hs.matrix("M").astype(np.float64)
    ct = hs.series("M[Corr]")
    R = tri_to_square(tri, 10)
    exp = [1.0] + [np.mean([np.exp(-R[j, j + i]) for j in range(10 - i)]) for i in range(1, 10)]
    assert np.allclose(ct, exp, rtol=1e-13, atol=0)
    assert len(text(corr).splitlines()) == 11


def test_cluster_cache_cmatrix(hs, golden, oracle, tmp_path):
    from cpptraj_amd.host import read_cmatrix
    g = load_tz2(hs, golden)
    path = str(tmp_path / "pd.cmatrix")
    hs.run(f"cluster crdset tz2 rms :3-7 mass sieve 3 pairdist {path} savepairdist")
    tri, actual, sieve, frames = read_cmatrix(path)
    assert actual == 10 and sieve == 3 and list(frames) == [0, 3, 6, 9]
    exp = oracle.cluster_cache(g["frames"], np.array([0, 3, 6, 9], np.int32), g["sel_3_7"],
                               g["mass"][g["sel_3_7"]], use_mass=True)
    assert np.abs(tri - exp).max() <= 1e-6
    raw = open(path, "rb").read()
    assert raw[:4] == b"CTM\x02"
    assert np.frombuffer(raw[4:28], dtype="<u8")[:2].tolist() == [10, 4]
    assert np.frombuffer(raw[20:28], dtype="<i8")[0] == 3
    assert raw[28 + 4 * 6:] == b"FTTFTTFTTF"


def test_cluster_cache_loaded_from_file_or_set(hs, golden, oracle, tmp_path):
    """MetricArray::setupPairwiseCache / CacheDistances: 'pairdist <name>
    pairdistfile <file>' reads the file when no cache set of that name exists
    (the planted values come back unchanged: nothing is recomputed); a cache
    whose frames cover the frames to cluster is used as it is (a sieve-1 cache
    for a sieve-3 run); a cache that does not cover them is an error, or is
    recomputed with 'pwrecalc' (and rewritten with 'pwrecalc savepairdist');
    a 'pairdist' file that does not exist is computed and saved; a cache set
    already in the session is used again."""
    from cpptraj_amd.host import RmsdError, read_cmatrix, write_cmatrix
    g = load_tz2(hs, golden)
    sel, m = g["sel_3_7"], g["mass"][g["sel_3_7"]]
    path = str(tmp_path / "planted.cmatrix")
    planted = np.arange(6, dtype=np.float32) * 0.25 + 1.0
    write_cmatrix(path, planted, 4, 10, 3, np.array([0, 3, 6, 9], np.int32))
    hs.run(f"cluster crdset tz2 rms :3-7 mass sieve 3 pairdist A pairdistfile {path}")
    assert np.array_equal(hs.matrix("A"), planted)
    # sieve 2 needs frames 2, 4, 8, which the file does not hold
    with pytest.raises(RmsdError, match="do not match"):
        hs.run(f"cluster crdset tz2 rms :3-7 mass sieve 2 pairdist B pairdistfile {path}")
    hs.run(f"cluster crdset tz2 rms :3-7 mass sieve 2 pairdist C pairdistfile {path} pwrecalc")
    exp2 = oracle.cluster_cache(g["frames"], np.arange(0, 10, 2, dtype=np.int32), sel, m,
                                use_mass=True)
    assert np.abs(hs.matrix("C") - exp2).max() <= 1e-6
    assert np.array_equal(read_cmatrix(path)[0], planted)   # loaded: not written back
    hs.run(f"cluster crdset tz2 rms :3-7 mass sieve 2 pairdist D pairdistfile {path} "
           "pwrecalc savepairdist")
    tri, actual, sieve, frames = read_cmatrix(path)
    assert (actual, sieve, list(frames)) == (10, 2, [0, 2, 4, 6, 8])
    assert np.array_equal(tri, hs.matrix("D"))
    # a sieve-1 cache holds every frame of a sieve-3 run: used as it is
    full = str(tmp_path / "full.cmatrix")
    hs.run(f"cluster crdset tz2 rms :3-7 mass pairdist {full}")   # not there: computed, saved
    exp1 = oracle.cluster_cache(g["frames"], np.arange(10, dtype=np.int32), sel, m, use_mass=True)
    tri, actual, sieve, frames = read_cmatrix(full)
    assert (actual, sieve, len(tri)) == (10, 1, 45)
    assert np.abs(tri - exp1).max() <= 1e-6
    write_cmatrix(full, planted[:1].repeat(45), 10, 10, 1, np.arange(10, dtype=np.int32))
    hs.run(f"cluster crdset tz2 rms :3-7 mass sieve 3 loadpairdist pairdist E pairdistfile {full}")
    assert np.array_equal(hs.matrix("E"), planted[:1].repeat(45))
    # the set 'A' of the session is used before any file of that name
    hs.run("cluster crdset tz2 rms :3-7 mass sieve 3 loadpairdist pairdist A")
    assert np.array_equal(hs.matrix("A"), planted)


@pytest.mark.parametrize("extra", ["", "mass savematrices savevectors separate", "norotate"])
def test_rms_streamed_chunks_match_one_batch(hs, extra):
    """The host layer's double-buffered streaming (chunks on two transfer
    slots) gives the same RMSD, matrices, vectors and coordinates as a single
    chunk: 1000 frames x 300 atoms in chunks of 64 frames (16 chunks, odd tail)."""
    from cpptraj_amd import synth
    X = synth.noisy(17, 1000, 300, 1.0).reshape(1000, -1)
    m = synth.masses(np.random.default_rng(17), 300)
    hs.add_coords("a", X, 300, m)
    hs.add_coords("b", X.copy(), 300, m)
    hs.add_mask(":1-20", np.arange(0, 300, 2))
    hs.run(f"rms A first :1-20 {extra} crdset a")
    hs.run(f"rms B first :1-20 {extra} chunk 64 crdset b")
    assert np.array_equal(hs.series("A"), hs.series("B"))
    if "savematrices" in extra:
        assert np.array_equal(hs.series("A[RM]"), hs.series("B[RM]"))
        assert np.array_equal(hs.series("A[TV]"), hs.series("B[TV]"))
    assert np.array_equal(hs.coords("a"), hs.coords("b"))


def test_align_move_mask(hs):
    """'align <mask> move <mask>' (Action_Align): the moved atoms get exactly the
    coordinates 'rms' gives them, the others stay where they were; without
    'move' align equals rms on every atom."""
    from cpptraj_amd import synth
    X = synth.noisy(23, 50, 120, 1.0).reshape(50, -1)
    for name in ("r", "a", "m"):
        hs.add_coords(name, X.copy(), 120)
    hs.add_mask(":1-5", np.arange(0, 60))
    hs.add_mask(":6-10", np.arange(60, 120))
    hs.run("rms R first :1-5 crdset r")
    hs.run("align :1-5 crdset a")
    hs.run("align :1-5 move :6-10 chunk 16 crdset m")
    r = hs.coords("r").reshape(50, 120, 3)
    a = hs.coords("a").reshape(50, 120, 3)
    m = hs.coords("m").reshape(50, 120, 3)
    assert np.array_equal(a, r)
    assert np.array_equal(m[:, 60:], r[:, 60:])
    assert np.array_equal(m[:, :60], X.reshape(50, 120, 3)[:, :60])


# -- Test_PerResRMSD (rmsd ... perres) ---------------------------------------------
def load_perres(hs, golden, frames=None):
    """tz2.truncoct solute frames as COORDS 'trunc' and frame 1 as reference
    'tz2.truncoct.nc' (refindex 0), both with residues; masks :2-11."""
    g = golden("tz2_perres")
    X = g["frames"] if frames is None else g["frames"][frames]
    n = g["frames"].shape[1] // 3
    names = [str(r) for r in g["res_names"]]
    hs.add_coords("trunc", X, n, g["mass"])
    hs.set_residues("trunc", g["res_first"], names)
    hs.add_reference("tz2.truncoct.nc", g["frames"][0], n, g["mass"])
    hs.set_residues("tz2.truncoct.nc", g["res_first"], names)
    hs.add_mask(":2-11", np.arange(g["res_first"][1], g["res_first"][11]))
    return g


def test_perres_grace_output(hs, golden, tmp_path):
    load_perres(hs, golden)
    out = str(tmp_path / "PerResRMSD.agr")
    hs.run(f"rmsd :2-11 refindex 0 perres perresout {out} crdset trunc")
    assert text(out) == expected("Test_PerResRMSD__PerResRMSD.agr.save")


def test_perres_average(hs, golden, tmp_path):
    # 'trajin ../tz2.truncoct.nc 2 4': frames 2-4 only.
    load_perres(hs, golden, frames=slice(1, 4))
    out = str(tmp_path / "perresavg.dat")
    hs.run(f"rmsd R2-11 :2-11 refindex 0 perres perresavg {out} crdset trunc")
    assert text(out) == expected("Test_PerResRMSD__perresavg.dat.save")


def test_perres_centered_range(hs, golden, tmp_path):
    g = load_perres(hs, golden)
    out = str(tmp_path / "center.agr")
    hs.run(f"rmsd :2-11 refindex 0 perres perresout {out} range 1 perrescenter crdset trunc")
    assert text(out) == expected("Test_PerResRMSD__center.agr.save")
    # The fit still moved every frame; the per-residue set is a data set too.
    assert np.abs(hs.series("RMSD_00000[res]:1") - g["center"][0]).max() <= 0.00005 + 1e-7


def test_perres_mask_ranges_and_nofit(hs, golden, oracle, tmp_path):
    """perresmask (heavy atoms picked by the caller's mask table), refrange
    onto other residues of the reference, nofit, and residues with atom
    count mismatches (kept as empty sets), against the oracle."""
    g = load_perres(hs, golden)
    X = g["frames"]
    n = X.shape[1] // 3
    first = np.append(g["res_first"], n)
    heavy = np.arange(0, n, 2)  # a stand-in selection for '&!@H='
    hs.add_mask("!@H=", heavy)
    hs.run("rmsd P :2-11 refindex 0 nofit perres range 2-4,6 refrange 3-6 perresmask !@H= "
           "perrescenter mass crdset trunc")
    pick = np.zeros(n, bool)
    pick[heavy] = True
    for t, r in zip([2, 3, 4, 6], [3, 4, 5, 6]):
        ta = np.arange(first[t - 1], first[t])[pick[first[t - 1]:first[t]]]
        ra = np.arange(first[r - 1], first[r])[pick[first[r - 1]:first[r]]]
        got = hs.series(f"P[res]:{t}")
        if len(ta) != len(ra) or len(ta) == 0:
            assert got.size == 0
            continue
        exp = oracle.perres(X, X[0], [0, len(ta)], ta, ref_atoms=ra, tmass=g["mass"][ta],
                            rmass=g["mass"][ra], use_mass=True, center=True)[:, 0]
        assert np.abs(got - exp).max() <= 1e-12


# -- Test_RmsAvgCorr (rmsavgcorr) ---------------------------------------------------
@pytest.mark.parametrize("cmd,save", [
    ("rmsavgcorr crdset ca out {out} R2-12 ref avg.CA.rst7", "Test_RmsAvgCorr__rmscorr.dat.save"),
    ("rmsavgcorr crdset ca out {out} Off10 offset 10 reference",
     "Test_RmsAvgCorr__rmscorr.10.dat.save"),
    ("rmsavgcorr crdset ca out {out} First first", "Test_RmsAvgCorr__rmscorr.first.dat.save"),
])
def test_rmsavgcorr_outputs(hs, golden, tmp_path, cmd, save):
    """The frames are the 101 tz2 frames stripped to :2-12@CA (11 atoms); the
    reference is avg.CA.rst7 on the stripped topology."""
    c, a = golden("tz2_camask"), golden("tz2_avgcorr")
    hs.add_coords("ca", c["frames"], 11, c["mass"])
    hs.add_reference("avg.CA.rst7", a["ref"].reshape(-1), 11)
    out = str(tmp_path / "rmscorr.dat")
    hs.run(cmd.format(out=out))
    assert text(out) == expected(save)


def test_rmsavgcorr_stop_mass_and_output_file(hs, golden, oracle, tmp_path):
    """'stop', 'mass', a mask and the 'output' file (window, mean, SD)."""
    c = golden("tz2_camask")
    X = c["frames"]
    hs.add_coords("ca", X, 11, c["mass"])
    sel = np.array([0, 2, 3, 5, 7, 8, 10], np.int32)
    hs.add_mask("@CA&!:5", sel)
    sep = str(tmp_path / "sep.txt")
    hs.run(f"rmsavgcorr crdset ca S @CA&!:5 mass stop 30 offset 4 first output {sep}")
    w, ct, sd = oracle.avgcorr(X, sel, c["mass"][sel], None, True, 30, 4)
    got = hs.series("S")
    assert got.shape == ct.shape and np.abs(got - ct).max() <= 5e-7
    assert np.abs(hs.series("S[SD]") - sd).max() <= 5e-7
    lines = text(sep).splitlines()
    assert [int(l.split()[0]) for l in lines] == list(w)
    assert all(abs(float(l.split()[1]) - v) <= 5e-7 for l, v in zip(lines, ct))


# Test_SymmRmsd through the host layer: 'symmrmsd' with the topology's
# symmetric groups (fixture data), reference modes and 'remap'.
def _topology_groups(d, sel=None):
    off, flat = d["goff"], d["gatoms"]
    g = [flat[off[k]:off[k + 1]] for k in range(len(off) - 1)]
    return [(np.asarray(sel)[x] if sel is not None else x).tolist() for x in g]


def test_symmrmsd_reference_and_output(hs, golden, tmp_path):
    d = golden("symm_afv")
    n = len(d["mass"])
    hs.add_coords("afv", d["rotated"].reshape(3, -1).astype(np.float32), n, d["mass"])
    hs.add_reference("AFV.rst7", d["ref"].reshape(-1).astype(np.float32), n, d["mass"])
    hs.set_symmetric_groups("afv", _topology_groups(d, d["sel"]))
    hs.add_mask("!@H=", d["sel"])
    out = str(tmp_path / "rmsd.dat")
    hs.run(f"symmrmsd crdset afv Symm !@H= reference out {out}")
    assert np.abs(hs.series("Symm") - d["expect_rotated"][:, 1]).max() <= 0.00005 + 1e-7
    lines = [l.split() for l in text(out).splitlines()]
    assert lines[0] == ["#Frame", "Symm"]
    assert [l[1] for l in lines[1:]] == ["%.4f" % v for v in d["expect_rotated"][:, 1]]


@pytest.mark.parametrize("res", ["asp", "tyr"])
def test_symmrmsd_first_remap_coordinates(hs, golden, res):
    d = golden(f"symm_{res}")
    fr = d["frames"]
    F, n = fr.shape[:2]
    hs.add_coords(res, fr.reshape(F, -1), n, d["mass"])
    hs.set_symmetric_groups(res, _topology_groups(d))
    hs.run(f"symmrmsd crdset {res} first SYMM remap")
    assert np.abs(hs.series("SYMM") - d["expect"][:, 1]).max() <= 0.00005 + 1e-7
    got = hs.coords(res).reshape(F, n, 3)
    assert np.abs(got - d["remap"]).max() <= 0.0005 + 1e-6


def test_symmrmsd_previous_and_nofit(hs, golden, oracle):
    # 'previous': each frame against the previous one as the action left it;
    # 'nofit' leaves the frames untouched.
    d = golden("symm_afv")
    traj = d["traj"][:20]
    F, n = traj.shape[:2]
    hs.add_coords("t", traj.reshape(F, -1), n, d["mass"])
    hs.set_symmetric_groups("t", _topology_groups(d, d["sel"]))
    hs.add_mask("!@H=", d["sel"])
    hs.run("symmrmsd crdset t P !@H= nofit previous")
    sel = d["sel"]
    groups = [g for g in _topology_groups(d)]
    exp = []
    for f in range(F):
        prev = traj[max(f - 1, 0)].astype(np.float64)[sel]
        r, _, _, _ = oracle.symm(traj[f:f + 1].reshape(1, -1), sel, None, prev, False, False,
                                 groups)
        exp.append(r[0])
    assert np.abs(hs.series("P") - np.array(exp)).max() <= 1e-12
    assert np.array_equal(hs.coords("t").reshape(F, n, 3), traj)


def test_rms2d_srmsd_matrix(hs, golden):
    # Test_Cluster_SymmRMSD's '2drms !@H= out 2drms.gnu srmsd' (150 frames).
    from conftest import tri_to_square
    d = golden("symm_afv")
    traj = d["traj"]
    F, n = traj.shape[:2]
    hs.add_coords("afv", traj.reshape(F, -1), n, d["mass"])
    hs.set_symmetric_groups("afv", _topology_groups(d, d["sel"]))
    hs.add_mask("!@H=", d["sel"])
    hs.run("rms2d crdset afv S2 !@H= srmsd")
    M = tri_to_square(hs.matrix("S2"), F)
    assert np.abs(M - d["expect_2drms"]).max() <= 0.0005 + 1e-6


def test_cluster_cache_srmsd_metric(hs, golden, oracle):
    # 'cluster ... srmsd !@H= sieve 3': element (i, j) of the cache is
    # Metric_SRMSD::FrameDist(f_i, f_j), frame f_i against f_j as the reference.
    d = golden("symm_afv")
    traj = d["traj"][:60]
    F, n = traj.shape[:2]
    hs.add_coords("afv", traj.reshape(F, -1), n, d["mass"])
    hs.set_symmetric_groups("afv", _topology_groups(d, d["sel"]))
    hs.add_mask("!@H=", d["sel"])
    hs.run("cluster crdset afv C srmsd !@H= sieve 3")
    got = hs.matrix("C")
    frames = list(range(0, F, 3))
    sel, groups = d["sel"], _topology_groups(d)
    k = 0
    for a in range(len(frames)):
        for b in range(a + 1, len(frames)):
            r, _ = oracle.center_on_origin(traj[frames[b]].astype(np.float64)[sel])
            e, _, _, _ = oracle.symm(traj[frames[a]].reshape(1, -1), sel, None, r, False, True,
                                     groups)
            assert abs(got[k] - np.float32(e[0])) <= 1e-6, (a, b)
            k += 1
    assert k == len(got)


def test_symmrmsd_reftraj_and_mass(hs, golden, oracle):
    # 'reftraj': frame f against frame f of another COORDS set (centred with
    # masses for the fit); 'mass' weights both fits.
    d = golden("symm_afv")
    traj = d["traj"][:30]
    F, n = traj.shape[:2]
    rng = np.random.default_rng(90)
    rtraj = (traj[::-1] + rng.normal(size=traj.shape) * 0.2).astype(np.float32)
    hs.add_coords("t", traj.reshape(F, -1), n, d["mass"])
    hs.add_coords("r", rtraj.reshape(F, -1), n, d["mass"])
    hs.set_symmetric_groups("t", _topology_groups(d, d["sel"]))
    hs.add_mask("!@H=", d["sel"])
    hs.run("symmrmsd crdset t RT !@H= reftraj r mass")
    sel, groups, m = d["sel"], _topology_groups(d), d["mass"][d["sel"]]
    exp = []
    for f in range(F):
        rc, _ = oracle.center_on_origin(rtraj[f].astype(np.float64)[sel], m, True)
        e, _, _, _ = oracle.symm(traj[f:f + 1].reshape(1, -1), sel, m, rc, True, True, groups)
        exp.append(e[0])
    assert np.abs(hs.series("RT") - np.array(exp)).max() <= 1e-9
