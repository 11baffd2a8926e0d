"""The host layer's writers, without a GPU: values parsed from the reference's
expected outputs, written back by the DataIO_Std restatement, must give the
same text byte for byte (square2d, nosquare2d, 1D series of one and several
components); the Cmatrix binary round-trips and has the reference's layout
(src/Cluster/Cmatrix_Binary.cpp: 'CTM' 2, uint64 frames, uint64 rows, int64
sieve, float triangle, 'T'/'F' per frame when sieved)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

TXT = os.path.join(GOLDEN, "text")


def read(name):
    with open(os.path.join(TXT, name)) as fh:
        return fh.read()


def rows(text):
    return np.array([[float(t) for t in l.split()] for l in text.splitlines()[1:]])


@pytest.mark.parametrize("save", ["Test_2DRMS__rmsd.dat.save", "Test_2DRMS__trp.dat.save",
                                  "Test_2DRMS__nofit.dat.save"])
def test_square2d_roundtrip(tmp_path, save):
    from cpptraj_amd.host import write_matrix
    M = rows(read(save))[:, 1:]  # drop the frame column
    out = str(tmp_path / "m.dat")
    write_matrix(out, M, M.shape[1], M.shape[0], tri=False, square2d=True)
    assert open(out).read() == read(save)


def test_square2d_triangle_roundtrip(tmp_path):
    """A TRI matrix prints its symmetric square with a zero diagonal."""
    from cpptraj_amd.host import write_matrix
    M = rows(read("Test_2DRMS__rmsd.dat.save"))[:, 1:]
    tri = M[np.triu_indices(10, 1)]
    out = str(tmp_path / "t.dat")
    write_matrix(out, tri, 10, tri=True, square2d=True)
    assert open(out).read() == read("Test_2DRMS__rmsd.dat.save")


def test_nosquare2d_roundtrip(tmp_path):
    from cpptraj_amd.host import write_matrix
    v = rows(read("Test_2DRMS__rgacc.dat.save"))
    M = v[:, 2].reshape(9, 9)
    out = str(tmp_path / "r.dat")
    write_matrix(out, M, 9, 9, tri=False, square2d=False, name="crd1")
    assert open(out).read() == read("Test_2DRMS__rgacc.dat.save")


@pytest.mark.parametrize("save,name,width,prec", [
    ("Test_RMSD__rmsd.dat.save", "Res2-11", 12, 4),
    ("Test_RMSD__NoMod.dat.save", "First_CA", 12, 4),
    ("Test_RMSD__rmatrices.dat.save", "ROT[RM]", 12, 9),
    ("Test_RMSD__vecs.dat.save", "Res2-11_mass[TV]", 8, 4),
])
def test_series_roundtrip(tmp_path, save, name, width, prec):
    from cpptraj_amd.host import write_series
    v = rows(read(save))[:, 1:]
    out = str(tmp_path / "s.dat")
    write_series(out, [(name, v, width, prec)])
    assert open(out).read() == read(save)


def test_cmatrix_roundtrip_and_layout(tmp_path):
    from cpptraj_amd.host import read_cmatrix, write_cmatrix
    rng = np.random.default_rng(3)
    frames = np.array([0, 3, 6, 9, 12], np.int32)
    tri = rng.random(10).astype(np.float32)
    p = str(tmp_path / "c.cmatrix")
    write_cmatrix(p, tri, 5, 14, 3, frames)
    t2, actual, sieve, fr = read_cmatrix(p)
    assert actual == 14 and sieve == 3 and np.array_equal(fr, frames)
    assert np.array_equal(t2, tri)
    raw = open(p, "rb").read()
    assert raw[:4] == b"CTM\x02"
    assert np.frombuffer(raw[4:20], "<u8").tolist() == [14, 5]
    assert np.frombuffer(raw[20:28], "<i8")[0] == 3
    assert np.array_equal(np.frombuffer(raw[28:68], "<f4"), tri)
    assert raw[68:] == b"FTTFTTFTTFTTFT"
    # no sieve: no presence flags
    write_cmatrix(p, tri, 5, 5, 1, np.arange(5))
    assert len(open(p, "rb").read()) == 28 + 40


def _agr_sets(text):
    legends, sets = [], []
    for line in text.splitlines():
        if line.startswith("@  s") and "legend" in line:
            legends.append(line.split('"')[1])
        elif line.startswith("@target"):
            sets.append([])
        elif line.strip() and not line.startswith("@"):
            sets[-1].append(float(line.split()[1]))
    return legends, sets


@pytest.mark.parametrize("save", ["Test_PerResRMSD__PerResRMSD.agr.save",
                                  "Test_PerResRMSD__center.agr.save"])
def test_grace_roundtrip(tmp_path, save):
    """Per-residue series in DataIO_Grace's normal layout (perresout)."""
    from cpptraj_amd.host import write_grace
    legends, sets = _agr_sets(read(save))
    out = str(tmp_path / "p.agr")
    write_grace(out, legends, sets)
    assert open(out).read() == read(save)


def test_grace_inverted_layout(tmp_path):
    """perresinvert: one xy set per frame, x = data set number from 0, each
    line ending with the set's legend (DataIO_Grace::WriteDataInverted)."""
    from cpptraj_amd.host import write_grace
    legends, sets = _agr_sets(read("Test_PerResRMSD__PerResRMSD.agr.save"))
    out = str(tmp_path / "i.agr")
    write_grace(out, legends, sets, invert=True)
    lines = open(out).read().splitlines()
    assert lines[:5] == ['@with g0', '@  xaxis label ""', '@  yaxis label "Frame"',
                         '@  legend 0.2, 0.995', '@  legend char size 0.60']
    body = lines[5:]
    nsets, nframes = len(sets), len(sets[0])
    assert len(body) == nframes * (2 + nsets)
    for f in range(nframes):
        blk = body[f * (2 + nsets):(f + 1) * (2 + nsets)]
        assert blk[0] == f"@target G0.S{f}" and blk[1] == "@type xy"
        for k, line in enumerate(blk[2:]):
            assert line == "%8.3f %12.4f \"%s\"" % (k, sets[k][f], legends[k])


def test_mesh_roundtrip(tmp_path):
    """perresavg: residue number, average and standard deviation."""
    from cpptraj_amd.host import write_mesh
    text = read("Test_PerResRMSD__perresavg.dat.save")
    R = rows(text)
    out = str(tmp_path / "avg.dat")
    write_mesh(out, "Residue", R[:, 0], ["R2-11[Avg]", "R2-11[Stdev]"], [R[:, 1], R[:, 2]])
    assert open(out).read() == text
