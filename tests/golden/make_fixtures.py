#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ from the reference's own test data.

Run in the build container (where the reference tree exists):

    python tests/golden/make_fixtures.py [--ref /root/reference]

Inputs (read as data, never executed):
  test/tz2.parm7, test/tz2.nc                  (223 atoms, 101 frames)
  test/tz2.truncoct.parm7, test/tz2.truncoct.nc (5827 atoms, 10 frames)
  test/tz2.ortho.parm7, test/tz2.ortho.nc       (5293 atoms, 10 frames, box)
  test/rGACC.full.parm7, test/rGACC.full.nc     (7622 atoms, 1 frame)
  test/Test_2DRMS/*.save, test/Test_RMSD/*.save (expected outputs)
  test/AFV.parm7, test/AFV.nc, test/Test_SymmRmsd/* (symmetric RMSD)

Outputs (numpy .npz, no pickles):
  tz2_2drms.npz    frames 1-10 of tz2.nc, masses, masks :3-7 / :2 / :11 /
                   :2-12@CA, and the Test_2DRMS expected matrices
  tz2_rmsd.npz     tz2.truncoct.nc :2-11 atoms and masses, plus expected RMSD
                   (plain, mass, reftraj) and combined translation vectors
  tz2_rotate.npz   reference = tz2.nc frame 1, targets = tz2.truncoct.nc with
                   waters stripped, @CA mask, expected rotation matrices and the
                   expected rotated / translated coordinates
  tz2_ortho.npz    tz2.ortho.nc (orthorhombic box) all 10 frames: the :1-13
                   solute and the atoms of residues 559 and 1412, masses,
                   box lengths and angles, and Test_RMSD's distances.agr.save
                   (distance between the two residues without / with imaging
                   after 'rms :1-13 first' rotated coordinates and box)
  tz2_camask.npz   tz2.nc all 101 frames, :2-12@CA atoms, expected 'nomod' and
                   'previous' RMSD series
  rgacc.npz        rGACC.full.nc frame 1, :1-4&!@H= atoms (9 identical frames ->
                   all-zero matrix)
  tz2_perres.npz   tz2.truncoct.nc solute atoms (13 residues, 223 atoms) of all
                   10 frames, masses, residue first atoms and names, and the
                   Test_PerResRMSD expected per-residue series, averages and
                   the centred residue-1 series
  tz2_avgcorr.npz  Test_RmsAvgCorr: the averaged CA reference (avg.CA.rst7,
                   11 atoms of :2-12@CA) and the expected RMS average
                   correlation series (reference, offset 10, first); the
                   frames are those of tz2_camask.npz
  symm_afv.npz     Test_SymmRmsd, AFV heavy atoms: reference AFV.rst7, the
                   three rotated structures and AFV.nc (150 frames), the
                   symmetric groups, masses, expected NoSymm / Symm RMSDs, and
                   Test_Cluster_SymmRMSD's '2drms srmsd' matrix (2drms.gnu.save)
  symm_asp.npz, symm_glu.npz, symm_tyr.npz
                   Test_SymmRmsd per residue: all frames and atoms, groups,
                   masses, expected NOSYMM / SYMM RMSDs and the remapped,
                   fitted coordinates (remap.crd.save)
  text/            the expected text outputs themselves (Test_2DRMS,
                   Test_RMSD, Test_PerResRMSD and Test_RmsAvgCorr .save
                   files), for byte-level checks of the host layer's
                   DataIO_Std / DataIO_Grace writers

The atom masks used by these tests are resolved here with a few lines of
parm7 parsing (residue ranges, atom names, hydrogen exclusion); the general
mask language is outside the scope of the RMSD path.
"""
from __future__ import annotations

import argparse
import os
import re

import numpy as np
from scipy.io import netcdf_file

OUT = os.path.dirname(os.path.abspath(__file__))


# --------------------------------------------------------------------------
def read_parm7(path):
    """Minimal Amber parm7 reader: returns dict flag -> list of tokens."""
    flags = {}
    cur = None
    fmt_w = None
    fmt_kind = None
    with open(path) as fh:
        for line in fh:
            line = line.rstrip("\n")
            if line.startswith("%FLAG"):
                cur = line.split()[1]
                flags[cur] = []
                continue
            if line.startswith("%FORMAT"):
                m = re.match(r"%FORMAT\((\d+)([aIE])(\d+)", line)
                fmt_kind = m.group(2)
                fmt_w = int(m.group(3))
                continue
            if line.startswith("%VERSION") or cur is None:
                continue
            for k in range(0, len(line), fmt_w):
                tok = line[k:k + fmt_w]
                if fmt_kind == "a":
                    flags[cur].append(tok.strip())
                elif tok.strip():
                    flags[cur].append(tok.strip())
    return flags


class Parm:
    def __init__(self, path):
        f = read_parm7(path)
        self.names = f["ATOM_NAME"]
        self.natom = len(self.names)
        self.mass = np.array([float(x) for x in f["MASS"]], dtype=np.float64)
        ptr = [int(x) - 1 for x in f["RESIDUE_POINTER"]]
        self.resnames = f["RESIDUE_LABEL"]
        self.res_of_atom = np.zeros(self.natom, dtype=np.int64)
        bounds = ptr + [self.natom]
        for r in range(len(ptr)):
            self.res_of_atom[bounds[r]:bounds[r + 1]] = r

    def select(self, res=None, names=None, no_h=False, exclude_res=None):
        """Residues are 1-based inclusive ranges [(a, b), ...]."""
        keep = np.ones(self.natom, dtype=bool)
        if res is not None:
            rk = np.zeros(self.natom, dtype=bool)
            for a, b in res:
                rk |= (self.res_of_atom >= a - 1) & (self.res_of_atom <= b - 1)
            keep &= rk
        if names is not None:
            keep &= np.array([n in names for n in self.names])
        if no_h:
            keep &= np.array([not n.startswith("H") for n in self.names])
        if exclude_res is not None:
            keep &= np.array([self.resnames[r] not in exclude_res for r in self.res_of_atom])
        return np.nonzero(keep)[0].astype(np.int32)


def read_nc(path):
    f = netcdf_file(path, "r", mmap=False)
    x = np.array(f.variables["coordinates"].data, dtype=np.float32)
    f.close()
    return x


def read_square2d(path):
    rows = []
    with open(path) as fh:
        for line in fh:
            if line.startswith("#"):
                continue
            tok = line.split()
            rows.append([float(t) for t in tok[1:]])
    return np.array(rows, dtype=np.float64)


def read_series(path, ncols=None):
    rows = []
    with open(path) as fh:
        for line in fh:
            if line.startswith("#"):
                continue
            tok = line.split()
            rows.append([float(t) for t in tok[1:]])
    return np.array(rows, dtype=np.float64)


def read_agr(path):
    """Grace xy sets -> (legends, array (nsets, npoints))."""
    legends, sets = [], []
    with open(path) as fh:
        for line in fh:
            if line.startswith("@  s") and "legend" in line:
                legends.append(line.split('"')[1])
            elif line.startswith("@target"):
                sets.append([])
            elif not line.startswith("@") and line.strip():
                sets[-1].append(float(line.split()[1]))
    return legends, np.array(sets, dtype=np.float64)


def read_rst7(path):
    """Amber ASCII restart: title, natom line, then 6 coordinates per line."""
    with open(path) as fh:
        fh.readline()
        natom = int(fh.readline().split()[0])
        vals = []
        for line in fh:
            vals.extend(float(t) for t in line.split())
            if len(vals) >= 3 * natom:
                break
    return np.array(vals[:3 * natom], dtype=np.float64).reshape(natom, 3)


def read_crd(path, natom, nbox=0):
    """Amber ASCII trajectory (%8.3f, 10 per line); nbox trailing box values
    per frame are dropped."""
    vals = []
    with open(path) as fh:
        fh.readline()
        for line in fh:
            line = line.rstrip("\n")
            for k in range(0, len(line), 8):
                tok = line[k:k + 8]
                if tok.strip():
                    vals.append(float(tok))
    a = np.array(vals, dtype=np.float64).reshape(-1, 3 * natom + nbox)
    return a[:, :3 * natom].reshape(-1, natom, 3)


def symmetric_groups(path, sel):
    """Groups of potentially symmetric selected atoms, as cpptraj's
    SymmetricRmsdCalc::SetupSymmRMSD (src/SymmetricRmsdCalc.cpp:28-119) forms
    them through AtomMap::SymmetricAtoms (src/AtomMap.cpp:374-439), written
    here from that behaviour to produce the inputs of the symmetric RMSD
    fixtures (the library takes the groups from its caller):
      * per residue up to the last selected one, bonds within the residue,
        neighbour lists in atom order;
      * a 'chiral' atom: 4 bonds and fewer than 3 bonded terminal atoms
        (AtomMap::CheckBonds, :269-298);
      * atom ID: element + sorted neighbour elements; unique ID: that plus the
        atom IDs one bond out, two bonds out for residues over 9 atoms and
        three for residues over 19, all but the first character sorted
        (AtomMap::AssignUniqueAtomIDs, :132-187);
      * a group: the atoms with one unique ID reached from its first atom
        without entering a chiral atom (FindSymmetricAtoms, :336-367); groups
        of one atom are dropped, the rest kept in visiting order;
      * then only selected atoms, as selection indices; groups left with at
        least one atom are kept.
    Elements come from ATOMIC_NUMBER (any one-to-one character map gives the
    same equalities)."""
    f = read_parm7(path)
    natom = len(f["ATOM_NAME"])
    z = [int(v) for v in f["ATOMIC_NUMBER"]]
    ch = [chr(0x100 + v) for v in z]
    nb = [[] for _ in range(natom)]
    for key in ("BONDS_INC_HYDROGEN", "BONDS_WITHOUT_HYDROGEN"):
        v = [int(t) for t in f.get(key, [])]
        for k in range(0, len(v), 3):
            a, b = v[k] // 3, v[k + 1] // 3
            nb[a].append(b)
            nb[b].append(a)
    ptr = [int(v) - 1 for v in f["RESIDUE_POINTER"]] + [natom]
    sel = [int(a) for a in sel]
    sel_idx = {a: i for i, a in enumerate(sel)}
    res_of = np.zeros(natom, np.int64)
    for r in range(len(ptr) - 1):
        res_of[ptr[r]:ptr[r + 1]] = r
    last_res = int(res_of[sel[-1]]) + 1
    groups = []
    for r in range(last_res):
        a0, a1 = ptr[r], ptr[r + 1]
        n = a1 - a0
        bl = [sorted(b - a0 for b in nb[a] if a0 <= b < a1) for a in range(a0, a1)]
        c = [ch[a] for a in range(a0, a1)]
        chiral = [len(bl[i]) == 4 and sum(len(bl[j]) == 1 for j in bl[i]) < 3 for i in range(n)]
        aid = [c[i] + "".join(sorted(c[j] for j in bl[i])) for i in range(n)]
        uid = []
        for i in range(n):
            u = aid[i]
            if n > 9:
                for b in bl[i]:
                    u += aid[b]
                    for c2 in bl[b]:
                        if c2 != i:
                            u += aid[c2]
                            if n > 19:
                                for d in bl[c2]:
                                    if d != i:
                                        u += aid[d]
            uid.append(u[0] + "".join(sorted(u[1:])))
        dup = [sum(uid[j] == uid[i] for j in range(n)) > 1 for i in range(n)]
        status = [0] * n  # 0 unselected, 1 nonsymm, 2 symm
        for at in range(n):
            if not dup[at]:
                status[at] = 1
            elif status[at] != 2:
                seen = [False] * n
                found = []

                def visit(x, want):
                    if seen[x]:
                        return
                    seen[x] = True
                    if uid[x] == want:
                        found.append(x)
                    for y in bl[x]:
                        if not chiral[y]:
                            visit(y, want)
                visit(at, uid[at])
                if len(found) == 1:
                    status[found[0]] = 1
                elif len(found) > 1:
                    for x in found:
                        status[x] = 2
                    g = [sel_idx[x + a0] for x in found if x + a0 in sel_idx]
                    if g:
                        groups.append(g)
    return groups


# --------------------------------------------------------------------------
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    args = ap.parse_args()
    T = os.path.join(args.ref, "test")

    # ---- Test_2DRMS --------------------------------------------------------
    p = Parm(os.path.join(T, "tz2.parm7"))
    x = read_nc(os.path.join(T, "tz2.nc"))  # (101, 223, 3)
    frames10 = x[:10].reshape(10, -1)
    np.savez_compressed(
        os.path.join(OUT, "tz2_2drms.npz"),
        frames=frames10,
        mass=p.mass,
        sel_3_7=p.select(res=[(3, 7)]),
        sel_2=p.select(res=[(2, 2)]),
        sel_11=p.select(res=[(11, 11)]),
        sel_2_12_ca=p.select(res=[(2, 12)], names={"CA"}),
        rmsd=read_square2d(os.path.join(T, "Test_2DRMS", "rmsd.dat.save")),
        rmsd_mass=read_square2d(os.path.join(T, "Test_2DRMS", "rmsd.mass.dat.save")),
        trp=read_square2d(os.path.join(T, "Test_2DRMS", "trp.dat.save")),
        nofit=read_square2d(os.path.join(T, "Test_2DRMS", "nofit.dat.save")),
    )

    # ---- Test_RMSD basic (truncoct, :2-11) -------------------------------------
    pt = Parm(os.path.join(T, "tz2.truncoct.parm7"))
    xt = read_nc(os.path.join(T, "tz2.truncoct.nc"))  # (10, 5827, 3)
    s = pt.select(res=[(2, 11)])
    vec = read_series(os.path.join(T, "Test_RMSD", "vecs.dat.save"))
    np.savez_compressed(
        os.path.join(OUT, "tz2_rmsd.npz"),
        frames=xt[:, s, :].reshape(xt.shape[0], -1),
        mass=pt.mass[s],
        rmsd=read_series(os.path.join(T, "Test_RMSD", "rmsd.dat.save"))[:, 0],
        rmsd_mass=read_series(os.path.join(T, "Test_RMSD", "rmsd.mass.dat.save"))[:, 0],
        rmsd_reftraj=read_series(os.path.join(T, "Test_RMSD", "rmsd.reftraj.dat.save"))[:, 0],
        vecs_mass=vec,
    )

    # ---- Test_RMSD rotation matrices / coordinates ----------------------------
    solute = pt.select(exclude_res={"WAT"})
    tgt = xt[:, solute, :]
    ca_ref = p.select(names={"CA"})
    ca_tgt = pt.select(names={"CA"}, exclude_res={"WAT"})
    assert len(ca_ref) == len(ca_tgt)
    # index of each target CA within the stripped (solute-only) frame
    ca_tgt = np.searchsorted(solute, ca_tgt).astype(np.int32)
    np.savez_compressed(
        os.path.join(OUT, "tz2_rotate.npz"),
        ref_frame=x[0].reshape(-1),
        frames=tgt.reshape(tgt.shape[0], -1),
        sel_ca_ref=ca_ref,
        sel_ca=ca_tgt,
        rmatrices=read_series(os.path.join(T, "Test_RMSD", "rmatrices.dat.save")),
        norotate=read_crd(os.path.join(T, "Test_RMSD", "tz2.norotate.crd.save"), len(solute), 3),
        rotate=read_crd(os.path.join(T, "Test_RMSD", "tz2.rotate.crd.save"), len(solute), 3),
    )

    # ---- Test_RMSD imaging after unit-cell rotation (tz2.ortho) ----------------
    # 'rms :1-13 first' rotates coordinates and box; the distances between the
    # centres of mass of residues 559 and 1412 without and with imaging
    # (distances.agr.save) then show whether the box was rotated with them.
    # Kept: the solute (:1-13, the fit mask) and the two residues' atoms.
    po = Parm(os.path.join(T, "tz2.ortho.parm7"))
    fo = netcdf_file(os.path.join(T, "tz2.ortho.nc"), "r", mmap=False)
    xo = np.array(fo.variables["coordinates"].data, dtype=np.float32)
    box = np.concatenate([np.array(fo.variables["cell_lengths"].data, dtype=np.float64),
                          np.array(fo.variables["cell_angles"].data, dtype=np.float64)], axis=1)
    fo.close()
    sol = po.select(res=[(1, 13)])
    r559 = po.select(res=[(559, 559)])
    r1412 = po.select(res=[(1412, 1412)])
    keep = np.concatenate([sol, r559, r1412])
    legends, dist = read_agr(os.path.join(T, "Test_RMSD", "distances.agr.save"))
    assert legends == ["NoImage", "Image"], legends
    np.savez_compressed(
        os.path.join(OUT, "tz2_ortho.npz"),
        frames=xo[:, keep, :].reshape(xo.shape[0], -1),
        box=box,
        nfit=np.int32(len(sol)),
        res_a=np.arange(len(sol), len(sol) + len(r559), dtype=np.int32),
        res_b=np.arange(len(sol) + len(r559), len(keep), dtype=np.int32),
        mass=po.mass[keep],
        dist_noimage=dist[0],
        dist_image=dist[1],
    )

    # ---- Test_RMSD nomod / previous (:2-12@CA over 101 frames) ---------------
    s_ca = p.select(res=[(2, 12)], names={"CA"})
    np.savez_compressed(
        os.path.join(OUT, "tz2_camask.npz"),
        frames=x[:, s_ca, :].reshape(x.shape[0], -1),
        mass=p.mass[s_ca],
        nomod=read_series(os.path.join(T, "Test_RMSD", "NoMod.dat.save"))[:, 0],
        previous=read_series(os.path.join(T, "Test_RMSD", "Previous.dat.save"))[:, 0],
    )

    # ---- Test_2DRMS rGACC (:1-4&!@H=) ------------------------------------------
    pg = Parm(os.path.join(T, "rGACC.full.parm7"))
    xg = read_nc(os.path.join(T, "rGACC.full.nc"))
    sg = pg.select(res=[(1, 4)], no_h=True)
    vals = []
    with open(os.path.join(T, "Test_2DRMS", "rgacc.dat.save")) as fh:
        for line in fh:
            if not line.startswith("#"):
                vals.append([float(t) for t in line.split()])
    np.savez_compressed(
        os.path.join(OUT, "rgacc.npz"),
        frame=xg[0, sg, :].reshape(-1),
        expected=np.array(vals),
    )
    # ---- Test_PerResRMSD (truncoct; rmsd :2-11 refindex 0 perres ...) ----------
    P = os.path.join(T, "Test_PerResRMSD")
    legends, perres = read_agr(os.path.join(P, "PerResRMSD.agr.save"))
    _, center = read_agr(os.path.join(P, "center.agr.save"))
    nsolute = len(solute)
    ptr = [int(v) - 1 for v in read_parm7(os.path.join(T, "tz2.truncoct.parm7"))["RESIDUE_POINTER"]]
    nres = int(np.count_nonzero(np.array(ptr) < nsolute))
    np.savez_compressed(
        os.path.join(OUT, "tz2_perres.npz"),
        frames=xt[:, solute, :].reshape(xt.shape[0], -1),
        mass=pt.mass[solute],
        res_first=np.array(ptr[:nres], dtype=np.int32),
        res_names=np.array(pt.resnames[:nres]),
        legends=np.array(legends),
        perres=perres,
        perresavg=read_series(os.path.join(P, "perresavg.dat.save")),
        center=center,
    )

    # ---- Test_RmsAvgCorr (strip !(:2-12@CA); reference avg.CA.rst7) -----------
    A = os.path.join(T, "Test_RmsAvgCorr")
    np.savez_compressed(
        os.path.join(OUT, "tz2_avgcorr.npz"),
        ref=read_rst7(os.path.join(A, "avg.CA.rst7")),
        ref_all=read_series(os.path.join(A, "rmscorr.dat.save")),
        off10=read_series(os.path.join(A, "rmscorr.10.dat.save")),
        first=read_series(os.path.join(A, "rmscorr.first.dat.save")),
    )

    # ---- expected text outputs, verbatim -------------------------------------------
    txt = os.path.join(OUT, "text")
    os.makedirs(txt, exist_ok=True)
    for sub, names in (("Test_2DRMS", ["rmsd.dat.save", "rmsd.mass.dat.save", "trp.dat.save",
                                       "nofit.dat.save", "rgacc.dat.save"]),
                       ("Test_RMSD", ["rmsd.dat.save", "rmsd.mass.dat.save",
                                      "rmsd.reftraj.dat.save", "vecs.dat.save",
                                      "rmatrices.dat.save", "NoMod.dat.save",
                                      "Previous.dat.save"]),
                      ("Test_PerResRMSD", ["PerResRMSD.agr.save", "perresavg.dat.save",
                                           "center.agr.save"]),
                      ("Test_RmsAvgCorr", ["rmscorr.dat.save", "rmscorr.10.dat.save",
                                           "rmscorr.first.dat.save"]),
                      ("Test_SymmRmsd", ["rmsd.dat.save", "afv.rmsd.dat.save",
                                         "ASP.rmsd.dat.save", "GLU.rmsd.dat.save",
                                         "TYR.rmsd.dat.save"])):
        for n in names:
            with open(os.path.join(T, sub, n), "rb") as fi:
                data = fi.read()
            with open(os.path.join(txt, sub + "__" + n), "wb") as fo:
                fo.write(data)
    # ---- Test_SymmRmsd -----------------------------------------------------
    S = os.path.join(T, "Test_SymmRmsd")

    def pack_groups(groups):
        off = np.zeros(len(groups) + 1, np.int32)
        for g, a in enumerate(groups):
            off[g + 1] = off[g] + len(a)
        return off, np.concatenate([np.asarray(a, np.int32) for a in groups])

    # AFV, heavy atoms (!@H=): reference AFV.rst7, three rotated structures
    # (rmsd.dat.save: NoSymm, Symm) and 150 frames of AFV.nc against its first
    # frame (afv.rmsd.dat.save).
    pa = Parm(os.path.join(T, "AFV.parm7"))
    sel = pa.select(no_h=True)
    off, flat = pack_groups(symmetric_groups(os.path.join(T, "AFV.parm7"), sel))
    rot = np.stack([read_rst7(os.path.join(S, n)) for n in
                    ("AFV.rotate2.rst7", "AFV.rotate3.rst7", "AFV.rotate23.rst7")])
    afv_nc = read_nc(os.path.join(T, "AFV.nc"))
    # Test_Cluster_SymmRMSD: '2drms !@H= out 2drms.gnu srmsd' over AFV.nc, the
    # gnuplot grid (x, y 1-based, padded with a 151st row and column of 0).
    nfr = afv_nc.shape[0]
    gnu = np.full((nfr, nfr), np.nan)
    with open(os.path.join(T, "Test_Cluster_SymmRMSD", "2drms.gnu.save")) as fh:
        for line in fh:
            tok = line.split()
            if len(tok) == 3 and tok[0][0].isdigit():
                x, y = int(float(tok[0])), int(float(tok[1]))
                if x <= nfr and y <= nfr:
                    gnu[x - 1, y - 1] = float(tok[2])
    assert not np.isnan(gnu).any()
    np.savez_compressed(
        os.path.join(OUT, "symm_afv.npz"),
        sel=sel, goff=off, gatoms=flat, mass=pa.mass,
        ref=read_rst7(os.path.join(S, "AFV.rst7")), rotated=rot,
        expect_rotated=read_series(os.path.join(S, "rmsd.dat.save")),
        traj=afv_nc, expect_traj=read_series(os.path.join(S, "afv.rmsd.dat.save"))[:, 0],
        expect_2drms=gnu)
    # ASP / GLU / TYR, all atoms, reference = first frame, with 'remap':
    # NOSYMM / SYMM columns and the remapped, fitted coordinates.
    for res in ("ASP", "GLU", "TYR"):
        pp = os.path.join(S, res + ".parm7")
        pr = Parm(pp)
        allsel = np.arange(pr.natom, dtype=np.int32)
        off, flat = pack_groups(symmetric_groups(pp, allsel))
        np.savez_compressed(
            os.path.join(OUT, f"symm_{res.lower()}.npz"),
            goff=off, gatoms=flat, mass=pr.mass, frames=read_nc(os.path.join(S, res + ".nc")),
            expect=read_series(os.path.join(S, res + ".rmsd.dat.save")),
            remap=read_crd(os.path.join(S, res + ".remap.crd.save"), pr.natom))
    for fn in sorted(os.listdir(OUT)):
        if fn.endswith(".npz"):
            print(fn, os.path.getsize(os.path.join(OUT, fn)))


if __name__ == "__main__":
    main()
