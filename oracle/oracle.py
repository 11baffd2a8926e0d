"""ctypes bindings for the test-only oracle libraries.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

* ``Oracle``  wraps ``_build/liboracle.so`` (rmsd_oracle.c, the C restatement).
* ``RefLib``  wraps ``_ref/libcpptraj_ref.so`` (the reference's own Frame /
  Matrix_3x3 / qcprot / QuaternionRMSD sources compiled in place, plus
  ref_shim.cpp).  It exists only where oracle/Makefile could build it.

All arrays are numpy; coordinates are AoS (n, 3) float64 for the per-frame
routines and AoS float32 frames (F, natom*3) for the loop drivers.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libcpptraj_ref.so")

MODE_FIT, MODE_NOFIT, MODE_QUAT = 0, 1, 2

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


def _d(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _f(a):
    return None if a is None else a.ctypes.data_as(_fp)


def _i(a):
    return None if a is None else a.ctypes.data_as(_ip)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ci(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)



def _groups(groups):
    """List of index lists -> (offsets, flat atoms) int32."""
    off = np.zeros(len(groups) + 1, np.int32)
    for g, a in enumerate(groups):
        off[g + 1] = off[g] + len(a)
    flat = np.concatenate([np.asarray(a, np.int32) for a in groups]) if groups else \
        np.zeros(0, np.int32)
    return off, np.ascontiguousarray(flat, np.int32)


def _hungarian(fn, M):
    M = np.array(M, dtype=np.float64, order="C")
    n = M.shape[0]
    m = np.full(max(n, 1), -1, np.int32)
    fn(_d(M), n, _i(m))
    return m[:n]


def _symm(fn, frames, sel, mass, ref_sel, use_mass, fit, groups):
    fr = _c32(frames)
    s = _ci(sel)
    nsel = len(s) if s is not None else fr.shape[1] // 3
    F = fr.shape[0]
    m = _c64(mass) if mass is not None else None
    rs = _c64(ref_sel)
    off, flat = _groups(groups)
    r = np.zeros(F)
    amap = np.zeros((F, nsel), np.int32)
    U = np.zeros((F, 9))
    tr = np.zeros((F, 3))
    err = fn(_f(fr), fr.shape[1], F, _i(s), nsel, _d(m), _d(rs), int(use_mass), int(fit),
             len(groups), _i(off), _i(flat), _d(r), _i(amap), _d(U), _d(tr))
    if err:
        raise MemoryError("symmetric RMSD failed")
    return r, amap, U, tr

class Oracle:
    """The C restatement (rmsd_oracle.c)."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"oracle library not built: {path} (run make -C oracle)")
        L = C.CDLL(path)
        self.L = L
        L.orc_center_on_origin.argtypes = [_dp, _dp, C.c_int, C.c_int, _dp]
        L.orc_jacobi3_sorted.argtypes = [_dp, _dp, _dp]
        L.orc_jacobi3_sorted.restype = C.c_int
        L.orc_kabsch_from_cov.argtypes = [_dp, C.c_double, C.c_double, _dp]
        L.orc_kabsch_from_cov.restype = C.c_double
        L.orc_rmsd_centered_ref.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, _dp, _dp]
        L.orc_rmsd_centered_ref.restype = C.c_double
        L.orc_qcp.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double]
        L.orc_qcp.restype = C.c_int
        L.orc_qrmsd_centered_ref.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, _dp, _dp]
        L.orc_qrmsd_centered_ref.restype = C.c_double
        L.orc_rmsd_nofit.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int]
        L.orc_rmsd_nofit.restype = C.c_double
        L.orc_tri_index.argtypes = [C.c_int64] * 3
        L.orc_tri_index.restype = C.c_int64
        L.orc_rotate_ucell.argtypes = [_dp, _dp, _dp]
        L.orc_rotate_ucell.restype = None
        L.orc_full_index.argtypes = [C.c_int64] * 3
        L.orc_full_index.restype = C.c_int64
        L.orc_rms2d.argtypes = [_fp, C.c_int64, C.c_int, _fp, C.c_int64, C.c_int,
                                _ip, _ip, C.c_int, _dp, _dp, C.c_int, C.c_int, C.c_int,
                                _fp, C.c_int]
        L.orc_rms2d.restype = C.c_int
        L.orc_cluster_cache.argtypes = [_fp, C.c_int64, _ip, C.c_int, _ip, C.c_int, _dp,
                                        C.c_int, C.c_int, C.c_double, _fp, C.c_int]
        L.orc_cluster_cache.restype = C.c_int
        L.orc_fit_frames.argtypes = [_fp, C.c_int64, C.c_int, C.c_int, _ip, C.c_int, _dp,
                                     _dp, _dp, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp,
                                     _fp, C.c_int]
        L.orc_fit_frames.restype = C.c_int
        L.orc_rmsd_grad_target.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, _dp]
        L.orc_rmsd_grad_target.restype = C.c_double
        L.orc_rms2d_rows.argtypes = [_fp, C.c_int64, C.c_int, _ip, C.c_int, _dp, C.c_int,
                                     C.c_int, _ip, C.c_int, _fp, C.c_int]
        L.orc_rms2d_rows.restype = C.c_int
        L.orc_perres.argtypes = [_fp, C.c_int64, C.c_int, _fp, C.c_int, _ip, _ip, _ip, _dp,
                                 _dp, C.c_int, C.c_int, _dp, _dp, C.c_int, _dp]
        L.orc_perres.restype = C.c_int
        L.orc_avgcorr_windows.argtypes = [C.c_int, C.c_int, C.c_int, _ip, C.c_int]
        L.orc_avgcorr_windows.restype = C.c_int
        L.orc_avgcorr.argtypes = [_fp, C.c_int64, C.c_int, _ip, C.c_int, _dp, _dp, C.c_int,
                                  C.c_int, C.c_int, _dp, _dp]
        L.orc_avgcorr.restype = C.c_int
        L.orc_hungarian.argtypes = [_dp, C.c_int, _ip]
        L.orc_hungarian.restype = C.c_int
        L.orc_symm_frames.argtypes = [_fp, C.c_int64, C.c_int, _ip, C.c_int, _dp, _dp, C.c_int,
                                      C.c_int, C.c_int, _ip, _ip, _dp, _ip, _dp, _dp]
        L.orc_symm_frames.restype = C.c_int

    # -- per-frame routines -------------------------------------------------
    def center_on_origin(self, xyz, mass=None, use_mass=False):
        x = _c64(xyz).copy()
        c = np.zeros(3)
        m = _c64(mass) if mass is not None else None
        self.L.orc_center_on_origin(_d(x), _d(m), len(x), int(use_mass), _d(c))
        return x, c

    def jacobi3_sorted(self, m):
        E = np.zeros(9)
        ev = np.zeros(3)
        err = self.L.orc_jacobi3_sorted(_d(_c64(m).ravel()), _d(E), _d(ev))
        return err, E.reshape(3, 3), ev

    def kabsch_from_cov(self, R, E0, total_mass):
        U = np.zeros(9)
        r = self.L.orc_kabsch_from_cov(_d(_c64(R).ravel()), E0, total_mass, _d(U))
        return r, U.reshape(3, 3)

    def rmsd_centered_ref(self, tgt, ref, mass=None, use_mass=False):
        t = _c64(tgt).copy()
        U = np.zeros(9)
        tr = np.zeros(3)
        m = _c64(mass) if mass is not None else None
        r = self.L.orc_rmsd_centered_ref(_d(t), _d(_c64(ref)), _d(m), len(t),
                                         int(use_mass), _d(U), _d(tr))
        return r, U.reshape(3, 3), tr, t

    def qcp(self, A, E0, length, want_rot=True):
        rot = np.zeros(9) if want_rot else None
        rmsd = C.c_double(0.0)
        ret = self.L.orc_qcp(_d(rot), _d(_c64(A).ravel()), C.byref(rmsd), E0, length)
        return ret, rmsd.value, (rot.reshape(3, 3) if want_rot else None)

    def qrmsd_centered_ref(self, ref, tgt, mass=None, use_mass=False, want_rot=False):
        t = _c64(tgt).copy()
        tr = np.zeros(3)
        rot = np.zeros(9) if want_rot else None
        m = _c64(mass) if mass is not None else None
        r = self.L.orc_qrmsd_centered_ref(_d(_c64(ref)), _d(t), _d(m), len(t),
                                          int(use_mass), _d(rot), _d(tr))
        return r, (rot.reshape(3, 3) if want_rot else None), tr

    def rmsd_nofit(self, tgt, ref, mass=None, use_mass=False):
        m = _c64(mass) if mass is not None else None
        return self.L.orc_rmsd_nofit(_d(_c64(tgt)), _d(_c64(ref)), _d(m), len(tgt),
                                     int(use_mass))

    def rmsd_grad_target(self, tgt, ref_centered, mass=None, use_mass=False):
        g = np.zeros((len(tgt), 3))
        m = _c64(mass) if mass is not None else None
        r = self.L.orc_rmsd_grad_target(_d(_c64(tgt)), _d(_c64(ref_centered)), _d(m),
                                        len(tgt), int(use_mass), _d(g))
        return r, g

    def tri_index(self, nx, x, y):
        return self.L.orc_tri_index(nx, x, y)

    # -- loop drivers ---------------------------------------------------------
    def rms2d(self, frames, sel=None, mass=None, mode=MODE_FIT, use_mass=False,
              ref_frames=None, ref_sel=None, ref_mass=None, full=None, nthreads=0):
        """Analysis_Rms2d::Calculate_2D.  frames: (F, natom*3) float32."""
        tf = _c32(frames)
        ntgt = tf.shape[0]
        tsel = _ci(sel)
        rsel = _ci(ref_sel) if ref_sel is not None else tsel
        nsel = len(tsel) if tsel is not None else tf.shape[1] // 3
        rf = _c32(ref_frames) if ref_frames is not None else None
        nref = rf.shape[0] if rf is not None else ntgt
        if full is None:
            full = rf is not None or (ref_sel is not None and not np.array_equal(ref_sel, sel))
        tm = _c64(mass) if mass is not None else np.ones(nsel)
        rm = _c64(ref_mass) if ref_mass is not None else tm
        n_out = nref * ntgt if full else ntgt * (ntgt - 1) // 2
        out = np.zeros(n_out, dtype=np.float32)
        err = self.L.orc_rms2d(_f(tf), tf.shape[1], ntgt, _f(rf),
                               rf.shape[1] if rf is not None else 0, nref,
                               _i(tsel), _i(rsel), nsel, _d(tm), _d(rm), mode,
                               int(use_mass), int(bool(full)), _f(out), nthreads)
        if err:
            raise MemoryError("orc_rms2d failed")
        return out

    def rms2d_rows(self, frames, rows, sel=None, mass=None, mode=MODE_FIT, use_mass=False,
                   nthreads=0):
        """Rows r of the triangle: pairs (r, t > r) concatenated in row order."""
        fr = _c32(frames)
        F = fr.shape[0]
        rw = np.ascontiguousarray(rows, dtype=np.int32)
        s = _ci(sel)
        nsel = len(s) if s is not None else fr.shape[1] // 3
        m = _c64(mass) if mass is not None else np.ones(nsel)
        out = np.zeros(int(sum(F - 1 - int(r) for r in rw)), dtype=np.float32)
        if self.L.orc_rms2d_rows(_f(fr), fr.shape[1], F, _i(s), nsel, _d(m), mode,
                                 int(use_mass), _i(rw), len(rw), _f(out), nthreads):
            raise MemoryError("orc_rms2d_rows failed")
        return out

    def cluster_cache(self, frames, frame_idx=None, sel=None, mass=None, nofit=False,
                      use_mass=False, weight=1.0, nthreads=0):
        fr = _c32(frames)
        idx = _ci(frame_idx)
        nc = len(idx) if idx is not None else fr.shape[0]
        s = _ci(sel)
        nsel = len(s) if s is not None else fr.shape[1] // 3
        m = _c64(mass) if mass is not None else np.ones(nsel)
        out = np.zeros(nc * (nc - 1) // 2, dtype=np.float32)
        err = self.L.orc_cluster_cache(_f(fr), fr.shape[1], _i(idx), nc, _i(s), nsel,
                                       _d(m), int(nofit), int(use_mass), weight, _f(out),
                                       nthreads)
        if err:
            raise MemoryError("orc_cluster_cache failed")
        return out

    def fit_frames(self, frames, ref_sel_centered, ref_trans, sel=None, mass=None,
                   fit=True, use_mass=False, modify=0, want_frames=False, nthreads=0):
        """Action_Rmsd::DoAction per frame against a fixed centred reference."""
        fr = _c32(frames)
        nf = fr.shape[0]
        natom = fr.shape[1] // 3
        s = _ci(sel)
        nsel = len(s) if s is not None else natom
        m = _c64(mass) if mass is not None else np.ones(nsel)
        rmsd = np.zeros(nf)
        U = np.zeros((nf, 9))
        tv = np.zeros((nf, 6))
        fo = np.zeros_like(fr) if want_frames else None
        err = self.L.orc_fit_frames(_f(fr), fr.shape[1], nf, natom, _i(s), nsel, _d(m),
                                    _d(_c64(ref_sel_centered)), _d(_c64(ref_trans)),
                                    int(fit), int(use_mass), modify, _d(rmsd), _d(U),
                                    _d(tv), _f(fo), nthreads)
        if err:
            raise MemoryError("orc_fit_frames failed")
        return rmsd, U.reshape(nf, 3, 3), tv, fo

    def perres(self, frames, ref, res_off, tgt_atoms, ref_atoms=None, tmass=None,
               rmass=None, use_mass=False, center=False, U=None, tvec=None, modify=0):
        """Per-residue no-fit RMSD (Action_Rmsd 'perres'): (nframes, nres) doubles."""
        fr = _c32(frames)
        nf = fr.shape[0]
        off = _ci(res_off)
        nres = len(off) - 1
        ta = _ci(tgt_atoms)
        ra = _ci(ref_atoms) if ref_atoms is not None else ta
        tm = _c64(tmass) if tmass is not None else None
        rm = _c64(rmass) if rmass is not None else tm
        Uc = _c64(U).reshape(nf, 9) if U is not None else None
        tv = _c64(tvec).reshape(nf, 6) if tvec is not None else None
        out = np.zeros((nf, nres))
        err = self.L.orc_perres(_f(fr), fr.shape[1], nf, _f(_c32(ref).reshape(-1)), nres,
                                _i(off), _i(ta), _i(ra), _d(tm), _d(rm), int(use_mass),
                                int(center), _d(Uc), _d(tv), modify, _d(out))
        if err:
            raise MemoryError("orc_perres failed")
        return out

    def avgcorr_windows(self, nframes, maxwindow=-1, offset=1):
        n = self.L.orc_avgcorr_windows(nframes, maxwindow, offset, None, 0)
        if n < 1:
            return np.zeros(0, np.int32)
        w = np.zeros(n, np.int32)
        self.L.orc_avgcorr_windows(nframes, maxwindow, offset, _i(w), n)
        return w

    def avgcorr(self, frames, sel=None, mass=None, ref_sel_centered=None, use_mass=False,
                maxwindow=-1, offset=1):
        """Analysis_RmsAvgCorr: (window sizes, mean RMSD, SD) per window size;
        ref_sel_centered None = 'first'."""
        fr = _c32(frames)
        s = _ci(sel)
        nsel = len(s) if s is not None else fr.shape[1] // 3
        m = _c64(mass) if mass is not None else np.ones(nsel)
        w = self.avgcorr_windows(fr.shape[0], maxwindow, offset)
        if len(w) == 0:
            raise ValueError("not enough frames")
        ct, sd = np.zeros(len(w)), np.zeros(len(w))
        rs = None if ref_sel_centered is None else _c64(ref_sel_centered)
        n = self.L.orc_avgcorr(_f(fr), fr.shape[1], fr.shape[0], _i(s), nsel, _d(m), _d(rs),
                               int(use_mass), maxwindow, offset, _d(ct), _d(sd))
        if n != len(w):
            raise MemoryError("orc_avgcorr failed")
        return w, ct, sd

    def rotate_ucell(self, U, ucell):
        """Box::RotateUcell: rows of ucell (3x3) rotated by the Kabsch U."""
        out = np.zeros(9)
        self.L.orc_rotate_ucell(_d(_c64(U).ravel()), _d(_c64(ucell).ravel()), _d(out))
        return out.reshape(3, 3)

    def hungarian(self, M):
        """Hungarian::Optimize restated: map[col] = row."""
        return _hungarian(self.L.orc_hungarian, M)

    def symm(self, frames, sel=None, mass=None, ref_sel=None, use_mass=False, fit=True,
             groups=()):
        """Symmetry-corrected RMSD per frame: (rmsd, amap, U, trans)."""
        return _symm(self.L.orc_symm_frames, frames, sel, mass, ref_sel, use_mass, fit, groups)



class RefLib:
    """The reference's own sources compiled in place (oracle/_ref)."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"reference library not built: {path}")
        L = C.CDLL(path)
        self.L = L
        L.cref_rmsd_centered_ref.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_int, _dp, _dp]
        L.cref_rmsd_centered_ref.restype = C.c_double
        L.cref_rmsd.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_int, _dp, _dp, _dp]
        L.cref_rmsd.restype = C.c_double
        L.cref_rmsd_nofit.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int]
        L.cref_rmsd_nofit.restype = C.c_double
        L.cref_center_on_origin.argtypes = [_dp, _dp, C.c_int, C.c_int, _dp]
        L.cref_qrmsd_centered_ref.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, _dp, _dp]
        L.cref_qrmsd_centered_ref.restype = C.c_double
        L.cref_qcp.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double]
        L.cref_qcp.restype = C.c_int
        L.cref_diagonalize_sort.argtypes = [_dp, _dp, _dp]
        L.cref_diagonalize_sort.restype = C.c_int
        L.cref_rms2d.argtypes = [_fp, C.c_int64, C.c_int, _fp, C.c_int64, C.c_int,
                                 _ip, _ip, C.c_int, _dp, _dp, C.c_int, C.c_int, C.c_int,
                                 _fp, C.c_int, C.c_int, C.c_int]
        L.cref_rms2d.restype = C.c_int
        L.cref_rms2d_rows.argtypes = [_fp, C.c_int64, C.c_int, _ip, C.c_int, _dp, C.c_int,
                                      C.c_int, _ip, C.c_int, _fp, C.c_int]
        L.cref_rms2d_rows.restype = C.c_int
        L.cref_fit_frames.argtypes = [_fp, C.c_int64, C.c_int, C.c_int, _ip, C.c_int, _dp,
                                      _dp, _dp, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp,
                                      _fp, C.c_int]
        L.cref_fit_frames.restype = C.c_int
        L.cref_fit_pairs.argtypes = [_fp, _fp, C.c_int64, C.c_int, _ip, C.c_int, _dp, C.c_int,
                                     _dp, C.c_int]
        L.cref_fit_pairs.restype = C.c_int
        L.cref_perres.argtypes = [_fp, C.c_int64, C.c_int, C.c_int, _fp, C.c_int, C.c_int, _ip,
                                  _ip, _ip, _dp, _dp, C.c_int, C.c_int, _dp, _dp, C.c_int, _dp]
        L.cref_perres.restype = C.c_int
        L.cref_avgcorr.argtypes = [_fp, C.c_int64, C.c_int, _ip, C.c_int, _dp, _dp, C.c_int,
                                   _ip, C.c_int, _dp, _dp, C.c_int]
        L.cref_avgcorr.restype = C.c_int
        L.cref_hungarian.argtypes = [_dp, C.c_int, _ip]
        L.cref_hungarian.restype = C.c_int
        L.cref_rotate_ucell.argtypes = [_dp, _dp, _dp, _dp]
        L.cref_rotate_ucell.restype = C.c_int
        L.cref_ucell_from_xyzabg.argtypes = [_dp, _dp, _dp]
        L.cref_ucell_from_xyzabg.restype = C.c_int
        L.cref_dist2_image_nonortho.argtypes = [_dp, _dp, _dp, _dp]
        L.cref_dist2_image_nonortho.restype = C.c_double
        L.cref_symm_frames.argtypes = [_fp, C.c_int64, C.c_int, _ip, C.c_int, _dp, _dp, C.c_int,
                                       C.c_int, C.c_int, _ip, _ip, _dp, _ip, _dp, _dp]
        L.cref_symm_frames.restype = C.c_int

    def rmsd_centered_ref(self, tgt, ref, mass=None, ref_mass=None, use_mass=False):
        t = _c64(tgt).copy()
        U = np.zeros(9)
        tr = np.zeros(3)
        m = _c64(mass) if mass is not None else None
        rm = _c64(ref_mass) if ref_mass is not None else m
        r = self.L.cref_rmsd_centered_ref(_d(t), _d(_c64(ref)), _d(m), _d(rm), len(t),
                                          int(use_mass), _d(U), _d(tr))
        return r, U.reshape(3, 3), tr, t

    def rmsd(self, tgt, ref, mass=None, ref_mass=None, use_mass=False):
        t = _c64(tgt).copy()
        rf = _c64(ref).copy()
        U = np.zeros(9)
        tr = np.zeros(3)
        rtr = np.zeros(3)
        m = _c64(mass) if mass is not None else None
        rm = _c64(ref_mass) if ref_mass is not None else m
        r = self.L.cref_rmsd(_d(t), _d(rf), _d(m), _d(rm), len(t), int(use_mass), _d(U),
                             _d(tr), _d(rtr))
        return r, U.reshape(3, 3), tr, rtr

    def rmsd_nofit(self, tgt, ref, mass=None, use_mass=False):
        m = _c64(mass) if mass is not None else None
        return self.L.cref_rmsd_nofit(_d(_c64(tgt)), _d(_c64(ref)), _d(m), len(tgt),
                                      int(use_mass))

    def center_on_origin(self, xyz, mass=None, use_mass=False):
        x = _c64(xyz).copy()
        c = np.zeros(3)
        m = _c64(mass) if mass is not None else None
        self.L.cref_center_on_origin(_d(x), _d(m), len(x), int(use_mass), _d(c))
        return x, c

    def qrmsd_centered_ref(self, ref, tgt, mass=None, use_mass=False, want_rot=False):
        t = _c64(tgt).copy()
        tr = np.zeros(3)
        rot = np.zeros(9) if want_rot else None
        m = _c64(mass) if mass is not None else None
        r = self.L.cref_qrmsd_centered_ref(_d(_c64(ref)), _d(t), _d(m), len(t),
                                           int(use_mass), _d(rot), _d(tr))
        return r, (rot.reshape(3, 3) if want_rot else None), tr

    def qcp(self, A, E0, length, want_rot=True):
        rot = np.zeros(9) if want_rot else None
        rmsd = C.c_double(0.0)
        ret = self.L.cref_qcp(_d(rot), _d(_c64(A).ravel()), C.byref(rmsd), E0, length)
        return ret, rmsd.value, (rot.reshape(3, 3) if want_rot else None)

    def diagonalize_sort(self, m):
        E = np.zeros(9)
        ev = np.zeros(3)
        err = self.L.cref_diagonalize_sort(_d(_c64(m).ravel()), _d(E), _d(ev))
        return err, E.reshape(3, 3), ev

    def rms2d(self, frames, sel=None, mass=None, mode=MODE_FIT, use_mass=False,
              ref_frames=None, ref_sel=None, ref_mass=None, full=None, nthreads=0,
              rows=None, want_output=True):
        tf = _c32(frames)
        ntgt = tf.shape[0]
        tsel = _ci(sel)
        rsel = _ci(ref_sel) if ref_sel is not None else tsel
        nsel = len(tsel) if tsel is not None else tf.shape[1] // 3
        rf = _c32(ref_frames) if ref_frames is not None else None
        nref = rf.shape[0] if rf is not None else ntgt
        if full is None:
            full = rf is not None or (ref_sel is not None and not np.array_equal(ref_sel, sel))
        tm = _c64(mass) if mass is not None else np.ones(nsel)
        rm = _c64(ref_mass) if ref_mass is not None else tm
        n_out = nref * ntgt if full else ntgt * (ntgt - 1) // 2
        out = np.zeros(n_out, dtype=np.float32) if want_output else None
        r0, r1 = rows if rows is not None else (0, -1)
        self.L.cref_rms2d(_f(tf), tf.shape[1], ntgt, _f(rf),
                          rf.shape[1] if rf is not None else 0, nref, _i(tsel), _i(rsel),
                          nsel, _d(tm), _d(rm), mode, int(use_mass), int(bool(full)),
                          _f(out), nthreads, r0, r1)
        return out

    def rms2d_rows(self, frames, rows, sel=None, mass=None, mode=MODE_FIT, use_mass=False,
                   nthreads=0, want_output=True):
        fr = _c32(frames)
        F = fr.shape[0]
        rw = np.ascontiguousarray(rows, dtype=np.int32)
        s = _ci(sel)
        nsel = len(s) if s is not None else fr.shape[1] // 3
        m = _c64(mass) if mass is not None else np.ones(nsel)
        n_out = int(sum(F - 1 - int(r) for r in rw))
        out = np.zeros(n_out, dtype=np.float32) if want_output else None
        self.L.cref_rms2d_rows(_f(fr), fr.shape[1], F, _i(s), nsel, _d(m), mode, int(use_mass),
                               _i(rw), len(rw), _f(out), nthreads)
        return out if want_output else n_out

    def fit_frames(self, frames, ref_sel_centered, ref_trans, sel=None, mass=None,
                   fit=True, use_mass=False, modify=0, want_frames=False, nthreads=0):
        fr = _c32(frames)
        nf = fr.shape[0]
        natom = fr.shape[1] // 3
        s = _ci(sel)
        nsel = len(s) if s is not None else natom
        m = _c64(mass) if mass is not None else np.ones(nsel)
        rmsd = np.zeros(nf)
        U = np.zeros((nf, 9))
        tv = np.zeros((nf, 6))
        fo = np.zeros_like(fr) if want_frames else None
        self.L.cref_fit_frames(_f(fr), fr.shape[1], nf, natom, _i(s), nsel, _d(m),
                               _d(_c64(ref_sel_centered)), _d(_c64(ref_trans)), int(fit),
                               int(use_mass), modify, _d(rmsd), _d(U), _d(tv), _f(fo),
                               nthreads)
        return rmsd, U.reshape(nf, 3, 3), tv, fo

    def fit_pairs(self, frames, ref_frames, sel=None, mass=None, use_mass=False, nthreads=0):
        """'reftraj': frame f against reference frame f (centred per pair), nomod."""
        fr = _c32(frames)
        rf = _c32(ref_frames)
        assert fr.shape == rf.shape
        nf, natom = fr.shape[0], fr.shape[1] // 3
        s = _ci(sel)
        nsel = len(s) if s is not None else natom
        m = _c64(mass) if mass is not None else np.ones(nsel)
        rmsd = np.zeros(nf)
        self.L.cref_fit_pairs(_f(fr), _f(rf), fr.shape[1], nf, _i(s), nsel, _d(m),
                              int(use_mass), _d(rmsd), nthreads)
        return rmsd

    def perres(self, frames, ref, res_off, tgt_atoms, ref_atoms=None, tmass=None,
               rmass=None, use_mass=False, center=False, U=None, tvec=None, modify=0):
        fr = _c32(frames)
        nf = fr.shape[0]
        rf = _c32(ref).reshape(-1)
        off = _ci(res_off)
        nres = len(off) - 1
        ta = _ci(tgt_atoms)
        ra = _ci(ref_atoms) if ref_atoms is not None else ta
        tm = _c64(tmass) if tmass is not None else None
        rm = _c64(rmass) if rmass is not None else tm
        Uc = _c64(U).reshape(nf, 9) if U is not None else None
        tv = _c64(tvec).reshape(nf, 6) if tvec is not None else None
        out = np.zeros((nf, nres))
        self.L.cref_perres(_f(fr), fr.shape[1], nf, fr.shape[1] // 3, _f(rf), len(rf) // 3,
                           nres, _i(off), _i(ta), _i(ra), _d(tm), _d(rm), int(use_mass),
                           int(center), _d(Uc), _d(tv), modify, _d(out))
        return out

    def avgcorr(self, frames, ws, sel=None, mass=None, ref_sel_centered=None, use_mass=False,
                nthreads=1):
        fr = _c32(frames)
        s = _ci(sel)
        nsel = len(s) if s is not None else fr.shape[1] // 3
        m = _c64(mass) if mass is not None else np.ones(nsel)
        w = _ci(ws)
        ct, sd = np.zeros(len(w)), np.zeros(len(w))
        rs = None if ref_sel_centered is None else _c64(ref_sel_centered)
        self.L.cref_avgcorr(_f(fr), fr.shape[1], fr.shape[0], _i(s), nsel, _d(m), _d(rs),
                            int(use_mass), _i(w), len(w), _d(ct), _d(sd), nthreads)
        return ct, sd

    def rotate_ucell(self, U, ucell):
        """Compiled Box::SetupFromUcell + RotateUcell: (unit cell, fractional cell)."""
        u, f = np.zeros(9), np.zeros(9)
        if self.L.cref_rotate_ucell(_d(_c64(ucell).ravel()), _d(_c64(U).ravel()), _d(u), _d(f)):
            raise ValueError("no box")
        return u.reshape(3, 3), f.reshape(3, 3)

    def ucell_from_xyzabg(self, box):
        u, f = np.zeros(9), np.zeros(9)
        if self.L.cref_ucell_from_xyzabg(_d(_c64(box).ravel()), _d(u), _d(f)):
            raise ValueError("no box")
        return u.reshape(3, 3), f.reshape(3, 3)

    def dist_image(self, a, b, ucell, frac):
        """Minimum-image distance (compiled DIST2_ImageNonOrtho)."""
        return float(np.sqrt(self.L.cref_dist2_image_nonortho(
            _d(_c64(a).ravel()), _d(_c64(b).ravel()), _d(_c64(ucell).ravel()),
            _d(_c64(frac).ravel()))))

    def hungarian(self, M):
        """The reference's Hungarian::Optimize: map[col] = row."""
        return _hungarian(self.L.cref_hungarian, M)

    def symm(self, frames, sel=None, mass=None, ref_sel=None, use_mass=False, fit=True,
             groups=()):
        """SymmRMSD_CenteredRef's steps on the reference's classes."""
        return _symm(self.L.cref_symm_frames, frames, sel, mass, ref_sel, use_mass, fit, groups)


def ref_available() -> bool:
    return os.path.exists(REF_SO)

