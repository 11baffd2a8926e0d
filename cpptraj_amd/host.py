"""ctypes binding of the C++ host layer (libcpptraj_rmsd_host.so).

A HostSession holds COORDS sets and reference frames in host memory (FP32,
as cpptraj keeps them), a mask table, and runs command lines in the
reference's syntax:

    s = HostSession()
    s.add_coords("crd1", frames, natom, masses)
    s.add_mask(":3-7", sel)
    s.run("rms2d crd1 :3-7 rmsout rmsd.dat")
    tri = s.matrix("crd1")

The commands execute on the GPU through libcpptraj_rmsd.so; there is no CPU
path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import RmsdError, library_path, load_library

_HOST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libcpptraj_rmsd_host.so")
_hl = None


def load_host():
    global _hl
    if _hl is not None:
        return _hl
    load_library()  # the kernels library first (same file the host library links)
    if not os.path.exists(_HOST):
        raise RmsdError(f"{_HOST} is missing: build it with python -m cpptraj_amd.build")
    L = C.CDLL(_HOST)
    P, I, I64 = C.c_void_p, C.c_int, C.c_int64
    FP, DP, IP = C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_int)
    L.chost_create.argtypes = [I, C.POINTER(P)]
    L.chost_destroy.argtypes = [P]
    L.chost_destroy.restype = None
    L.chost_last_error.argtypes = [P]
    L.chost_last_error.restype = C.c_char_p
    L.chost_add_coords.argtypes = [P, C.c_char_p, FP, I64, I64, I, DP]
    L.chost_add_reference.argtypes = [P, C.c_char_p, FP, I, DP]
    L.chost_add_mask.argtypes = [P, C.c_char_p, C.c_char_p, IP, I]
    L.chost_run.argtypes = [P, C.c_char_p]
    L.chost_get_matrix.argtypes = [P, C.c_char_p, FP, C.POINTER(I64), C.POINTER(I64), IP]
    L.chost_get_series.argtypes = [P, C.c_char_p, DP, C.POINTER(I64), IP]
    L.chost_get_coords.argtypes = [P, C.c_char_p, FP, C.POINTER(I64)]
    L.chost_set_box.argtypes = [P, C.c_char_p, DP, I64]
    L.chost_get_box.argtypes = [P, C.c_char_p, DP, C.POINTER(I64)]
    L.chost_read_cmatrix.argtypes = [C.c_char_p, FP, C.POINTER(I64), C.POINTER(I64), IP, IP]
    L.chost_write_matrix.argtypes = [C.c_char_p, FP, I64, I64, I, I, C.c_char_p]
    L.chost_write_series.argtypes = [C.c_char_p, I, C.POINTER(C.c_char_p), IP, IP, IP,
                                     C.POINTER(DP), I64]
    L.chost_write_cmatrix.argtypes = [C.c_char_p, FP, I64, I64, I, IP]
    L.chost_set_residues.argtypes = [P, C.c_char_p, I, IP, C.POINTER(C.c_char_p), I, IP]
    L.chost_set_symmetric_groups.argtypes = [P, C.c_char_p, I, IP, IP]
    L.chost_write_grace.argtypes = [C.c_char_p, I, C.POINTER(C.c_char_p), C.POINTER(DP), I64, I]
    L.chost_write_mesh.argtypes = [C.c_char_p, C.c_char_p, DP, I64, I, C.POINTER(C.c_char_p),
                                   C.POINTER(DP)]
    for f in ("chost_write_matrix", "chost_write_series", "chost_write_cmatrix",
              "chost_write_grace", "chost_write_mesh"):
        getattr(L, f).restype = I
    for f in ("chost_create", "chost_add_coords", "chost_add_reference", "chost_add_mask",
              "chost_run", "chost_get_matrix", "chost_get_series", "chost_get_coords",
              "chost_read_cmatrix", "chost_set_residues", "chost_set_symmetric_groups",
              "chost_set_box", "chost_get_box"):
        getattr(L, f).restype = I
    _hl = L
    return L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


class HostSession:
    def __init__(self, device: int = 0):
        self._L = load_host()
        h = C.c_void_p()
        rc = self._L.chost_create(device, C.byref(h))
        self._h = h
        if rc:
            msg = self._L.chost_last_error(h).decode() if h else "chost_create failed"
            self.close()
            raise RmsdError(msg)

    def close(self):
        if getattr(self, "_h", None):
            self._L.chost_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise RmsdError(self._L.chost_last_error(self._h).decode())

    def add_coords(self, name, frames, natom, mass=None, stride=None):
        x = np.ascontiguousarray(frames, dtype=np.float32).reshape(len(frames), -1)
        st = x.shape[1] if stride is None else stride
        m = None if mass is None else np.ascontiguousarray(mass, dtype=np.float64)
        self._check(self._L.chost_add_coords(self._h, name.encode(), _fp(x), x.shape[0], st,
                                             natom, _dp(m)))

    def set_box(self, name, ucell):
        """Unit cells of a COORDS set's frames: (nframes, 3, 3), rows = cell vectors."""
        u = np.ascontiguousarray(ucell, dtype=np.float64).reshape(-1, 9)
        self._check(self._L.chost_set_box(self._h, name.encode(), _dp(u), u.shape[0]))

    def box(self, name):
        n = C.c_int64()
        self._check(self._L.chost_get_box(self._h, name.encode(), None, C.byref(n)))
        out = np.zeros((max(n.value, 1), 9))
        self._check(self._L.chost_get_box(self._h, name.encode(), _dp(out), None))
        return out[: n.value].reshape(-1, 3, 3)

    def add_reference(self, name, frame, natom, mass=None):
        x = np.ascontiguousarray(frame, dtype=np.float32).reshape(-1)
        m = None if mass is None else np.ascontiguousarray(mass, dtype=np.float64)
        self._check(self._L.chost_add_reference(self._h, name.encode(), _fp(x), natom, _dp(m)))

    def set_residues(self, name, first, names, solute=None):
        """Residues of a set's topology: first atoms (0-based), names, and the
        solute residues (0-based; default every residue)."""
        f = np.ascontiguousarray(first, dtype=np.int32)
        nm = (C.c_char_p * max(len(f), 1))(*[str(n).encode() for n in names])
        sol = None if solute is None else np.ascontiguousarray(solute, dtype=np.int32)
        IP = C.POINTER(C.c_int)
        self._check(self._L.chost_set_residues(self._h, name.encode(), len(f),
                                               f.ctypes.data_as(IP), nm,
                                               0 if sol is None else len(sol),
                                               None if sol is None else sol.ctypes.data_as(IP)))

    def set_symmetric_groups(self, name, groups):
        """Symmetric atom groups of a set's topology (lists of 0-based atom
        indices, residue by residue)."""
        off = np.zeros(len(groups) + 1, np.int32)
        for k, g in enumerate(groups):
            off[k + 1] = off[k] + len(g)
        flat = np.ascontiguousarray(
            np.concatenate([np.asarray(g, np.int32) for g in groups]) if len(groups)
            else np.zeros(1, np.int32), dtype=np.int32)
        IP = C.POINTER(C.c_int)
        self._check(self._L.chost_set_symmetric_groups(self._h, name.encode(), len(groups),
                                                       off.ctypes.data_as(IP),
                                                       flat.ctypes.data_as(IP)))

    def add_mask(self, expr, sel, set_name=None):
        s = np.ascontiguousarray(sel, dtype=np.int32)
        self._check(self._L.chost_add_mask(self._h, None if set_name is None else set_name.encode(),
                                           expr.encode(), s.ctypes.data_as(C.POINTER(C.c_int)),
                                           len(s)))

    def run(self, line):
        self._check(self._L.chost_run(self._h, line.encode()))

    def matrix(self, name):
        nc, nr, tri = C.c_int64(), C.c_int64(), C.c_int()
        self._check(self._L.chost_get_matrix(self._h, name.encode(), None, C.byref(nc),
                                             C.byref(nr), C.byref(tri)))
        n = nc.value * (nc.value - 1) // 2 if tri.value else nc.value * nr.value
        out = np.zeros(max(n, 1), dtype=np.float32)
        self._check(self._L.chost_get_matrix(self._h, name.encode(), _fp(out), None, None, None))
        out = out[:n]
        return out if tri.value else out.reshape(nr.value, nc.value)

    def series(self, name):
        n, k = C.c_int64(), C.c_int()
        self._check(self._L.chost_get_series(self._h, name.encode(), None, C.byref(n),
                                             C.byref(k)))
        out = np.zeros(max(n.value * k.value, 1))
        self._check(self._L.chost_get_series(self._h, name.encode(), _dp(out), None, None))
        out = out[: n.value * k.value]
        return out if k.value == 1 else out.reshape(n.value, k.value)

    def coords(self, name):
        n = C.c_int64()
        self._check(self._L.chost_get_coords(self._h, name.encode(), None, C.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=np.float32)
        self._check(self._L.chost_get_coords(self._h, name.encode(), _fp(out), None))
        return out[: n.value]


def read_cmatrix(path):
    """(triangle float32, actual_nrows, sieve, cached frame list) of a Cmatrix file."""
    L = load_host()
    nr, an, sv = C.c_int64(), C.c_int64(), C.c_int()
    if L.chost_read_cmatrix(path.encode(), None, C.byref(nr), C.byref(an), C.byref(sv), None):
        raise RmsdError(f"cannot read {path}")
    n = nr.value
    tri = np.zeros(max(n * (n - 1) // 2, 1), dtype=np.float32)
    fr = np.zeros(max(n, 1), dtype=np.int32)
    if L.chost_read_cmatrix(path.encode(), _fp(tri), None, None, None,
                            fr.ctypes.data_as(C.POINTER(C.c_int))):
        raise RmsdError(f"cannot read {path}")
    return tri[: n * (n - 1) // 2], an.value, sv.value, fr[:n]


def write_matrix(path, data, ncols, nrows=None, tri=True, square2d=True, name=""):
    """DataIO_Std text of a float matrix (host layer writer, no GPU needed)."""
    L = load_host()
    a = np.ascontiguousarray(data, dtype=np.float32).reshape(-1)
    if L.chost_write_matrix(path.encode(), _fp(a), ncols, ncols if nrows is None else nrows,
                            int(tri), int(square2d), name.encode()):
        raise RmsdError(f"cannot write {path}")


def write_series(path, sets):
    """sets: list of (name, values (nframes, ncomp), width, precision)."""
    L = load_host()
    n = len(sets)
    vals = [np.ascontiguousarray(np.asarray(v, np.float64).reshape(len(v), -1)) for _, v, _, _ in sets]
    names = (C.c_char_p * n)(*[s[0].encode() for s in sets])
    ncomp = (C.c_int * n)(*[v.shape[1] for v in vals])
    width = (C.c_int * n)(*[s[2] for s in sets])
    prec = (C.c_int * n)(*[s[3] for s in sets])
    ptrs = (C.POINTER(C.c_double) * n)(*[_dp(v) for v in vals])
    nf = vals[0].shape[0] if vals else 0
    if L.chost_write_series(path.encode(), n, names, ncomp, width, prec, ptrs, nf):
        raise RmsdError(f"cannot write {path}")


def write_cmatrix(path, tri, nrows, actual_nrows, sieve, frames):
    L = load_host()
    t = np.ascontiguousarray(tri, dtype=np.float32)
    fr = np.ascontiguousarray(frames, dtype=np.int32)
    if L.chost_write_cmatrix(path.encode(), _fp(t), nrows, actual_nrows, sieve,
                             fr.ctypes.data_as(C.POINTER(C.c_int))):
        raise RmsdError(f"cannot write {path}")


def write_grace(path, legends, values, invert=False):
    """DataIO_Grace xy text of per-frame series (values: (nsets, nframes))."""
    L = load_host()
    v = [np.ascontiguousarray(x, dtype=np.float64) for x in values]
    n = len(v)
    lg = (C.c_char_p * n)(*[s.encode() for s in legends])
    ptrs = (C.POINTER(C.c_double) * n)(*[_dp(x) for x in v])
    if L.chost_write_grace(path.encode(), n, lg, ptrs, len(v[0]) if v else 0, int(invert)):
        raise RmsdError(f"cannot write {path}")


def write_mesh(path, xlabel, x, names, values):
    """DataIO_Std text of XY mesh sets sharing the x column."""
    L = load_host()
    xa = np.ascontiguousarray(x, dtype=np.float64)
    v = [np.ascontiguousarray(y, dtype=np.float64) for y in values]
    n = len(v)
    nm = (C.c_char_p * max(n, 1))(*[s.encode() for s in names])
    ptrs = (C.POINTER(C.c_double) * max(n, 1))(*[_dp(y) for y in v])
    if L.chost_write_mesh(path.encode(), xlabel.encode(), _dp(xa), len(xa), n, nm, ptrs):
        raise RmsdError(f"cannot write {path}")
