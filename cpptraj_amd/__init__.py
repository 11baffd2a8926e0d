"""cpptraj_amd -- MI355X (gfx950) best-fit RMSD for cpptraj's rms2d / cluster / rmsd path.

The product is the C ABI library ``_lib/libcpptraj_rmsd.so`` (include/cpptraj_rmsd.h).
This module is a thin ctypes binding over it for tests, benchmarks and Python
callers.  It mirrors the reference's command options:

* ``rms2d(...)``         Analysis_Rms2d (src/Analysis_Rms2d.cpp): mask / refmask,
                         ``mass``, ``nofit`` / ``qrmsd``, reference trajectory,
                         triangle or full matrix.
* ``cluster_cache(...)`` Cluster::MetricArray::CacheDistances
                         (src/Cluster/MetricArray.cpp:722-801).
* ``rmsd_fit(...)``      Action_Rmsd (src/Action_Rmsd.cpp): ``nofit``, ``mass``,
                         ``norotate`` / ``nomod``, ``savematrices``, ``savevectors``.

Tensors are torch CUDA tensors (HIP on ROCm).  There is no CPU fallback: if the
library is missing or no GPU is visible every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

__all__ = ["RmsdError", "Context", "Batch", "frames_desc", "frame_layout", "rms2d",
           "rms2d_batch", "cluster_cache", "rmsd_fit", "library_path", "load_library", "MODES", "PRECISIONS",
           "MODIFY", "BATCH_GRAPH", "BATCH_AUTOTUNE"]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_lib", "libcpptraj_rmsd.so")

MODES = {"fit": 0, "nofit": 1, "qrmsd": 2}
PRECISIONS = {"fp64": 0, "fp32": 1, "bf16x3": 2}
MODIFY = {"nomod": 0, "rottrans": 1, "norotate": 2}
BATCH_GRAPH = 1
BATCH_AUTOTUNE = 2


class RmsdError(RuntimeError):
    pass


class Frames(C.Structure):
    """rmsd_frames (include/cpptraj_rmsd.h)."""
    _fields_ = [
        ("xyz", C.c_void_p),
        ("stride", C.c_int64),
        ("ld", C.c_int64),
        ("nframes", C.c_int32),
        ("natom", C.c_int32),
        ("sel", C.c_void_p),
        ("nsel", C.c_int32),
        ("mass", C.c_void_p),
        ("fidx", C.c_void_p),
    ]


_lib = None


def library_path() -> str:
    return _LIB_PATH


def load_library():
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise RmsdError(f"{_LIB_PATH} is missing: build it with python -m cpptraj_amd.build")
    # One HIP runtime per process: torch ships its own libamdhip64 and asks for
    # it by file name, so if this library were loaded first (binding the ROCm
    # install's copy by soname) torch would load a second runtime next to it and
    # find no device.  Loading torch first makes this library bind to torch's.
    import torch  # noqa: F401
    L = C.CDLL(_LIB_PATH)
    P = C.c_void_p
    L.rmsd_ctx_create.argtypes = [C.POINTER(P), C.c_int, P]
    L.rmsd_ctx_create.restype = C.c_int
    L.rmsd_ctx_destroy.argtypes = [P]
    L.rmsd_ctx_destroy.restype = None
    L.rmsd_last_error.argtypes = [P]
    L.rmsd_last_error.restype = C.c_char_p
    L.rmsd_ctx_synchronize.argtypes = [P]
    L.rmsd_ctx_synchronize.restype = C.c_int
    L.rmsd_ctx_scratch_bytes.argtypes = [P]
    L.rmsd_ctx_scratch_bytes.restype = C.c_int64
    L.rmsd_ctx_set_profiling.argtypes = [P, C.c_int]
    L.rmsd_ctx_set_profiling.restype = C.c_int
    L.rmsd_ctx_stage_ms.argtypes = [P, C.POINTER(C.c_double)]
    L.rmsd_ctx_stage_ms.restype = C.c_int
    L.rmsd_ctx_set_tolerance.argtypes = [P, C.c_double]
    L.rmsd_ctx_set_tolerance.restype = C.c_int
    L.rmsd_ctx_error_report.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    L.rmsd_ctx_error_report.restype = C.c_int
    FP = C.POINTER(Frames)
    L.rmsd_pairwise.argtypes = [P, FP, FP, C.c_int, C.c_int, C.c_int, P]
    L.rmsd_pairwise.restype = C.c_int
    L.rmsd_cluster_cache.argtypes = [P, FP, C.c_int, C.c_int, C.c_int, C.c_double, P]
    L.rmsd_cluster_cache.restype = C.c_int
    L.rmsd_fit_to_ref.argtypes = [P, FP, FP, C.c_int, C.c_int, C.c_int, C.c_int, P, P, P]
    L.rmsd_fit_to_ref.restype = C.c_int
    L.rmsd_fit_pairs.argtypes = [P, FP, FP, C.c_int, C.c_int, C.c_int, P, P, P]
    L.rmsd_fit_pairs.restype = C.c_int
    _optional(L)
    _lib = L
    return L


def _optional(L):
    """Entry points that later parts of the library add (batch, backward)."""
    P = C.c_void_p
    FP = C.POINTER(Frames)
    if hasattr(L, "rmsd_batch_create"):
        L.rmsd_batch_create.argtypes = [P, FP, C.c_int, C.c_int, C.c_int, C.c_int, P, C.c_int,
                                        C.POINTER(P)]
        L.rmsd_batch_create.restype = C.c_int
        L.rmsd_batch_run.argtypes = [P]
        L.rmsd_batch_run.restype = C.c_int
        L.rmsd_batch_set_variant.argtypes = [P, C.c_int]
        L.rmsd_batch_set_variant.restype = C.c_int
        L.rmsd_batch_info.argtypes = [P, C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        L.rmsd_batch_info.restype = C.c_int
        L.rmsd_batch_destroy.argtypes = [P]
        L.rmsd_batch_destroy.restype = None
        L.rmsd_pairwise_batch.argtypes = [P, FP, C.c_int, C.c_int, C.c_int, C.c_int, P]
        L.rmsd_pairwise_batch.restype = C.c_int
    if hasattr(L, "rmsd_fit_backward"):
        L.rmsd_fit_backward.argtypes = [P, FP, FP, C.c_int, C.c_int, P, P]
        L.rmsd_fit_backward.restype = C.c_int
    if hasattr(L, "rmsd_pairwise_backward"):
        L.rmsd_pairwise_backward.argtypes = [P, FP, FP, C.c_int, P, P, P]
        L.rmsd_pairwise_backward.restype = C.c_int
    if hasattr(L, "rmsd_avgcorr"):
        L.rmsd_avgcorr_windows.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int32),
                                           C.c_int]
        L.rmsd_avgcorr_windows.restype = C.c_int
        L.rmsd_avgcorr.argtypes = [P, FP, FP, C.c_int, C.c_int, C.c_int, C.c_int, P, P]
        L.rmsd_avgcorr.restype = C.c_int
        L.rmsd_avgcorr_set_method.argtypes = [P, C.c_int]
        L.rmsd_avgcorr_set_method.restype = C.c_int
    if hasattr(L, "rmsd_perres_nofit"):
        IP = C.POINTER(C.c_int32)
        DP = C.POINTER(C.c_double)
        L.rmsd_perres_nofit.argtypes = [P, FP, FP, C.c_int, C.c_int, IP, IP, IP, DP, DP,
                                        C.c_int, C.c_int, C.c_int, P, P, P]
        L.rmsd_perres_nofit.restype = C.c_int
    if hasattr(L, "rmsd_rotate_ucell"):
        L.rmsd_rotate_ucell.argtypes = [P, P, C.c_int, P]
        L.rmsd_rotate_ucell.restype = C.c_int
    if hasattr(L, "rmsd_symm"):
        IP = C.POINTER(C.c_int32)
        L.rmsd_symm.argtypes = [P, FP, FP, C.c_int, C.c_int, IP, IP, C.c_int, C.c_int, P, P, P,
                                P]
        L.rmsd_symm.restype = C.c_int


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def frames_desc(xyz, sel=None, mass=None, fidx=None, soa=False, nsel=None, natom=None):
    """Builds an rmsd_frames for a CUDA tensor.

    The descriptor holds raw device pointers, not references: the caller keeps
    xyz, sel, mass and fidx alive until the library call that uses it has been
    enqueued (a freed tensor's block can be handed to the next allocation).

    xyz: AoS (F, natom*3) or (F, natom, 3) float32, or SoA (F, 3, ld) with soa=True;
         natom atoms are used of each (padded) row (default: the row length / 3
         for AoS, ld for SoA).
    sel: int32 CUDA tensor of selected atoms; mass: float64 CUDA tensor (nsel)."""
    import torch

    if not xyz.is_cuda or xyz.dtype != torch.float32:
        raise RmsdError("frames must be a float32 CUDA tensor")
    stride, ld, natom = frame_layout(xyz, soa, natom)
    for t, dt in ((sel, torch.int32), (mass, torch.float64), (fidx, torch.int32)):
        if t is not None and (not t.is_cuda or t.dtype != dt or not t.is_contiguous()):
            raise RmsdError("sel/fidx must be contiguous int32 and mass float64 CUDA tensors")
    if nsel is None:
        nsel = sel.numel() if sel is not None else natom
    nfr = fidx.numel() if fidx is not None else xyz.shape[0]
    d = Frames(xyz.data_ptr(), stride, ld, nfr, natom, None if sel is None else sel.data_ptr(),
               nsel, None if mass is None else mass.data_ptr(),
               None if fidx is None else fidx.data_ptr())
    return d


def frame_layout(xyz, soa=False, natom=None):
    """(frame stride, ld, natom) of a frame tensor as the library will read it.

    The library reads the tensor's own storage through (data_ptr, stride, ld),
    so the layout is checked on xyz itself: a view whose rows overlap or
    whose atoms are not packed is rejected, never silently copied."""
    F = xyz.shape[0]
    empty = xyz.numel() == 0
    if soa:
        if xyz.dim() != 3 or xyz.shape[1] != 3 or (
                not empty and (xyz.stride(2) != 1 or xyz.stride(1) < xyz.shape[2]
                               or (F > 1 and xyz.stride(0) < 3 * xyz.stride(1)))):
            raise RmsdError("SoA frames must be (F, 3, ld) with unit atom stride, "
                            "ld >= atoms and non-overlapping frames")
        ld = xyz.stride(1) if not empty else max(xyz.shape[2], 1)
        natom = xyz.shape[2] if natom is None else natom
        stride = xyz.stride(0) if F > 1 else 3 * ld
    else:
        if xyz.dim() == 3:
            if xyz.shape[2] != 3:
                raise RmsdError("AoS frames must be (F, natom*3) or (F, natom, 3)")
            if not empty and (xyz.stride(2) != 1 or (xyz.shape[1] > 1 and xyz.stride(1) != 3)):
                raise RmsdError("AoS frames (F, natom, 3) need packed atoms (strides (., 3, 1))")
            row = 3 * xyz.shape[1]
        elif xyz.dim() == 2:
            if not empty and xyz.shape[1] > 1 and xyz.stride(1) != 1:
                raise RmsdError("AoS frames need unit coordinate stride")
            row = xyz.shape[1]
        else:
            raise RmsdError("AoS frames must be (F, natom*3) or (F, natom, 3)")
        ld = 0
        if natom is None:
            natom = row // 3
        elif 3 * natom > row:
            raise RmsdError(f"{natom} atoms do not fit AoS rows of {row} floats")
        if not empty and F > 1 and xyz.stride(0) < row:
            raise RmsdError("AoS frames overlap (frame stride below the row length)")
        stride = xyz.stride(0) if F > 1 else 3 * natom
    return stride, ld, natom


class Context:
    """An rmsd_ctx bound to the current torch stream of `device`."""

    def __init__(self, device: int = 0, stream=None):
        import torch

        L = load_library()
        if not torch.cuda.is_available():
            raise RmsdError("no GPU visible: the RMSD library has no CPU path")
        self.device = device
        if stream is None:
            stream = torch.cuda.current_stream(device)
        self.stream = stream
        h = C.c_void_p()
        rc = L.rmsd_ctx_create(C.byref(h), device, C.c_void_p(stream.cuda_stream))
        if rc != 0:
            raise RmsdError(f"rmsd_ctx_create failed ({rc})")
        self._h = h
        self._L = L
        self._plans = weakref.WeakSet()

    def close(self):
        """Destroys the context and every batch plan still bound to it."""
        for p in list(getattr(self, "_plans", ())):
            p.close()
        if getattr(self, "_h", None):
            self._L.rmsd_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        if rc != 0:
            msg = self._L.rmsd_last_error(self._h).decode()
            raise RmsdError(f"rc={rc}: {msg}")

    @property
    def handle(self):
        return self._h

    def synchronize(self):
        self.check(self._L.rmsd_ctx_synchronize(self._h))

    def scratch_bytes(self) -> int:
        return int(self._L.rmsd_ctx_scratch_bytes(self._h))

    def set_profiling(self, on: bool = True):
        self.check(self._L.rmsd_ctx_set_profiling(self._h, int(on)))

    def set_tolerance(self, tol: float):
        """Error tolerance of the mixed-precision pairwise forms (Angstrom): pairs
        whose stated bound exceeds it are recomputed in FP64; 0 switches the
        bound off, a negative value reports the bound without recomputing."""
        self.check(self._L.rmsd_ctx_set_tolerance(self._h, float(tol)))

    def error_report(self):
        """(largest stated bound of the results kept in mixed precision,
        pairs recomputed in FP64) for the last pairwise / cluster-cache call."""
        b = C.c_double()
        n = C.c_int64()
        self.check(self._L.rmsd_ctx_error_report(self._h, C.byref(b), C.byref(n)))
        return b.value, n.value

    def stage_ms(self):
        """(per-frame stage, main kernel, epilogue) milliseconds of the last call."""
        ms = (C.c_double * 3)()
        self.check(self._L.rmsd_ctx_stage_ms(self._h, ms))
        return tuple(ms)

    # -- raw entry points on descriptors ----------------------------------------
    def pairwise_desc(self, tgt: Frames, ref: Frames | None, mode: int, use_mass: bool,
                      precision: int, out):
        self.check(self._L.rmsd_pairwise(self._h, C.byref(tgt),
                                         None if ref is None else C.byref(ref), mode,
                                         int(use_mass), precision, _ptr(out)))

    def cluster_desc(self, frames: Frames, nofit: bool, use_mass: bool, precision: int,
                     weight: float, out):
        self.check(self._L.rmsd_cluster_cache(self._h, C.byref(frames), int(nofit),
                                              int(use_mass), precision, weight, _ptr(out)))

    def fit_desc(self, frames: Frames, ref: Frames, ref_frame: int, fit: bool, use_mass: bool,
                 modify: int, rmsd, U=None, tvec=None):
        self.check(self._L.rmsd_fit_to_ref(self._h, C.byref(frames), C.byref(ref), ref_frame,
                                           int(fit), int(use_mass), modify, _ptr(rmsd), _ptr(U),
                                           _ptr(tvec)))


_default_ctx = {}


def _ctx(device=0) -> Context:
    """The default context for `device` on torch's current stream there.

    A context enqueues its work on the stream it was made with, so the cache is
    keyed on (device, stream): a call made under another torch stream gets a
    context on that stream and stays ordered with the work that produced its
    inputs and will consume its outputs."""
    import torch

    stream = torch.cuda.current_stream(device)
    key = (device, stream.cuda_stream)
    c = _default_ctx.get(key)
    if c is None:
        c = Context(device, stream)
        _default_ctx[key] = c
    return c


def rms2d(frames, mask=None, mass=None, mode="fit", use_mass=False, precision="fp64",
          ref_frames=None, ref_mask=None, ref_mass=None, soa=False, natom=None,
          ctx: Context | None = None):
    """All-pairs RMSD (Analysis_Rms2d::Calculate_2D).

    Returns the packed upper triangle (same frames and mask) or the full
    (n_tgt, n_ref) matrix (reference trajectory or different masks) as a float32
    CUDA tensor."""
    import torch

    ctx = ctx or _ctx(frames.device.index or 0)
    tgt = frames_desc(frames, mask, mass, soa=soa, natom=natom)
    full = ref_frames is not None or (ref_mask is not None and (
        mask is None or ref_mask.numel() != mask.numel() or not torch.equal(ref_mask, mask)))
    if not full:
        F = tgt.nframes
        out = torch.empty(F * (F - 1) // 2, dtype=torch.float32, device=frames.device)
        ctx.pairwise_desc(tgt, None, MODES[mode], use_mass, PRECISIONS[precision], out)
        return out
    rf = ref_frames if ref_frames is not None else frames
    ref = frames_desc(rf, ref_mask if ref_mask is not None else mask,
                      ref_mass if ref_mass is not None else mass, soa=soa, natom=natom)
    out = torch.empty((tgt.nframes, ref.nframes), dtype=torch.float32, device=frames.device)
    ctx.pairwise_desc(tgt, ref, MODES[mode], use_mass, PRECISIONS[precision], out)
    return out


def cluster_cache(frames, frames_to_cache=None, mask=None, mass=None, nofit=False,
                  use_mass=False, precision="fp64", ctx: Context | None = None):
    """Pairwise-distance cache over `frames_to_cache` (int32 CUDA tensor of frame
    indices, or all frames): packed upper triangle, float32."""
    import torch

    ctx = ctx or _ctx(frames.device.index or 0)
    d = frames_desc(frames, mask, mass, fidx=frames_to_cache)
    n = d.nframes
    out = torch.empty(n * (n - 1) // 2, dtype=torch.float32, device=frames.device)
    ctx.cluster_desc(d, nofit, use_mass, PRECISIONS[precision], 1.0, out)
    return out


def rmsd_fit(frames, ref_frames=None, ref_index=0, mask=None, mass=None, ref_mask=None,
             ref_mass=None, fit=True, use_mass=False, modify="rottrans", save_matrices=False,
             save_vectors=False, soa=False, natom=None, ctx: Context | None = None):
    """Action_Rmsd over a batch of frames against one reference frame.

    modify: 'rottrans' (default action), 'norotate', 'nomod'; frames are
    rewritten in place unless 'nomod'.  Returns (rmsd, U or None, tvec or None)."""
    import torch

    ctx = ctx or _ctx(frames.device.index or 0)
    d = frames_desc(frames, mask, mass, soa=soa, natom=natom)
    rf = frames if ref_frames is None else ref_frames
    r = frames_desc(rf, ref_mask if ref_mask is not None else mask,
                    ref_mass if ref_mass is not None else mass, soa=soa, natom=natom)
    F = d.nframes
    dev = frames.device
    rmsd = torch.empty(F, dtype=torch.float64, device=dev)
    U = torch.empty((F, 3, 3), dtype=torch.float64, device=dev) if save_matrices else None
    tv = torch.empty((F, 6), dtype=torch.float64, device=dev) if save_vectors else None
    ctx.fit_desc(d, r, ref_index, fit, use_mass, MODIFY[modify], rmsd, U, tv)
    return rmsd, U, tv


def rmsd_fit_pairs(frames, ref_frames, mask=None, mass=None, ref_mask=None, ref_mass=None,
                   fit=True, use_mass=False, modify="nomod", save_matrices=False,
                   save_vectors=False, soa=False, natom=None, ctx: Context | None = None):
    """Frame f of `frames` fitted to frame f of `ref_frames` (rmsd_fit_pairs:
    the 'reftraj' reference mode; 'previous' on frames as stored is
    ref_frames = frames shifted by one).  Same outputs as rmsd_fit."""
    import torch

    ctx = ctx or _ctx(frames.device.index or 0)
    d = frames_desc(frames, mask, mass, soa=soa, natom=natom)
    r = frames_desc(ref_frames, ref_mask if ref_mask is not None else mask,
                    ref_mass if ref_mass is not None else mass, soa=soa, natom=natom)
    F = d.nframes
    dev = frames.device
    rmsd = torch.empty(F, dtype=torch.float64, device=dev)
    U = torch.empty((F, 3, 3), dtype=torch.float64, device=dev) if save_matrices else None
    tv = torch.empty((F, 6), dtype=torch.float64, device=dev) if save_vectors else None
    ctx.check(ctx._L.rmsd_fit_pairs(ctx.handle, C.byref(d), C.byref(r), int(bool(fit)),
                                    int(bool(use_mass)), MODIFY[modify], _ptr(rmsd), _ptr(U),
                                    _ptr(tv)))
    return rmsd, U, tv


def rotate_ucell(U, ucell, ctx: Context | None = None):
    """Box::RotateUcell of the fit's update, in place: ucell (F, 3, 3) float64
    CUDA tensor of unit cells (rows = cell vectors), U (F, 3, 3) the fit's
    Kabsch rotations (rmsd_fit(..., save_matrices=True))."""
    import torch

    ctx = ctx or _ctx(ucell.device.index or 0)
    if (ucell.dtype != torch.float64 or U.dtype != torch.float64 or not ucell.is_contiguous()
            or ucell.shape[0] != U.shape[0]):
        raise RmsdError("U and ucell must be (F, 3, 3) float64, ucell contiguous")
    Uc = U.contiguous()
    ctx.check(ctx._L.rmsd_rotate_ucell(ctx.handle, _ptr(Uc), int(ucell.shape[0]), _ptr(ucell)))
    return ucell


def symm_groups(groups):
    """List of index lists -> (group_off, group_atoms) int32 host arrays."""
    import numpy as np

    off = np.zeros(len(groups) + 1, np.int32)
    for g, a in enumerate(groups):
        off[g + 1] = off[g] + len(a)
    flat = (np.concatenate([np.asarray(a, np.int32) for a in groups]) if len(groups)
            else np.zeros(0, np.int32))
    return off, np.ascontiguousarray(flat, dtype=np.int32)


def rmsd_symm(frames, groups, ref_frames=None, ref_index=0, mask=None, mass=None,
              ref_mask=None, ref_mass=None, fit=True, use_mass=False, want_map=False,
              want_xform=False, soa=False, natom=None, ctx: Context | None = None):
    """Symmetry-corrected RMSD ('symmrmsd', rmsd_symm) of every frame against
    one reference frame.  groups: lists of selection indices (or a
    (group_off, group_atoms) pair).  Returns (rmsd, amap or None, U or None,
    tvec or None) as CUDA tensors; amap is (F, nsel) int32."""
    import numpy as np
    import torch

    ctx = ctx or _ctx(frames.device.index or 0)
    d = frames_desc(frames, mask, mass, soa=soa, natom=natom)
    rf = frames if ref_frames is None else ref_frames
    r = frames_desc(rf, ref_mask if ref_mask is not None else mask,
                    ref_mass if ref_mass is not None else mass, soa=soa, natom=natom)
    if isinstance(groups, tuple) and len(groups) == 2 and isinstance(groups[0], np.ndarray):
        off = np.ascontiguousarray(groups[0], dtype=np.int32)
        flat = np.ascontiguousarray(groups[1], dtype=np.int32)
    else:
        off, flat = symm_groups(groups)
    F, dev = d.nframes, frames.device
    rmsd = torch.empty(F, dtype=torch.float64, device=dev)
    amap = torch.empty((F, d.nsel), dtype=torch.int32, device=dev) if want_map else None
    U = torch.empty((F, 3, 3), dtype=torch.float64, device=dev) if want_xform else None
    tv = torch.empty((F, 6), dtype=torch.float64, device=dev) if want_xform else None
    IP = C.POINTER(C.c_int32)
    ng = len(off) - 1
    ctx.check(ctx._L.rmsd_symm(ctx.handle, C.byref(d), C.byref(r), ref_index, ng,
                               off.ctypes.data_as(IP), (flat if len(flat) else
                                                        np.zeros(1, np.int32)).ctypes.data_as(IP),
                               int(fit), int(use_mass), _ptr(rmsd), _ptr(amap), _ptr(U), _ptr(tv)))
    return rmsd, amap, U, tv


def rmsd_perres(frames, res_off, tgt_atoms, ref_frames=None, ref_index=0, ref_atoms=None,
                tmass=None, rmass=None, use_mass=False, center=False, modify="nomod", U=None,
                tvec=None, soa=False, natom=None, ctx: Context | None = None):
    """Per-residue no-fit RMSD of the rmsd action ('perres').

    res_off (nres+1), tgt_atoms / ref_atoms (res_off[-1] atom indices) and the
    per-entry masses tmass / rmass are host arrays (numpy or lists).  modify
    'rottrans' / 'norotate' applies the fit's U and tvec (CUDA float64 tensors
    from rmsd_fit with modify='nomod') to the target in FP64 first.  Returns a
    (nframes, nres) float64 CUDA tensor."""
    import numpy as np
    import torch

    ctx = ctx or _ctx(frames.device.index or 0)
    d = frames_desc(frames, soa=soa, natom=natom)
    r = frames_desc(frames if ref_frames is None else ref_frames, soa=soa, natom=natom)
    off = np.ascontiguousarray(res_off, dtype=np.int32)
    ta = np.ascontiguousarray(tgt_atoms, dtype=np.int32)
    ra = ta if ref_atoms is None else np.ascontiguousarray(ref_atoms, dtype=np.int32)
    tm = None if tmass is None else np.ascontiguousarray(tmass, dtype=np.float64)
    rm = tm if rmass is None else np.ascontiguousarray(rmass, dtype=np.float64)
    nres = len(off) - 1
    out = torch.empty((d.nframes, max(nres, 0)), dtype=torch.float64, device=frames.device)
    IP = C.POINTER(C.c_int32)
    DP = C.POINTER(C.c_double)

    def cp(a, t):
        return None if a is None else a.ctypes.data_as(t)

    Ut = None if U is None else U.contiguous()
    tvt = None if tvec is None else tvec.contiguous()
    ctx.check(ctx._L.rmsd_perres_nofit(ctx.handle, C.byref(d), C.byref(r), ref_index, nres,
                                       cp(off, IP), cp(ta, IP), cp(ra, IP), cp(tm, DP),
                                       cp(rm, DP), int(use_mass), int(center), MODIFY[modify],
                                       _ptr(Ut), _ptr(tvt), _ptr(out)))
    return out


def avgcorr_windows(nframes, maxwindow=-1, offset=1):
    """Window sizes of the RMS average correlation (rmsd_avgcorr_windows)."""
    import numpy as np
    L = load_library()
    n = L.rmsd_avgcorr_windows(nframes, maxwindow, offset, None, 0)
    if n < 1:
        raise RmsdError("Not enough frames to perform calculation.")
    w = np.zeros(n, np.int32)
    L.rmsd_avgcorr_windows(nframes, maxwindow, offset, w.ctypes.data_as(C.POINTER(C.c_int32)), n)
    return w


AVGCORR_METHODS = {"auto": 0, "running": 1, "prefix": 2}


def rms_avgcorr(frames, mask=None, mass=None, ref_frames=None, ref_index=0, ref_mask=None,
                ref_mass=None, use_mass=False, maxwindow=-1, offset=1, soa=False, natom=None,
                method=None, ctx: Context | None = None):
    """RMS average correlation ('rmsavgcorr').  ref_frames None = 'first'.
    method ("auto", "running", "prefix"; rmsd_avgcorr_set_method), when given,
    is set on the context first and stays there.
    Returns (window sizes, mean RMSD, SD), the last two float64 CUDA tensors."""
    import torch

    ctx = ctx or _ctx(frames.device.index or 0)
    d = frames_desc(frames, mask, mass, soa=soa, natom=natom)
    r = None
    if ref_frames is not None:
        r = frames_desc(ref_frames, ref_mask if ref_mask is not None else mask,
                        ref_mass if ref_mass is not None else mass, soa=soa, natom=natom)
    if method is not None:
        if method not in AVGCORR_METHODS:
            raise RmsdError(f"unknown rmsavgcorr method {method!r}")
        ctx.check(ctx._L.rmsd_avgcorr_set_method(ctx.handle, AVGCORR_METHODS[method]))
    w = avgcorr_windows(d.nframes, maxwindow, offset)
    ct = torch.empty(len(w), dtype=torch.float64, device=frames.device)
    sd = torch.empty(len(w), dtype=torch.float64, device=frames.device)
    ctx.check(ctx._L.rmsd_avgcorr(ctx.handle, C.byref(d), None if r is None else C.byref(r),
                                  ref_index, int(use_mass), maxwindow, offset, _ptr(ct),
                                  _ptr(sd)))
    return w, ct, sd


class Batch:
    """An rmsd_batch plan: packed upper triangles of many frame sets from one
    batched launch (rmsd_batch_create / rmsd_batch_run)."""

    def __init__(self, ctx: Context, sets, outs, mode="fit", use_mass=False, precision="bf16x3",
                 graph=True, autotune=False):
        if len(sets) != len(outs):
            raise RmsdError("one output per set")
        self.ctx = ctx
        self._L = ctx._L
        n = len(sets)
        self._sets = (Frames * max(n, 1))(*sets)
        self._outs_t = list(outs)
        self._outs = (C.c_void_p * max(n, 1))(*[o.data_ptr() for o in outs])
        flags = (BATCH_GRAPH if graph else 0) | (BATCH_AUTOTUNE if autotune else 0)
        h = C.c_void_p()
        ctx.check(self._L.rmsd_batch_create(ctx.handle, self._sets, n, MODES[mode], int(use_mass),
                                            PRECISIONS[precision], self._outs, flags, C.byref(h)))
        self._h = h
        ctx._plans.add(self)

    def run(self):
        self.ctx.check(self._L.rmsd_batch_run(self._h))

    def set_variant(self, v: int):
        self.ctx.check(self._L.rmsd_batch_set_variant(self._h, v))

    def info(self) -> dict:
        i = (C.c_int64 * 6)()
        t = (C.c_double * 4)()
        self.ctx.check(self._L.rmsd_batch_info(self._h, i, t))
        return {"variant": i[0], "tile": i[1], "tiles": i[2], "graph": bool(i[3]), "queues": i[4],
                "longest_queue": i[5], "tune_ms": list(t)}

    def close(self):
        # A plan whose context is gone was freed with it.
        if getattr(self, "_h", None) and getattr(self.ctx, "_h", None):
            self._L.rmsd_batch_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rms2d_batch(frame_sets, masks=None, masses=None, mode="fit", use_mass=False,
                precision="bf16x3", graph=False, autotune=False, ctx: Context | None = None):
    """Packed upper-triangle rms2d matrices of several trajectories in one batched
    pass.  frame_sets: list of CUDA tensors (AoS or (F, 3, ld) SoA with unit atom
    stride detected by shape); returns the list of float32 triangles."""
    import torch

    ctx = ctx or _ctx(frame_sets[0].device.index or 0) if frame_sets else None
    n = len(frame_sets)
    masks = masks or [None] * n
    masses = masses or [None] * n
    descs, outs = [], []
    for X, m, w in zip(frame_sets, masks, masses):
        soa = X.dim() == 3 and X.shape[1] == 3 and X.shape[2] != 3
        d = frames_desc(X, m, w, soa=soa)
        descs.append(d)
        F = d.nframes
        outs.append(torch.empty(max(F * (F - 1) // 2, 1), dtype=torch.float32, device=X.device))
    if n == 0:
        return []
    b = Batch(ctx, descs, outs, mode, use_mass, precision, graph=graph, autotune=autotune)
    b.run()
    b.close()
    return [o[: d.nframes * (d.nframes - 1) // 2] for o, d in zip(outs, descs)]
