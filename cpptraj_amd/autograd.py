"""torch.autograd bindings of the RMSD kernels (thin test binding).

* ``RmsdToReference``  RMSD of every frame to one reference frame (the rmsd
  action's fit, nomod); backward through ``rmsd_fit_backward``.
* ``PairwiseRmsd``     full (A x B) best-fit RMSD block; backward through
  ``rmsd_pairwise_backward`` (rotations recomputed in FP64).

Coordinates are float32 CUDA tensors of shape (F, n, 3) (the selected atoms,
AoS).  Forward and backward run in the HIP library; the FP64 gradient is cast to
the input dtype for autograd.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import Context, RmsdError, frames_desc, load_library
from . import _ctx as _device_ctx


def _ctx(dev) -> Context:
    """The default context of `dev` on torch's current stream there."""
    return _device_ctx(dev.index or 0)


def fit_backward(frames, ref, grad_rmsd, mass=None, use_mass=False, ctx: Context | None = None):
    """FP64 d(sum_f g_f RMSD_f)/d frames, shape (F, n, 3)."""
    ctx = ctx or _ctx(frames.device)
    L = load_library()
    F, n = frames.shape[0], frames.shape[1]
    # The descriptors hold raw pointers: keep the contiguous copies alive.
    fc, rc = frames.contiguous(), ref.contiguous()
    d = frames_desc(fc, None, mass)
    r = frames_desc(rc, None, mass)
    g = torch.empty((F, n, 3), dtype=torch.float64, device=frames.device)
    gr = grad_rmsd.to(torch.float64).contiguous()
    ctx.check(L.rmsd_fit_backward(ctx.handle, C.byref(d), C.byref(r), 0, int(use_mass),
                                  C.c_void_p(gr.data_ptr()), C.c_void_p(g.data_ptr())))
    return g


def pairwise_backward(A, B, grad_out, mass=None, use_mass=False, ctx: Context | None = None):
    """FP64 gradients of sum_ij G_ij RMSD(A_i, B_j) with respect to A and B."""
    ctx = ctx or _ctx(A.device)
    L = load_library()
    Ac, Bc = A.contiguous(), B.contiguous()
    da = frames_desc(Ac, None, mass)
    db = frames_desc(Bc, None, mass)
    ga = torch.empty((A.shape[0], A.shape[1], 3), dtype=torch.float64, device=A.device)
    gb = torch.empty((B.shape[0], B.shape[1], 3), dtype=torch.float64, device=B.device)
    G = grad_out.to(torch.float64).contiguous()
    ctx.check(L.rmsd_pairwise_backward(ctx.handle, C.byref(da), C.byref(db), int(use_mass),
                                       C.c_void_p(G.data_ptr()), C.c_void_p(ga.data_ptr()),
                                       C.c_void_p(gb.data_ptr())))
    return ga, gb


class RmsdToReference(torch.autograd.Function):
    @staticmethod
    def forward(fctx, frames, ref, mass=None, use_mass=False):
        import cpptraj_amd as ca
        if frames.dim() != 3 or ref.dim() != 3 or ref.shape[0] != 1:
            raise RmsdError("frames (F, n, 3) and ref (1, n, 3) expected")
        if ref.requires_grad:
            # rmsd_fit_backward differentiates the frames only; a silent None
            # for the reference would read as a zero gradient.
            raise RmsdError("RmsdToReference has no gradient for the reference: "
                            "pass ref.detach(), or use PairwiseRmsd for both sides")
        ctx = _ctx(frames.device)
        fc, rc = frames.contiguous(), ref.contiguous()
        d = frames_desc(fc, None, mass)
        r = frames_desc(rc, None, mass)
        out = torch.empty(frames.shape[0], dtype=torch.float64, device=frames.device)
        ctx.fit_desc(d, r, 0, True, use_mass, ca.MODIFY["nomod"], out)
        fctx.save_for_backward(frames, ref)
        fctx.mass = mass
        fctx.use_mass = use_mass
        return out

    @staticmethod
    def backward(fctx, g):
        frames, ref = fctx.saved_tensors
        grad = fit_backward(frames, ref, g, fctx.mass, fctx.use_mass)
        return grad.to(frames.dtype), None, None, None


class PairwiseRmsd(torch.autograd.Function):
    @staticmethod
    def forward(fctx, A, B, mass=None, use_mass=False):
        import cpptraj_amd as ca
        ctx = _ctx(A.device)
        Ac, Bc = A.contiguous(), B.contiguous()
        da = frames_desc(Ac, None, mass)
        db = frames_desc(Bc, None, mass)
        out = torch.empty((A.shape[0], B.shape[0]), dtype=torch.float32, device=A.device)
        ctx.pairwise_desc(da, db, ca.MODES["fit"], use_mass, ca.PRECISIONS["fp64"], out)
        fctx.save_for_backward(A, B)
        fctx.mass = mass
        fctx.use_mass = use_mass
        return out

    @staticmethod
    def backward(fctx, G):
        A, B = fctx.saved_tensors
        ga, gb = pairwise_backward(A, B, G, fctx.mass, fctx.use_mass)
        return ga.to(A.dtype), gb.to(B.dtype), None, None
