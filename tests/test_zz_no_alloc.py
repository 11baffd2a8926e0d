"""rmsd_fit_to_ref_noalloc: the allocation-free fit against rmsd_fit_to_ref.

Both run the same kernels, so RMSD, U, the translation vectors and the
updated coordinates must agree bit for bit.  Two calls in a row, the second
with fewer frames and atoms (the shared per-GPU scratch keeps the first call's
contents), weighted and unweighted, fit and no fit, with and without the
coordinate update; a call beyond the scratch size is rejected by the host
before anything runs on the device.
"""
import ctypes as C

import numpy as np
import pytest

from cpptraj_amd import synth

pytestmark = pytest.mark.gpu


def _bind(L):
    P = C.c_void_p
    FP = C.POINTER(__import__("cpptraj_amd").Frames)
    L.rmsd_fit_to_ref_noalloc.argtypes = [P, FP, FP, C.c_int, C.c_int, C.c_int, C.c_int, P, P,
                                          P]
    L.rmsd_fit_to_ref_noalloc.restype = C.c_int
    return L


def _fit(ctx, fn, X, R, m, fit, use_mass, modify):
    """One call of `fn` on copies of X (AoS (F, n, 3)) and R; returns the outputs
    and the frames after the call."""
    import torch
    import cpptraj_amd as ca
    Xd = torch.as_tensor(X).cuda()
    Rd = torch.as_tensor(R).cuda()
    md = torch.as_tensor(m).cuda()
    d = ca.frames_desc(Xd, None, md)
    r = ca.frames_desc(Rd, None, md)
    F = X.shape[0]
    rmsd = torch.full((F,), -7.0, dtype=torch.float64, device="cuda")
    U = torch.full((F, 9), -7.0, dtype=torch.float64, device="cuda")
    tv = torch.full((F, 6), -7.0, dtype=torch.float64, device="cuda")
    ctx.check(fn(ctx.handle, C.byref(d), C.byref(r), 0, int(fit), int(use_mass), modify,
                 C.c_void_p(rmsd.data_ptr()), C.c_void_p(U.data_ptr()),
                 C.c_void_p(tv.data_ptr())))
    torch.cuda.synchronize()
    return rmsd.cpu().numpy(), U.cpu().numpy(), tv.cpu().numpy(), Xd.cpu().numpy()


@pytest.mark.parametrize("use_mass", [False, True])
@pytest.mark.parametrize("fit,modify", [(True, 0), (True, 1), (True, 2), (False, 0)])
def test_noalloc_matches_fit_to_ref(ctx, use_mass, fit, modify):
    L = _bind(ctx._L)
    rng = np.random.default_rng(90 + 3 * modify + use_mass)
    for F, n in ((3000, 700), (257, 130)):   # second call smaller
        X = synth.noisy(F + n, F, n, 0.7)
        R = synth.noisy(F + n + 1, 2, n, 0.7)[:1]
        m = synth.masses(rng, n)
        a = _fit(ctx, L.rmsd_fit_to_ref_noalloc, X, R, m, fit, use_mass, modify)
        b = _fit(ctx, L.rmsd_fit_to_ref, X, R, m, fit, use_mass, modify)
        assert np.array_equal(a[0], b[0]), (F, n)
        if fit:
            assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]), (F, n)
        assert np.array_equal(a[3], b[3]), (F, n)
        assert np.all(a[0] >= 0.0)


def test_noalloc_rejects_calls_beyond_its_scratch(ctx):
    import torch
    import cpptraj_amd as ca
    L = _bind(ctx._L)
    X = torch.zeros((4, 30), dtype=torch.float32, device="cuda")
    d = ca.frames_desc(X)
    d.nframes = 600000              # 600,000 x 128 B > 64 MiB: rejected on the host
    out = torch.zeros(4, dtype=torch.float64, device="cuda")
    rc = L.rmsd_fit_to_ref_noalloc(ctx.handle, C.byref(d), C.byref(ca.frames_desc(X)), 0, 1, 0,
                                   0, C.c_void_p(out.data_ptr()), None, None)
    assert rc == 1
    assert "RMSD_NOALLOC_SCRATCH_BYTES" in ctx._L.rmsd_last_error(ctx.handle).decode()
