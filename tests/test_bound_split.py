"""The keep test of the mixed-precision contraction, checked against the bound
it must respect.

The contraction keeps a float RMSD r of a pair whose covariance error bound is
eps when r lies in a float interval [lo, hi] (keep_limits in rmsd_device.hpp,
from a split of the tolerance made once on the host, rmsd_keep_split in
capi.hip); a kept r must satisfy r^2 > eps and
  B(r) = eps r / (2 r^2 - eps) (1 + 2^-23) + r 2^-23 <= tol.
Here both steps are evaluated in float32 the way the device does, with the
hardware square root and reciprocal (1 ulp) taken one ulp low, and B is
evaluated in FP64 at the ends of every interval (B is convex above
sqrt(eps / 2), so its largest value over [lo, hi] is at an end).  The exact
lowest r with B(r) <= tol (bisection on B) shows how much the split gives up.
The per-lane report bounds (bound_kept, bound_any) must not undercut B at the
ends of their ranges.  CPU only.
"""
import numpy as np
import pytest

f32 = np.float32
INF = np.float32(np.inf)


def f_ru(x):
    f = f32(x)
    return np.nextafter(f, INF) if float(f) < x else f


def f_rd(x):
    f = f32(x)
    return np.nextafter(f, -INF) if float(f) > x else f


def keep_split(tol):
    """rmsd_keep_split: (inv4t, c8, h1, h2, e16, tol_f)."""
    a = 1.0 + 2.0 ** -23
    h1 = f_rd(min(16.0, tol * 2.0 ** 22))
    tol1 = tol - float(h1) * 2.0 ** -23
    inv4t = f_ru(1.0 / (4.0 * tol1))
    c8 = f_ru(8.0 * tol1 * tol1)
    tol2 = tol - 256.0 * 2.0 ** -23
    e16 = f_rd(2.0 * tol2 * float(h1) ** 2 / (a * float(h1) + tol2) * (1.0 - 2.0 ** -40)) \
        if tol2 > 0 else f32(-1.0)
    return inv4t, c8, h1, f32(256.0), e16, f_ru(tol)


def sqrt_low(x):
    """The hardware square root, one ulp low (the unsafe side)."""
    return np.nextafter(np.sqrt(f32(x)), f32(0))


def fmaf(a, b, c):
    return f32(float(a) * float(b) + float(c))


def keep_limits(e, k):
    inv4t, c8, h1, h2, e16, _ = k
    t1 = f32(e * f32(1.0 + 2.0 ** -22))
    s = sqrt_low(fmaf(t1, t1, f32(c8 * e)))
    r0 = f32(f32(f32(t1 + s) * inv4t) * f32(1.0 + 2.0 ** -18))
    rs = f32(sqrt_low(e) * f32(1.0 + 2.0 ** -18))
    return max(r0, rs), (h2 if e <= e16 else h1)


def bound_kept(e, kmin, kmax):
    den = f32(fmaf(f32(2.0) * kmin, kmin, -e) * f32(1.0 - 2.0 ** -20))
    rcp = np.nextafter(f32(1.0) / den, f32(0))
    b = f32(f32(f32(f32(e * f32(1.0 + 2.0 ** -22)) * kmin) * rcp) * f32(1.0 + 2.0 ** -18))
    return f32(b + f32(kmax * f32(2.0 ** -23 * (1.0 + 2.0 ** -20))))


def bound_any(e, r):
    b = f32(sqrt_low(fmaf(r, r, e)) * f32(1.0 + 2.0 ** -18))
    return f32(f32(b + f32(f32(r + b) * f32(2.0 ** -23))) * f32(1.0 + 2.0 ** -20))


def B(eps, r):
    """The stated bound of a kept float r (FP64; r^2 > eps)."""
    return eps * r / (2.0 * r * r - eps) * (1.0 + 2.0 ** -23) + r * 2.0 ** -23


def lowest_kept_exact(eps, tol):
    lo, hi = np.sqrt(eps) * (1 + 1e-12), 256.0
    if B(eps, lo) <= tol:
        return lo
    if B(eps, hi) > tol:
        return None
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        lo, hi = (lo, mid) if B(eps, mid) <= tol else (mid, hi)
    return hi


TOLS = [1e-6, 1e-5, 3e-5, 1e-4, 3e-4, 1e-3, 1e-2]
EPS = np.concatenate([[0.0], np.logspace(-20, 1, 400)])


@pytest.mark.parametrize("tol", TOLS)
def test_every_kept_rmsd_has_its_bound_within_tol(tol):
    k = keep_split(tol)
    kept = 0
    for eps in EPS:
        e = f_ru(eps)
        lo, hi = keep_limits(e, k)
        if not lo <= hi:
            continue
        kept += 1
        for r in (float(lo), float(hi)):
            if eps == 0.0:
                assert r * 2.0 ** -23 <= tol
                continue
            assert r * r > eps, (tol, eps, r)
            assert B(eps, r) <= tol, (tol, eps, r, B(eps, r))
    assert kept > 0


@pytest.mark.parametrize("tol", [1e-5, 1e-4, 1e-3])
def test_the_split_gives_up_little(tol):
    k = keep_split(tol)
    worst = 1.0
    for eps in np.logspace(-16, -4, 200):
        exact = lowest_kept_exact(eps, tol)
        lo, hi = keep_limits(f_ru(eps), k)
        if exact is None or exact > float(k[2]):
            continue
        worst = max(worst, float(lo) / exact)
    # tol1 = tol - 16 A 2^-23 costs 1.9e-6 of tol (20% at 1e-5, 2% at 1e-4)
    assert worst <= 1.0 + 1.25 * 16 * 2.0 ** -23 / tol + 1e-5, worst


def test_report_only_keeps_from_sqrt_eps():
    k = (f32(0), f32(0), INF, INF, INF, INF)  # rmsd_keep_split(tol < 0)
    for eps in np.logspace(-16, 1, 50):
        lo, hi = keep_limits(f_ru(eps), k)
        assert float(lo) ** 2 > eps and float(lo) <= np.sqrt(eps) * (1 + 2.0 ** -16)
        assert hi == INF


@pytest.mark.parametrize("tol", [1e-5, 1e-4, 1e-3])
def test_report_bounds_cover_the_kept_and_the_rest(tol):
    rng = np.random.default_rng(5)
    k = keep_split(tol)
    for eps in np.logspace(-14, -2, 120):
        e = f_ru(eps)
        lo, hi = keep_limits(e, k)
        if not lo <= hi:
            continue
        top = min(float(hi), 64.0)
        if top < float(lo):
            continue
        a, b = np.sort(rng.uniform(float(lo), top, 2)).astype(np.float32)
        kmin, kmax = max(a, lo), b
        rep = float(bound_kept(e, kmin, kmax))
        assert rep >= B(eps, float(kmin)) and rep >= B(eps, float(kmax))
        r = f32(rng.uniform(0, float(lo)))
        exact = np.sqrt(float(r) ** 2 + eps)
        assert float(bound_any(e, r)) >= exact + (float(r) + exact) * 2.0 ** -23
