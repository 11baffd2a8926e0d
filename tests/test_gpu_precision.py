"""The mixed-precision pairwise forms and their stated error bound.

Every result of the FP32-MFMA and split-BF16 forms carries a bound on its
distance from the FP64 computation; pairs whose bound exceeds the context's
tolerance are recomputed in FP64 inside the same call.  The checks:
  * a two-state trajectory (two conformations 20 A apart, sigma 0.1 A inside a
    state), where pairs inside the state far from the template lose most to
    the mixed precision: every pair within 1e-4 A of the FP64 kernel, with
    pairs recomputed (triangle, full matrix, masses, qrmsd, sieved cache);
  * a ladder of tolerances on a drifting trajectory: with tolerance t every
    pair is within t of FP64 (so every bound that passed t held);
  * report-only mode (negative tolerance): the largest error is below the
    largest stated bound; tolerance 0 switches the bound off;
  * the FP64 recomputation itself against the oracle.
"""
import numpy as np
import pytest

from cpptraj_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-4


def dev(a, dt=None):
    import torch
    if a is None:
        return None
    a = np.ascontiguousarray(a)
    if dt is None:
        dt = {np.float32: torch.float32, np.float64: torch.float64,
              np.int32: torch.int32}[a.dtype.type]
    return torch.as_tensor(a, dtype=dt).cuda()


def rms2d(ctx, X, precision, **kw):
    import cpptraj_amd as ca
    out = ca.rms2d(dev(X), precision=precision, ctx=ctx, **kw)
    return out.cpu().numpy().astype(np.float64)


@pytest.fixture(scope="module")
def two_state():
    X, state = synth.two_state(31, 512, 1200, separation=20.0, sigma=0.1)
    return X.reshape(len(X), -1), state


@pytest.mark.parametrize("precision", ["bf16x3", "fp32"])
@pytest.mark.parametrize("mode", ["fit", "qrmsd"])
def test_two_state_every_pair_within_tolerance(ctx, two_state, precision, mode):
    X, state = two_state
    ctx.set_tolerance(TOL)
    exp = rms2d(ctx, X, "fp64", mode=mode)
    got = rms2d(ctx, X, precision, mode=mode)
    bound, nfp64 = ctx.error_report()
    err = np.abs(got - exp)
    print(f"{precision} {mode}: max err {err.max():.3e}, stated max bound {bound:.3e}, "
          f"{nfp64} of {len(got)} pairs recomputed in FP64")
    assert err.max() <= TOL
    assert 0.0 < bound <= TOL
    # Pairs inside the state far from the template need FP64 with split-BF16.
    if precision == "bf16x3":
        assert nfp64 > 0
    # Without the recomputation the same call is measurably worse for split-BF16:
    # the bound is doing work.
    if precision == "bf16x3" and mode == "fit":
        ctx.set_tolerance(0.0)
        raw = rms2d(ctx, X, precision, mode=mode)
        assert ctx.error_report() == (-1.0, 0)
        print(f"  without recomputation: max err {np.abs(raw - exp).max():.3e}")
        ctx.set_tolerance(TOL)


@pytest.mark.parametrize("precision", ["bf16x3", "fp32"])
def test_two_state_full_matrix_and_masses(ctx, two_state, precision):
    """Rows = targets (one state first), columns = references; masses weight
    the covariance and centre each set (the fallback must follow the same
    row / column roles and sqrt(mass) scaling as the contraction)."""
    X, state = two_state
    rng = np.random.default_rng(7)
    m = synth.masses(rng, 1200)
    T = X[:200]
    R = X[256:400]
    ctx.set_tolerance(TOL)
    kw = dict(mass=dev(m), use_mass=True, ref_frames=dev(R), ref_mass=dev(m))
    exp = rms2d(ctx, T, "fp64", **kw)
    got = rms2d(ctx, T, precision, **kw)
    bound, nfp64 = ctx.error_report()
    assert got.shape == (200, 144)
    print(f"{precision} full: max err {np.abs(got - exp).max():.3e}, bound {bound:.3e}, "
          f"{nfp64} recomputed")
    assert np.abs(got - exp).max() <= TOL and bound <= TOL


def test_two_state_sieved_cluster_cache(ctx, two_state):
    import cpptraj_amd as ca
    X, state = two_state
    fidx = np.arange(3, len(X), 3, dtype=np.int32)
    ctx.set_tolerance(TOL)
    a = ca.cluster_cache(dev(X), dev(fidx), precision="fp64", ctx=ctx).cpu().numpy()
    b = ca.cluster_cache(dev(X), dev(fidx), precision="bf16x3", ctx=ctx).cpu().numpy()
    bound, nfp64 = ctx.error_report()
    assert np.abs(a.astype(np.float64) - b).max() <= TOL and nfp64 > 0


def test_recomputed_pairs_match_the_oracle(oracle, ctx, two_state):
    """With a tolerance below every bound, every pair is recomputed by the FP64
    fallback: the whole matrix is then the FP64 result (<= 1e-6 A of the
    oracle)."""
    X, state = two_state
    Y = X[:96]
    ctx.set_tolerance(1e-12)
    got = rms2d(ctx, Y, "bf16x3")
    bound, nfp64 = ctx.error_report()
    assert nfp64 == len(got)
    exp = oracle.rms2d(Y)
    assert np.abs(got - exp).max() <= 1e-6
    ctx.set_tolerance(TOL)


@pytest.mark.parametrize("precision", ["bf16x3", "fp32"])
@pytest.mark.parametrize("layout", ["aos", "aos_tail", "soa", "soa_odd_ld", "mask"])
def test_recomputation_in_every_frame_layout(oracle, ctx, two_state, layout, precision):
    """The FP64 recomputation loads four atoms per 16-byte load where the frames
    allow it (AoS with a row stride, or SoA with a row length, that is a
    multiple of 4 floats, and no selection; atom by atom past the last whole
    quad) and one atom at a time otherwise.  With every pair recomputed, each
    layout gives the FP64 result (<= 1e-6 A of the oracle)."""
    import torch
    import cpptraj_amd as ca
    X, state = two_state
    F, n = 72, 1200
    Y = np.ascontiguousarray(X[:F].reshape(F, n, 3))
    kw, used, sel = {}, Y, None
    if layout == "aos":
        inp = dev(Y)
    elif layout == "aos_tail":  # 1,197 atoms of 1,200-atom rows: 299 quads + 1
        inp, kw, used = dev(Y), {"natom": 1197}, Y[:, :1197]
    elif layout == "soa":  # (F, 3, 1200): ld a multiple of 4
        inp, kw = dev(np.ascontiguousarray(Y.transpose(0, 2, 1))), {"soa": True}
    elif layout == "soa_odd_ld":  # rows of 1,203 floats, 1,200 atoms used
        S = np.zeros((F, 3, 1203), np.float32)
        S[:, :, :n] = Y.transpose(0, 2, 1)
        inp, kw = dev(S), {"soa": True, "natom": n}
    else:  # every third atom selected
        sel = np.arange(0, n, 3, dtype=np.int32)
        inp, kw, used = dev(Y), {"mask": dev(sel)}, Y[:, sel]
    ctx.set_tolerance(1e-12)
    try:
        got = ca.rms2d(inp, precision=precision, ctx=ctx, **kw).cpu().numpy().astype(np.float64)
        bound, nfp64 = ctx.error_report()
    finally:
        ctx.set_tolerance(TOL)
    assert nfp64 == len(got) == F * (F - 1) // 2
    exp = oracle.rms2d(np.ascontiguousarray(used).reshape(F, -1))
    assert np.abs(got - exp).max() <= 1e-6


@pytest.mark.parametrize("precision", ["bf16x3", "fp32"])
def test_tolerance_ladder_on_drift(ctx, precision):
    """Drifting trajectory: with tolerance t every pair is within t of FP64, for
    t from 1e-5 to 1e-4; report-only mode states a bound above the largest
    error."""
    F, n = 1536, 1200
    X = synth.drift(2, F, n).reshape(F, -1)
    ctx.set_tolerance(TOL)
    exp = rms2d(ctx, X, "fp64")
    ctx.set_tolerance(-1.0)
    raw = rms2d(ctx, X, precision)
    bound, nfp64 = ctx.error_report()
    assert nfp64 == 0
    assert np.abs(raw - exp).max() <= bound
    print(f"{precision} report only: max err {np.abs(raw - exp).max():.3e}, "
          f"max bound {bound:.3e}")
    for t in (1e-5, 3e-5, 1e-4):
        ctx.set_tolerance(t)
        got = rms2d(ctx, X, precision)
        b, nf = ctx.error_report()
        err = np.abs(got - exp).max()
        print(f"  tol {t:.0e}: max err {err:.3e}, bound {b:.3e}, {nf} recomputed "
              f"({nf / len(got):.2%})")
        assert err <= t and b <= t
    ctx.set_tolerance(TOL)


def test_fp64_and_nofit_report_zero(ctx):
    X = synth.noisy(4, 100, 60, 0.5).reshape(100, -1)
    for prec, mode in (("fp64", "fit"), ("bf16x3", "nofit")):
        rms2d(ctx, X, prec, mode=mode)
        assert ctx.error_report() == (0.0, 0)
