"""Seeded synthetic trajectories for the benchmark workloads and parity tests.

Every generator returns FP32 coordinates (so the library, the FP64 oracle and
the reference code all see identical, FP32-representable inputs) plus the
selected-atom masses where the workload is mass-weighted.

Recipes (one per workload):

* chain(n)        3D random walk, x_0 = 0, x_{k+1} = x_k + 3.8 * u_k with u_k
                  uniform on the sphere (normalised N(0, I3)).
* rigid(x)        Q x + t, Q a uniform random rotation (Shoemake's uniform
                  quaternion), t uniform in [-20, 20]^3.
* noisy(F, n, s)  F frames rigid(ref + N(0, s^2)), ref = chain(n).
* drift(F, n)     frame_k = rigid(ref + d_k), d_k = 0.99 d_{k-1} + N(0, 0.3^2)
                  per coordinate, d_0 drawn from the stationary law
                  N(0, 0.3^2 / (1 - 0.99^2)).
* two_state(F, n) two conformations 20 A apart (a hinge), sigma 0.1 A noise,
                  alternating frames, rigid motion (precision stress case).
* masses(n)       i.i.d. draws from {1, 12, 14, 16, 32} with probabilities
                  0.50 / 0.30 / 0.08 / 0.10 / 0.02 (H / C / N / O / S-like).
"""
from __future__ import annotations

import numpy as np

MASS_VALUES = np.array([1.0, 12.0, 14.0, 16.0, 32.0])
MASS_PROBS = np.array([0.50, 0.30, 0.08, 0.10, 0.02])


def chain(rng: np.random.Generator, n: int, step: float = 3.8) -> np.ndarray:
    u = rng.normal(size=(n - 1, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    x = np.zeros((n, 3))
    x[1:] = np.cumsum(step * u, axis=0)
    return x


def rotations(rng: np.random.Generator, k: int) -> np.ndarray:
    """k uniform random rotation matrices (Shoemake 1992)."""
    u1, u2, u3 = rng.random(k), rng.random(k), rng.random(k)
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    w = a * np.sin(2 * np.pi * u2)
    x = a * np.cos(2 * np.pi * u2)
    y = b * np.sin(2 * np.pi * u3)
    z = b * np.cos(2 * np.pi * u3)
    R = np.empty((k, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def masses(rng: np.random.Generator, n: int) -> np.ndarray:
    return rng.choice(MASS_VALUES, size=n, p=MASS_PROBS)


def _rigid(rng: np.random.Generator, X: np.ndarray) -> np.ndarray:
    F = X.shape[0]
    R = rotations(rng, F)
    t = rng.uniform(-20.0, 20.0, size=(F, 1, 3))
    return np.einsum("fij,fnj->fni", R, X) + t


def noisy(seed: int, F: int, n: int, sigma: float) -> np.ndarray:
    """(F, n, 3) float32: rigid(ref + N(0, sigma^2))."""
    rng = np.random.default_rng(seed)
    ref = chain(rng, n)
    X = ref[None] + rng.normal(scale=sigma, size=(F, n, 3))
    return _rigid(rng, X).astype(np.float32)


def drift(seed: int, F: int, n: int, rho: float = 0.99, sigma: float = 0.3) -> np.ndarray:
    """(F, n, 3) float32: rigid(ref + d_k), AR(1) displacement d_k."""
    rng = np.random.default_rng(seed)
    ref = chain(rng, n)
    X = np.empty((F, n, 3))
    d = rng.normal(scale=sigma / np.sqrt(1.0 - rho * rho), size=(n, 3))
    for k in range(F):
        if k:
            d = rho * d + rng.normal(scale=sigma, size=(n, 3))
        X[k] = ref + d
    out = np.empty((F, n, 3), dtype=np.float32)
    blk = 4096
    for s in range(0, F, blk):
        out[s:s + blk] = _rigid(rng, X[s:s + blk])
    return out


def two_state(seed: int, F: int, n: int, separation: float = 20.0, sigma: float = 0.1):
    """(F, n, 3) float32 frames of two conformations `separation` Angstrom apart
    (best-fit RMSD): state A is a chain, state B the same chain with its second
    half rotated about the middle atom (a hinge), the angle found by bisection.
    Frame k is in state k % 2, plus N(0, sigma^2) noise, then a rigid motion.
    Returns (frames, state per frame)."""
    rng = np.random.default_rng(seed)
    A = chain(rng, n)
    mid = n // 2
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)

    def hinge(theta):
        K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
        R = np.eye(3) + np.sin(theta) * K + (1 - np.cos(theta)) * (K @ K)
        B = A.copy()
        B[mid:] = (A[mid:] - A[mid]) @ R.T + A[mid]
        return B

    def fit_rmsd(X, Y):
        x = X - X.mean(0)
        y = Y - Y.mean(0)
        V, S, Wt = np.linalg.svd(x.T @ y)
        d = np.sign(np.linalg.det(V @ Wt))
        S[-1] *= d
        return np.sqrt(max((x * x).sum() + (y * y).sum() - 2 * S.sum(), 0.0) / len(X))

    lo, hi = 0.0, np.pi
    for _ in range(60):
        th = 0.5 * (lo + hi)
        if fit_rmsd(A, hinge(th)) < separation:
            lo = th
        else:
            hi = th
    B = hinge(hi)
    state = np.arange(F) % 2
    Y = np.where(state[:, None, None] == 0, A[None], B[None]) + rng.normal(scale=sigma,
                                                                            size=(F, n, 3))
    return _rigid(rng, Y).astype(np.float32), state


def batch_sizes(seed: int, ntraj: int = 64, f_range=(500, 20000), n_range=(50, 5000)):
    """Per-trajectory (frames, atoms), log-uniform."""
    rng = np.random.default_rng(seed)
    F = np.exp(rng.uniform(np.log(f_range[0]), np.log(f_range[1]), ntraj)).astype(int)
    N = np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]), ntraj)).astype(int)
    return F, N


def to_soa(X: np.ndarray, ld: int | None = None) -> np.ndarray:
    """(F, n, 3) -> (F, 3, ld) float32 with zero padding."""
    F, n, _ = X.shape
    ld = ld or n
    out = np.zeros((F, 3, ld), dtype=np.float32)
    out[:, :, :n] = np.transpose(X, (0, 2, 1))
    return out


# ---- device-side generator for the streaming-fit workload ------------------------
def noisy_soa_torch(seed: int, F: int, n: int, sigma: float, ld: int, device="cuda",
                    chunk: int = 65536):
    """(F, 3, ld) float32 on `device`: rigid(ref + N(0, sigma^2)) built with torch's
    seeded generator in chunks (1e6 x 3000 frames do not fit a host round trip).
    Returns (frames, ref (n, 3) float64 numpy)."""
    import torch

    rng = np.random.default_rng(seed)
    ref = chain(rng, n)
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    out = torch.zeros((F, 3, ld), dtype=torch.float32, device=device)
    ref_t = torch.as_tensor(ref.T.copy(), dtype=torch.float32, device=device)  # (3, n)
    for s in range(0, F, chunk):
        e = min(F, s + chunk)
        k = e - s
        R = torch.as_tensor(rotations(rng, k), dtype=torch.float32, device=device)
        t = torch.as_tensor(rng.uniform(-20.0, 20.0, size=(k, 3, 1)), dtype=torch.float32,
                            device=device)
        X = ref_t.unsqueeze(0) + sigma * torch.randn((k, 3, n), generator=g, device=device)
        out[s:e, :, :n] = torch.bmm(R, X) + t
    return out, ref


def symm_case(seed: int, F: int, n: int, sizes=(2, 3, 2, 4, 3, 2), ngroups: int = 40,
              sigma: float = 0.3, rigid: bool = True, compact: bool = True):
    """Frames for the symmetry-corrected RMSD: a reference chain of n atoms
    with `ngroups` disjoint groups (sizes cycled from `sizes`, atoms picked
    over the whole chain; with `compact` each group's atoms are then moved to
    1.5 A from its first atom, as symmetric atoms sit in a molecule), masses
    equal within each group, and F frames that are
    the reference plus N(0, sigma^2) noise with the atoms of every group
    permuted at random (a different permutation per frame), then (rigid)
    rotated and translated.  Returns (frames (F, 3n) float32, ref (n, 3),
    groups (lists of atom indices), masses (n), perms (F, n): frame atom k
    holds the reference's atom perms[f, k])."""
    rng = np.random.default_rng(seed)
    ref = chain(rng, n)
    m = masses(rng, n)
    order = rng.permutation(n)
    groups, p = [], 0
    for g in range(ngroups):
        k = int(sizes[g % len(sizes)])
        if p + k > n:
            break
        grp = sorted(int(a) for a in order[p:p + k])
        groups.append(grp)
        m[grp] = m[grp[0]]
        if compact:
            for a in grp[1:]:
                u = rng.normal(size=3)
                ref[a] = ref[grp[0]] + 1.5 * u / np.linalg.norm(u)
        p += k
    X = np.empty((F, n, 3))
    perms = np.tile(np.arange(n), (F, 1))
    for f in range(F):
        for grp in groups:
            perms[f, grp] = np.asarray(grp)[rng.permutation(len(grp))]
        Y = ref[perms[f]] + rng.normal(size=(n, 3)) * sigma
        X[f] = _rigid(rng, Y[None])[0] if rigid else Y
    return X.reshape(F, -1).astype(np.float32), ref, groups, m, perms
