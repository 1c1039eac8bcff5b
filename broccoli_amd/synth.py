"""Synthetic second-level studies with the shapes of the benchmark workloads.

Grids are (D, H, W) arrays flattened as x + y*W + z*W*H; data are
(N, D, H, W), i.e. Y[x + y*W + z*W*H + s*W*H*D] as the reference expects.
Randomness comes from numpy's PCG64 (numpy.random.default_rng) seeded with
20260000 + workload number, with independent child streams for data,
covariates and permutations.  A `scale` < 1 shrinks the grid and the mask
(same shapes, fewer voxels) for quick parity runs.
"""
from __future__ import annotations

import numpy as np

from . import capi

GRID_2MM = (91, 109, 91)
GRID_1MM = (182, 218, 182)


def ellipsoid_mask(W, H, D, center, semi) -> np.ndarray:
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    r = ((x - center[0]) / semi[0]) ** 2 + ((y - center[1]) / semi[1]) ** 2 + ((z - center[2]) / semi[2]) ** 2
    return (r <= 1.0).astype(np.float32)


def sphere(W, H, D, center, radius) -> np.ndarray:
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    return ((x - center[0]) ** 2 + (y - center[1]) ** 2 + (z - center[2]) ** 2) <= radius ** 2


def smooth_noise(rng, N, W, H, D, fwhm=3.0) -> np.ndarray:
    """Gaussian-smoothed white noise (separable, truncated at 4 sigma), scaled
    by the analytic factor that gives unit variance away from the borders."""
    from scipy.ndimage import correlate1d

    sigma = fwhm / 2.3548
    rad = int(np.ceil(4 * sigma))
    k = np.exp(-0.5 * (np.arange(-rad, rad + 1) / sigma) ** 2)
    k /= k.sum()
    norm = np.sqrt(np.sum(k ** 2)) ** 3
    out = np.empty((N, D, H, W), np.float32)
    for s in range(N):
        v = rng.standard_normal((D, H, W))
        for ax in range(3):
            v = correlate1d(v, k, axis=ax, mode="constant")
        out[s] = (v / norm).astype(np.float32)
    return out


def _scaled(grid, scale):
    return tuple(max(8, int(round(g * scale))) for g in grid)


def study(workload: int, scale: float = 1.0, seed: int | None = None, contrast: str = "t") -> dict:
    """Synthetic study for workloads 1, 2, 3 and 5 (see the module doc).
    Workload 5 has two statistical maps on the same data: contrast "t" (the
    first covariate) or "F" (the 2-row F contrast over group dummies 1 and
    2)."""
    seed = 20260000 + workload if seed is None else seed
    ss = np.random.SeedSequence(seed)
    r_data, r_cov, r_perm = (np.random.default_rng(c) for c in ss.spawn(3))
    if workload == 5:
        W, H, D = _scaled(GRID_1MM, scale)
        f = W / GRID_1MM[0]
        center, semi = (90.5 * f, 108.5 * f, 90.5 * f), (72 * f, 88 * f, 68.8 * f)
    else:
        W, H, D = _scaled(GRID_2MM, scale)
        f = W / GRID_2MM[0]
        center, semi = (45 * f, 54 * f, 45 * f), (36 * f, 44 * f, 34.4 * f)
    mask = ellipsoid_mask(W, H, D, center, semi)
    st = dict(W=W, H=H, D=D, mask=mask, workload=workload, seed=seed, perm_rng=r_perm)
    if workload == 1:
        # one-sample t by sign flipping, N=20, intercept only
        N = 20
        Y = r_data.standard_normal((N, D, H, W), dtype=np.float32)
        Y[:, sphere(W, H, D, (60 * f, 54 * f, 45 * f), 5 * f)] += 0.5
        st.update(N=N, p=1, nc=1, test=capi.GROUP_MEAN, X=np.ones((N, 1), np.float32),
                  C=np.ones((1, 1), np.float32), Y=Y, K=1000)
    elif workload == 2:
        # two-sample GLM with covariates, N=60 (30/30), t contrast [0 1 0 0]
        N = 60
        g = np.r_[np.ones(30), np.zeros(30)]
        age = r_cov.uniform(20, 70, N)
        age -= age.mean()
        sex = (r_cov.random(N) < 0.5).astype(np.float64)
        X = np.c_[np.ones(N), g, age, sex].astype(np.float32)
        Y = r_data.standard_normal((N, D, H, W), dtype=np.float32)
        eff = sphere(W, H, D, (30 * f, 54 * f, 45 * f), 5 * f) | sphere(W, H, D, (60 * f, 54 * f, 45 * f), 5 * f)
        Y[:30][:, eff] += 0.4
        st.update(N=N, p=4, nc=1, test=capi.TTEST, X=X, C=np.array([[0, 1, 0, 0]], np.float32), Y=Y, K=10000)
    elif workload == 3:
        # F test over 3 group dummies (4 groups x 25), 2 covariates, smooth noise
        N = 100
        grp = np.repeat(np.arange(4), 25)
        dum = np.stack([(grp == j).astype(np.float64) for j in (1, 2, 3)], 1)
        cov = r_cov.standard_normal((N, 2))
        X = np.c_[np.ones(N), dum, cov].astype(np.float32)
        C = np.zeros((3, 6), np.float32)
        C[0, 1] = C[1, 2] = C[2, 3] = 1
        Y = smooth_noise(r_data, N, W, H, D)
        st.update(N=N, p=6, nc=3, test=capi.FTEST, X=X, C=C, Y=Y, K=10000,
                  inference=capi.CLUSTER_EXTENT, cdt=4.0)
    elif workload == 5:
        N = 200
        grp = np.repeat(np.arange(4), 50)
        dum = np.stack([(grp == j).astype(np.float64) for j in (1, 2, 3)], 1)
        cov = r_cov.standard_normal((N, 4))
        X = np.c_[np.ones(N), dum, cov].astype(np.float32)
        Y = r_data.standard_normal((N, D, H, W), dtype=np.float32)
        if contrast == "F":
            C = np.zeros((2, 8), np.float32)
            C[0, 1] = C[1, 2] = 1
            st.update(N=N, p=8, nc=2, test=capi.FTEST, X=X, C=C, Y=Y, K=10000)
        else:
            C = np.zeros((1, 8), np.float32)
            C[0, 4] = 1
            st.update(N=N, p=8, nc=1, test=capi.TTEST, X=X, C=C, Y=Y, K=10000)
    else:
        raise ValueError(f"unknown workload {workload}")
    st["V"] = int(mask.sum())
    return st


def permutations(st: dict, K: int, compat: bool = False, seed: int | None = None):
    """Permutation (t, F) or sign (group mean) matrix for a study, from the
    library's generators, with the design detection of the CLI."""
    if seed is None:
        seed = int(st["perm_rng"].integers(1, 2**62))
    N = st["N"]
    if st["test"] == capi.GROUP_MEAN:
        return dict(sign=capi.gen_sign_matrix(N, K, compat=compat, seed=seed))
    designs, n1, n2 = capi.detect_designs(st["X"], st["C"], st["test"] == capi.FTEST)
    if designs[0] == capi.TWOSAMPLE:
        return dict(perm=capi.gen_perm_two_sample(N, int(n1[0]), K, compat=compat, seed=seed))
    return dict(perm=capi.gen_perm_correlation(N, K, compat=compat, seed=seed))


def studies_batch(n: int = 64, scale: float = 1.0, seed: int | None = None, K: int = 5000) -> list:
    """Workload 4: n independent studies on the 91x109x91 grid, each with its
    own ellipsoid mask (50k-230k voxels at full scale), N ~ U{12..120},
    p ~ U{1..6} (intercept + Gaussian covariates), one t contrast on the last
    covariate (p = 1: the intercept, tested by sign flipping).  The data are
    N(0, 1); to keep host memory small the studies take the first N volumes
    of one shared pool of 120 noise volumes."""
    seed = 20260004 if seed is None else seed
    rng = np.random.default_rng(seed)
    W, H, D = _scaled(GRID_2MM, scale)
    f = W / GRID_2MM[0]
    pool = rng.standard_normal((120, D, H, W), dtype=np.float32)
    out = []
    for i in range(n):
        s = rng.uniform(0.61, 1.0)
        center = (45 * f + rng.uniform(-4, 4) * f, 54 * f + rng.uniform(-4, 4) * f, 45 * f + rng.uniform(-4, 4) * f)
        semi = (36 * f * s, 44 * f * s, 34.4 * f * s)
        mask = ellipsoid_mask(W, H, D, center, semi)
        N = int(rng.integers(12, 121))
        p = int(rng.integers(1, 7))
        X = np.c_[np.ones(N), rng.standard_normal((N, p - 1))].astype(np.float32)
        C = np.zeros((1, p), np.float32)
        C[0, p - 1] = 1.0
        st = dict(W=W, H=H, D=D, mask=mask, workload=4, seed=seed, N=N, p=p, nc=1, X=X, C=C, Y=pool[:N], K=K,
                  test=capi.GROUP_MEAN if p == 1 else capi.TTEST, inference=capi.VOXEL,
                  perm_rng=np.random.default_rng(rng.integers(1, 2**62)))
        st["V"] = int(mask.sum())
        out.append(st)
    return out


# ------------------------------------------------------------- first level
# Single-subject fMRI data for the first-level permutation test: volumes are
# (T, D, H, W) arrays, i.e. volumes[x + y*W + z*W*H + t*W*H*D].

def ar_series(coefs, T: int, shape, seed: int = 0, rng=None) -> np.ndarray:
    """AR(4) noise v_t = e_t + sum_k a_k v_{t-k} from zero history, per
    voxel; coefs is 4 values or a (4, *shape) field."""
    rng = rng if rng is not None else np.random.default_rng(seed)
    a = np.asarray(coefs, np.float64)
    if a.ndim == 1:
        a = a.reshape((4,) + (1,) * len(shape))
    v = np.zeros((T,) + tuple(shape))
    for t in range(T):
        acc = rng.standard_normal(shape)
        for k in range(1, 5):
            if t - k >= 0:
                acc = acc + a[k - 1] * v[t - k]
        v[t] = acc
    return v


def first_level_design(T: int, p: int = 5, seed: int = 0) -> np.ndarray:
    """(T, p) design: intercept, linear trend, a block task (on/off blocks of
    10), a sinusoid, then Gaussian nuisance columns (e.g. motion)."""
    rng = np.random.default_rng(seed)
    t = np.arange(T, dtype=np.float64)
    cols = [np.ones(T), (t - t.mean()) / T, ((t // 10) % 2).astype(np.float64), np.sin(2 * np.pi * t / 37.0)]
    while len(cols) < p:
        cols.append(rng.standard_normal(T) * 0.5)
    return np.stack(cols[:p], axis=1).astype(np.float32)


def first_level_study(seed: int = 0, W: int = 12, H: int = 14, D: int = 10, T: int = 64, p: int = 5,
                      ar=(0.3, 0.1, 0.0, 0.0), noise_only: bool = False, voxel=(3.0, 3.0, 3.0),
                      ar_gradient: float = 0.0) -> dict:
    """A synthetic single-subject study: ellipsoid brain mask, AR(4) noise
    (AR1 varying by `ar_gradient` across x), activation 0.6 * task in a
    sphere unless noise_only (and p >= 3), baseline 100.  Contrast: the task
    column (the intercept when p == 1)."""
    rng = np.random.default_rng(seed)
    mask = ellipsoid_mask(W, H, D, ((W - 1) / 2, (H - 1) / 2, (D - 1) / 2), (W * 0.45, H * 0.45, D * 0.45))
    X = first_level_design(T, p, seed + 1)
    a = np.zeros((4, D, H, W))
    for k in range(4):
        a[k] = ar[k]
    if ar_gradient:
        a[0] += ar_gradient * (np.arange(W) / max(W - 1, 1) - 0.5)[None, None, :]
    Y = ar_series(a, T, (D, H, W), rng=rng)
    task = min(2, p - 1)  # the block-task column (the intercept when p == 1)
    if not noise_only and task == 2:
        act = sphere(W, H, D, (W * 0.3, H * 0.5, D * 0.5), max(2.0, W * 0.15))
        Y += 0.6 * X[:, 2][:, None, None, None] * act[None]
    Y += 100.0
    C = np.zeros((1, p), np.float32)
    C[0, task] = 1.0
    return {"Y": Y.astype(np.float32), "mask": mask, "X": X, "C": C, "voxel": tuple(voxel), "W": W, "H": H,
            "D": D, "T": T, "p": p}


def first_level_workload(seed: int = 20260006) -> dict:
    """The first-level benchmark study: 64x64x33 EPI grid (3.75 x 3.75 x 4
    mm voxels), ellipsoid brain mask, T = 240, p = 12, AR(4) noise whose AR1
    varies across the brain, block-task activation."""
    return first_level_study(seed=seed, W=64, H=64, D=33, T=240, p=12, ar=(0.35, 0.1, 0.05, 0.0),
                             voxel=(3.75, 3.75, 4.0), ar_gradient=0.15)
