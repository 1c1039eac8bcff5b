"""Pins the first-level oracle (oracle/fl_oracle.c) against closed forms,
library routines, round trips and the reference's own generator.

Reference lines: LIB:16363-16558 (whitening and the permutation loop),
LIB:16877-16922 (generator), LIB:11055-11090 (filters),
kernelConvolution.cpp:52-91/531-569/1201-1240 (separable smoothing),
kernelWhitening.cpp:103-535 (AR(4) estimation, whitening, inverse
whitening), kernelStatistics3.cpp:2107-2190 / kernelStatistics5.cpp:
1943-2060 (first-level permutation t / F).
"""
import numpy as np
import pytest
import scipy.linalg
import scipy.ndimage

from oracle import oracle as orc
from tests.fl_data import ar_series, first_level_study


def test_filters_are_normalised_gaussians():
    fwhm, vox = 8.0, (3.0, 3.5, 4.0)
    fx, fy, fz = orc.fl_filters(fwhm, *vox)
    for f, v in zip((fx, fy, fz), vox):
        s = fwhm / 2.354 / v
        u = np.arange(9) - 4.0
        g = np.exp(-u * u / (2 * s * s))
        np.testing.assert_allclose(f, g / g.sum(), rtol=2e-7)
        assert np.array_equal(f, f[::-1])


def test_smoothing_matches_scipy_with_zero_padding():
    rng = np.random.default_rng(1)
    vol = rng.standard_normal((7, 9, 11))
    cert = (rng.random((7, 9, 11)) > 0.3).astype(np.float32)
    scert = rng.random((7, 9, 11)) + 0.5
    filt = orc.fl_filters(6.0, 2.0, 3.0, 2.5)
    got = orc.fl_smooth(vol, filt, cert, scert)
    fx, fy, fz = (f.astype(np.float64) for f in filt)
    a = scipy.ndimage.correlate1d(vol * cert, fy, axis=1, mode="constant")
    a = scipy.ndimage.correlate1d(a, fx, axis=2, mode="constant")
    a = scipy.ndimage.correlate1d(a, fz, axis=0, mode="constant") / scert
    np.testing.assert_allclose(got, a, rtol=1e-12, atol=1e-12)


def test_normalised_smoothing_keeps_a_constant_field():
    mask = np.zeros((8, 10, 12), np.float32)
    mask[2:6, 1:9, 3:10] = 1
    filt = orc.fl_filters(8.0, 3.0, 3.0, 3.0)
    smask = orc.fl_smooth(mask.astype(np.float64), filt)
    out = orc.fl_smooth(np.full(mask.shape, 0.37), filt, mask, smask)
    np.testing.assert_allclose(out[mask == 1], 0.37, rtol=1e-12)


def test_regression_residuals_match_lstsq():
    st = first_level_study(seed=3)
    Y, mask, X = st["Y"].astype(np.float64), st["mask"], st["X"]
    e = orc.fl_regress(Y, mask, X)
    idx = np.nonzero(mask.ravel() == 1)[0]
    y = Y.reshape(Y.shape[0], -1)[:, idx]
    beta, *_ = np.linalg.lstsq(X.astype(np.float64), y, rcond=None)
    np.testing.assert_allclose(e.reshape(Y.shape[0], -1)[:, idx], y - X @ beta, atol=1e-9)
    assert np.all(e.reshape(Y.shape[0], -1)[:, mask.ravel() != 1] == 0)


def test_ar4_is_the_offset_yule_walker_solution():
    """Levinson (scipy.linalg.solve_toeplitz) on the lag sums with the
    0.001 off-diagonal offsets, scaled by det / (det + 0.001)."""
    rng = np.random.default_rng(5)
    T, invalid = 120, 4
    data = rng.standard_normal((T, 2, 3, 4)).cumsum(axis=0) * 0.2 + rng.standard_normal((T, 2, 3, 4))
    mask = np.ones((2, 3, 4), np.float32)
    mask[0, 0, 0] = 0
    ar = orc.fl_ar4(data, mask, invalid)
    assert np.all(ar[:, 0, 0, 0] == 0)
    for (z, y, x) in [(0, 1, 2), (1, 2, 3), (1, 0, 0)]:
        v = data[invalid:, z, y, x]
        n = len(v)
        c = np.array([np.dot(v[k:], v[:n - k]) / (T - 1 - k - invalid) for k in range(5)])
        r = c[1:] / c[0]
        col = np.concatenate([[1.0], r[:3] + 0.001])
        a = scipy.linalg.solve_toeplitz(col, r)
        det = np.linalg.det(scipy.linalg.toeplitz(col))
        np.testing.assert_allclose(ar[:, z, y, x], a * det / (det + 0.001), rtol=1e-10, atol=1e-13)


def test_ar4_recovers_known_coefficients():
    true = np.array([0.5, -0.2, 0.1, 0.05])
    data = ar_series(true, T=40000, shape=(1, 1, 2), seed=7)
    ar = orc.fl_ar4(data, np.ones((1, 1, 2), np.float32), 0)
    np.testing.assert_allclose(ar[:, 0, 0, 0], true, atol=0.02)
    np.testing.assert_allclose(ar[:, 0, 0, 1], true, atol=0.02)


def test_whitening_returns_the_innovations():
    true = np.array([0.4, 0.1, -0.15, 0.05])
    rng = np.random.default_rng(11)
    T = 50
    eps = rng.standard_normal((T, 1, 2, 3))
    v = np.zeros_like(eps)
    for t in range(T):
        v[t] = eps[t] + sum(true[k - 1] * v[t - k] for k in range(1, 5) if t - k >= 0)
    ar = np.broadcast_to(true[:, None, None, None], (4, 1, 2, 3)).copy()
    mask = np.ones((1, 2, 3), np.float32)
    np.testing.assert_allclose(orc.fl_whiten(v, ar, mask), eps, atol=1e-12)
    # inverse whitening with the identity rebuilds the series; a permutation
    # of the whitened input is the same as permuting before
    ident = np.arange(T, dtype=np.uint16)
    np.testing.assert_allclose(orc.fl_color(eps, ar, mask, ident), v, atol=1e-12)
    perm = rng.permutation(T).astype(np.uint16)
    np.testing.assert_allclose(orc.fl_color(eps, ar, mask, perm), orc.fl_color(eps[perm], ar, mask, ident),
                               atol=1e-12)


def test_first_level_t_and_f_closed_forms():
    st = first_level_study(seed=13)
    X, mask = st["X"].astype(np.float64), st["mask"]
    rng = np.random.default_rng(2)
    T = X.shape[0]
    x = rng.standard_normal((T,) + mask.shape)
    C = np.array([[0, 1, 0, 0, 0], [0, 0, 1, -1, 0]], np.float64)
    idx = np.nonzero(mask.ravel() == 1)[0]
    xs = x.reshape(T, -1)[:, idx]
    beta, *_ = np.linalg.lstsq(X, xs, rcond=None)
    e = xs - X @ beta
    var = e.var(axis=0, ddof=1)
    XtXi = np.linalg.inv(X.T @ X)
    for c in range(2):
        t = orc.fl_stat(0, x, mask, X, C, c).ravel()[idx]
        want = (C[c] @ beta) / np.sqrt(var * (C[c] @ XtXi @ C[c]))
        np.testing.assert_allclose(t, want, rtol=1e-9)
        # one F row is t^2
        f1 = orc.fl_stat(1, x, mask, X, C[c:c + 1]).ravel()[idx]
        np.testing.assert_allclose(f1, want ** 2, rtol=1e-9)
    # F * var * nc is the extra sum of squares of the model restricted to
    # C beta = 0 (fitted in the null space of C)
    F = orc.fl_stat(1, x, mask, X, C).ravel()[idx]
    Nc = scipy.linalg.null_space(C)
    Xr = X @ Nc
    br, *_ = np.linalg.lstsq(Xr, xs, rcond=None)
    extra = ((xs - Xr @ br) ** 2).sum(axis=0) - (e ** 2).sum(axis=0)
    np.testing.assert_allclose(F * var * 2, extra, rtol=1e-8)


def test_prepare_inverse_whitening_with_identity_gives_the_residuals():
    st = first_level_study(seed=17)
    wh, ar = orc.fl_prepare(st["Y"], st["mask"], st["X"], 8.0, st["voxel"])
    e = orc.fl_regress(st["Y"].astype(np.float64), st["mask"], st["X"])
    ident = np.arange(st["Y"].shape[0], dtype=np.uint16)
    m = st["mask"] == 1
    back = orc.fl_color(wh, ar, st["mask"], ident)
    np.testing.assert_allclose(back[:, m], e[:, m], atol=1e-9)
    assert np.all(ar[:, ~m] == 0)
    # the identity permutation refits residuals: every t is 0
    md = orc.fl_null(0, 0, wh, ar, st["mask"], st["X"], st["C"], ident[None, :])
    assert md[0] == 0.0


def test_prepare_estimates_a_uniform_ar1_process():
    st = first_level_study(seed=19, T=1500, ar=(0.4, 0.0, 0.0, 0.0), noise_only=True)
    _, ar = orc.fl_prepare(st["Y"], st["mask"], st["X"], 8.0, st["voxel"])
    m = st["mask"] == 1
    assert abs(np.median(ar[0][m]) - 0.4) < 0.03
    assert np.all(np.abs(np.median(ar[1:, m], axis=1)) < 0.03)


def test_generator_matches_the_reference_and_its_invariants():
    T, K = 40, 300
    perm = orc.gen_perm_first_level(T, K, seed=1)
    ident = np.arange(T)
    assert all(np.array_equal(np.sort(r), ident) for r in perm)
    assert len({r.tobytes() for r in perm}) == K
    assert not any(np.array_equal(r, ident) for r in perm)
    ref = orc.ref_gen_perm_first_level(T, K, seed=1)
    if ref is None:
        pytest.skip("oracle/_ref not built")
    assert np.array_equal(ref, perm)


def test_null_voxel_max_is_the_max_of_the_maps():
    st = first_level_study(seed=23)
    wh, ar = orc.fl_prepare(st["Y"], st["mask"], st["X"], 8.0, st["voxel"])
    perm = orc.gen_perm_first_level(st["Y"].shape[0], 6, seed=3)
    md, stats = orc.fl_null(0, 0, wh, ar, st["mask"], st["X"], st["C"], perm, want_stats=True)
    m = st["mask"] == 1
    want = np.trunc(stats[:, m].max(axis=1) * 1e4) / 1e4
    np.testing.assert_allclose(md, want.astype(np.float32), rtol=1e-7)
