"""Pins the CPU oracle before it is trusted to check the HIP path.

Each check is chosen so a plausible mistake (a dropped term, a wrong sign or
index, a transposed operand) fails it:
- reference vectors: outputs of the reference's own compiled routines
  (tests/golden/reference_vectors.npz, written by make_golden.py from
  oracle/_ref) and the inputs the reference's MATLAB test holds;
- library routines: glibc rand(), libstdc++ std::random_shuffle,
  numpy.linalg, scipy.stats, scipy.ndimage.label;
- closed forms and invariants: one-/two-sample t, F = t^2, one-way ANOVA,
  Frisch-Waugh-Lovell, projector identities, sign-flip symmetry.
"""
import ctypes
import subprocess
import tempfile
from pathlib import Path

import numpy as np
import pytest
from scipy import ndimage, stats

from oracle import oracle

GOLD = np.load(Path(__file__).resolve().parent / "golden" / "reference_vectors.npz")


def grid_study(N, X, C, test, W=9, H=7, D=5, seed=0, mask=None):
    rng = np.random.default_rng(seed)
    if mask is None:
        mask = np.ones((D, H, W), np.float32)
        mask[0, 0, :3] = 0.0
    Y = rng.standard_normal((N, D, H, W)).astype(np.float32)
    return dict(W=W, H=H, D=D, N=N, p=X.shape[1], nc=C.shape[0], test=test, X=X.astype(np.float32),
                C=C.astype(np.float32), Y=Y, mask=mask)


def vox(st):
    return st["mask"].ravel() == 1.0


# ------------------------------------------------------------------ generators

def test_glibc_rand_known_answer():
    """The generators draw from glibc rand(); its default stream is published."""
    lib = oracle.lib()
    lib.orc_srand(1)
    assert [lib.orc_rand() for _ in range(3)] == [1804289383, 846930886, 1681692777]


def test_shuffle_matches_libstdcxx_random_shuffle():
    """orc_random_shuffle_u16 restates std::random_shuffle (LIB:16978, 17042);
    compare with the real libstdc++ on this host."""
    src = r'''
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>
int main() { srand(7); std::vector<unsigned short> v(23);
  for (int i = 0; i < 23; i++) v[i] = i;
  for (int r = 0; r < 50; r++) { std::random_shuffle(v.begin(), v.end());
    for (auto x : v) printf("%d ", x); } return 0; }'''
    with tempfile.TemporaryDirectory() as d:
        cpp = Path(d) / "rs.cpp"
        cpp.write_text(src)
        exe = Path(d) / "rs"
        subprocess.run(["g++", "-std=c++11", "-w", str(cpp), "-o", str(exe)], check=True)
        ref = np.array([int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                                        check=True).stdout.split()])
    lib = oracle.lib()
    lib.orc_srand(7)
    v = np.arange(23, dtype=np.uint16)
    got = []
    for _ in range(50):
        lib.orc_random_shuffle_u16(oracle.P(v), 23)
        got.extend(v.tolist())
    assert np.array_equal(np.array(got), ref)


@pytest.mark.parametrize("key", [k for k in GOLD.files if k.startswith("sign_")])
def test_sign_generator_matches_reference(key):
    N, K = [int(t[1:]) for t in key.split("_")[1:]]
    got = oracle.gen_sign_matrix(N, K, seed=1)
    assert np.array_equal(got, GOLD[key])
    # invariants: row 0 all +1, entries +-1, rows distinct
    assert np.all(got[0] == 1) and np.all(np.abs(got) == 1)
    assert len({r.tobytes() for r in got}) == K


@pytest.mark.parametrize("key", [k for k in GOLD.files if k.startswith("twosample_")])
def test_two_sample_generator_matches_reference(key):
    _, n, nn, k = key.split("_")
    N, n1, K = int(n[1:]), int(nn[1:]), int(k[1:])
    got = oracle.gen_perm_two_sample(N, n1, K, seed=1)
    assert np.array_equal(got, GOLD[key])
    assert np.array_equal(got[0], np.arange(N))
    for row in got:
        assert sorted(row.tolist()) == list(range(N))
        g1 = row[row < n1]
        g2 = row[row >= n1]
        assert np.all(np.diff(g1.astype(int)) > 0) and np.all(np.diff(g2.astype(int)) > 0)
    labels = (got < n1)
    assert len({r.tobytes() for r in labels}) == K


@pytest.mark.parametrize("key", [k for k in GOLD.files if k.startswith("correlation_")])
def test_correlation_generator_matches_reference(key):
    N, K = [int(t[1:]) for t in key.split("_")[1:]]
    got = oracle.gen_perm_correlation(N, K, seed=1)
    assert np.array_equal(got, GOLD[key])
    assert np.array_equal(got[0], np.arange(N))
    assert all(sorted(r.tolist()) == list(range(N)) for r in got)
    assert len({r.tobytes() for r in got}) == K


# ------------------------------------------------------------------ model math

def test_pinv_matches_reference_and_numpy():
    X = GOLD["tt_design"]
    got = oracle.pinv64(X)
    assert np.allclose(got, GOLD["tt_pinv"], rtol=0, atol=1e-10)
    assert np.allclose(got, np.linalg.pinv(X), atol=1e-12)
    assert np.allclose(got @ X, np.eye(9), atol=1e-12)


def reference_projector(X, Crows):
    """LIB:16657-16689 / 16695-16721 with numpy: C_full = [C; U(:,0:p-r)'] from
    the SVD of I - C' pinv(C'), W = X C_full^-1, R = I - W2 pinv(W2)."""
    p = X.shape[1]
    r = Crows.shape[0]
    Ct = Crows.T
    tmp = np.eye(p) - Ct @ np.linalg.inv(Ct.T @ Ct) @ Ct.T
    U, _, _ = np.linalg.svd(tmp)
    Cf = np.vstack([Crows, U[:, :p - r].T])
    Wm = X @ np.linalg.inv(Cf)
    W2 = Wm[:, r:]
    return np.eye(X.shape[0]) - W2 @ np.linalg.inv(W2.T @ W2) @ W2.T, W2


@pytest.mark.parametrize("rows", [[0], [4], [0, 1, 2]])
def test_nuisance_projector(rows):
    X = GOLD["tt_design"]
    if len(rows) == 1:
        C = GOLD["tt_contrast"][None, :] if rows == [0] else np.eye(9)[rows]
    else:
        C = np.eye(9)[rows]
    R = oracle.nuisance_projector64(X, C)
    Rref, W2 = reference_projector(X, C)
    assert np.allclose(R, Rref, atol=1e-11)
    assert np.allclose(R, R.T, atol=1e-12)
    assert np.allclose(R @ R, R, atol=1e-11)
    assert np.allclose(R @ W2, 0.0, atol=1e-10)
    assert round(np.trace(R)) == 10 - (9 - C.shape[0])


def test_model_precompute_matches_cli_formulas():
    """RGL:1137-1194: xtxxt = inv(X'X) X', c inv(X'X) c', (C inv(X'X) C')^-1."""
    X = GOLD["tt_design"].astype(np.float32)
    c = GOLD["tt_contrast"][None, :].astype(np.float32)
    xtxxt, ct = oracle.model_precompute(X, c, False)
    Xd = X.astype(np.float64)
    ixtx = np.linalg.inv(Xd.T @ Xd)
    assert np.allclose(xtxxt.reshape(9, 10), (ixtx @ Xd.T), rtol=1e-5, atol=1e-5)
    assert np.isclose(ct[0], (c @ ixtx @ c.T)[0, 0], rtol=1e-6)
    C3 = np.eye(9, dtype=np.float32)[[0, 1, 2]]
    _, M = oracle.model_precompute(X, C3, True)
    assert np.allclose(M.reshape(3, 3), np.linalg.inv(C3 @ ixtx @ C3.T), rtol=1e-5)


def test_design_detection_and_permutation_counts():
    rng = np.random.default_rng(1)
    N = 60
    g = np.r_[np.ones(30), np.zeros(30)]
    X = np.c_[np.ones(N), g, rng.standard_normal(N)]
    d, n1, n2 = oracle.detect_designs(X, np.array([[0, 1, 0], [0, 0, 1], [1, 0, 0]], float), False)
    assert d.tolist() == [0, 1, 2] and n1[0] == 30 and n2[0] == 30
    # F over group dummies: X C' holds only 0/1 -> the CLI's "two-sample" (RGL:1010-1014)
    grp = np.repeat(np.arange(4), 25)
    Xf = np.column_stack([np.ones(100)] + [(grp == j) for j in (1, 2, 3)]).astype(float)
    d, n1, n2 = oracle.detect_designs(Xf, np.eye(4)[1:], True)
    assert d[0] == 0 and n1[0] == 75 and n2[0] == 25
    lib = oracle.lib()
    assert lib.orc_max_permutations(2, 0, 10, 0) == 1024
    assert lib.orc_max_permutations(0, 0, 12, 5) == 792
    assert lib.orc_stirling_factorial(10) == round(np.sqrt(2 * 3.14 * 10) * (10 / 2.7183) ** 10)
    assert lib.orc_max_permutations(0, 2, 10, 0) == 1


# ------------------------------------------------------------------ statistics

def test_group_mean_sign_flip_is_one_sample_t():
    N = 12
    st = grid_study(N, np.ones((N, 1)), np.ones((1, 1)), 2, seed=3)
    signs = np.ones((3, N), np.float32)
    signs[1, ::2] = -1
    signs[2, :5] = -1
    perms = dict(sign=signs)
    out = oracle.perm_stats(st, perms, 3)
    Y = st["Y"].reshape(N, -1)[:, vox(st)].astype(np.float64)
    for k in range(3):
        ref = stats.ttest_1samp(signs[k][:, None] * Y, 0.0).statistic
        assert np.allclose(out[k][vox(st)], ref, rtol=1e-10)


def test_two_sample_t_and_f_equal_t_squared():
    N = 14
    g = np.r_[np.ones(6), np.zeros(8)]
    X = np.c_[np.ones(N), g]
    st = grid_study(N, X, np.array([[0.0, 1.0]]), 0, seed=4)
    maps, _, _ = oracle.randomise(st, dict(perm=np.arange(N, dtype=np.uint16)[None]), 1)
    Y = st["Y"].reshape(N, -1)[:, vox(st)].astype(np.float64)
    ref = stats.ttest_ind(Y[:6], Y[6:], equal_var=True).statistic
    assert np.allclose(maps[0][vox(st)], ref, rtol=1e-5)
    # F with the same single row: F = t^2 for every permutation
    perm = np.array([np.arange(N), np.random.default_rng(2).permutation(N)], np.uint16)
    tt = oracle.perm_stats(st, dict(perm=perm), 2)
    stF = dict(st, test=1)
    ff = oracle.perm_stats(stF, dict(perm=perm), 2)
    assert np.allclose(ff, tt ** 2, rtol=1e-10)


def test_one_way_anova_f_and_observed_quirk():
    N = 40
    grp = np.repeat(np.arange(4), 10)
    X = np.column_stack([np.ones(N)] + [(grp == j) for j in (1, 2, 3)]).astype(float)
    C = np.eye(4)[1:]
    st = grid_study(N, X, C, 1, seed=5)
    ident = np.arange(N, dtype=np.uint16)[None]
    f_id = oracle.perm_stats(st, dict(perm=ident), 1)[0][vox(st)]
    Y = st["Y"].reshape(N, -1)[:, vox(st)].astype(np.float64)
    ref = stats.f_oneway(*[Y[grp == j] for j in range(4)]).statistic
    assert np.allclose(f_id, ref, rtol=1e-10)
    # observed F divides the centred residuals by N - p - 1 (K1:1634-1636)
    maps, _, _ = oracle.randomise(st, dict(perm=ident), 1)
    p = 4
    assert np.allclose(maps[0][vox(st)], ref * (N - p - 1) / (N - p), rtol=1e-5)


def brute_force_t(Ystar, X, c, perm):
    """OLS of the residualised data on the row-permuted design (A.2)."""
    Xp = X[perm]
    beta, *_ = np.linalg.lstsq(Xp, Ystar, rcond=None)
    e = Ystar - Xp @ beta
    s2 = (e ** 2).sum(0) / (X.shape[0] - X.shape[1])
    return (c @ beta) / np.sqrt(s2 * (c @ np.linalg.inv(X.T @ X) @ c))


def test_permuted_t_with_nuisance_matches_brute_force_and_fwl():
    X = GOLD["tt_design"]
    c = GOLD["tt_contrast"]
    N = 10
    st = grid_study(N, X, c[None, :], 0, seed=6)
    X = st["X"].astype(np.float64)  # the oracle fits the float32 design
    perms = GOLD["tt_permutations"]
    out = oracle.perm_stats(st, dict(perm=perms), 6)
    R = oracle.nuisance_projector64(X, c[None, :])
    Y = st["Y"].reshape(N, -1)[:, vox(st)].astype(np.float64)
    Ystar = R @ Y
    for k in range(6):
        ref = brute_force_t(Ystar, X, c, perms[k].astype(int))
        assert np.allclose(out[k][vox(st)], ref, rtol=1e-9, atol=1e-9)
    # Frisch-Waugh-Lovell: the identity permutation reproduces the observed t
    maps, _, _ = oracle.randomise(st, dict(perm=perms), 6)
    assert np.allclose(out[0][vox(st)], maps[0][vox(st)], rtol=1e-5, atol=1e-5)


def test_exhaustive_sign_flips_are_symmetric():
    """All 2^6 sign flips: the null distribution of t is symmetric."""
    N = 6
    st = grid_study(N, np.ones((N, 1)), np.ones((1, 1)), 2, seed=8)
    signs = GOLD["sign_N6_K64"]
    out = oracle.perm_stats(st, dict(sign=signs), 64)[:, vox(st)]
    key = {r.tobytes(): i for i, r in enumerate(signs)}
    for i, r in enumerate(signs):
        j = key[(-r).tobytes()]
        assert np.allclose(out[i], -out[j], rtol=1e-12)


@pytest.mark.parametrize("test", [0, 1, 2])
def test_ref32_restatement_within_tolerance_of_fp64(test):
    N = 30
    rng = np.random.default_rng(9)
    if test == 2:
        X, C = np.ones((N, 1)), np.ones((1, 1))
        perms = dict(sign=np.where(rng.random((8, N)) < 0.5, -1, 1).astype(np.float32))
    else:
        X = np.c_[np.ones(N), rng.standard_normal((N, 3))]
        C = np.array([[0, 1, 0, 0.0]]) if test == 0 else np.eye(4)[1:3]
        perms = dict(perm=np.array([rng.permutation(N) for _ in range(8)], np.uint16))
    st = grid_study(N, X, C, test, W=11, H=9, D=7, seed=10)
    a = oracle.perm_stats(st, perms, 8, precision=oracle.REF32)[:, vox(st)]
    b = oracle.perm_stats(st, perms, 8, precision=oracle.F64)[:, vox(st)]
    assert np.max(np.abs(a - b) / (1 + np.abs(b))) < 1e-5


# ------------------------------------------------------------------ max-T

def test_quantised_max_semantics():
    mask = np.ones((1, 1, 5), np.float32)
    mask[0, 0, 4] = 0.0
    m = np.array([[[1.23456789, -0.5, 2.00019, 0.1, 99.0]]], np.float32)
    # trunc(T*1e4)/1e4 over the mask (KM:299, LIB:10988); voxel 4 is outside
    assert oracle.max_quantized(m, mask) == np.float32(np.float32(20001) / np.float32(10000))
    neg = np.array([[[-1.23456, -2.5, -0.00019, -7.0, 0.0]]], np.float32)
    assert oracle.max_quantized(neg, mask) == np.float32(np.float32(-1) / np.float32(10000))
    # empty mask: the int buffer keeps the bit pattern of -1e6f (LIB:10971)
    sentinel = np.frombuffer(np.float32(-1e6).tobytes(), np.int32)[0]
    assert oracle.max_quantized(m, np.zeros_like(mask)) == np.float32(np.float32(sentinel) / np.float32(10000))


# ------------------------------------------------------------------ clusters

@pytest.mark.parametrize("i", [0, 1, 2])
def test_ccl_matches_reference_clusterize(i):
    data, mask, thr = GOLD[f"clu{i}_data"], GOLD[f"clu{i}_mask"], float(GOLD[f"clu{i}_thr"])
    ref = GOLD[f"clu{i}_labels"].ravel()
    lab, sz, best = oracle.ccl(data, mask, thr, 26)
    assert best == int(GOLD[f"clu{i}_max"])
    active = ref > 0
    G = lab.size
    assert np.array_equal(lab != 3 * G, active)
    # same partition: the reference numbers clusters 1..n in raster discovery
    # order, whose first voxel is the component minimum (= our label)
    firsts = {}
    for idx in np.flatnonzero(active):
        firsts.setdefault(ref[idx], idx)
    mapped = np.array([firsts[r] for r in ref[active]])
    assert np.array_equal(lab[active], mapped)
    assert len(firsts) == int(GOLD[f"clu{i}_n"])
    for root in np.unique(lab[active]):
        assert sz[root] == np.sum(lab == root)


@pytest.mark.parametrize("conn", [6, 26])
def test_ccl_matches_scipy_label(conn):
    rng = np.random.default_rng(conn)
    vol = ndimage.gaussian_filter(rng.standard_normal((17, 19, 23)), 1.2).astype(np.float32)
    mask = np.ones_like(vol)
    thr = 0.1
    lab, sz, best = oracle.ccl(vol, mask, thr, conn)
    struct = ndimage.generate_binary_structure(3, 1 if conn == 6 else 3)
    ref, n = ndimage.label(vol > thr, structure=struct)
    counts = np.bincount(ref.ravel())[1:]
    assert best == counts.max()
    assert len(np.unique(lab[lab != 3 * vol.size])) == n


def test_cluster_mass_matches_scipy_label_sums():
    """Mass of a component = sum of (uint)(T*1e4) over its voxels
    (KC:432-458), checked on scipy's own labelling with ndimage.sum."""
    rng = np.random.default_rng(11)
    vol = (3 * ndimage.gaussian_filter(rng.standard_normal((15, 18, 21)), 1.5)).astype(np.float32)
    mask = (rng.random(vol.shape) < 0.95).astype(np.float32)
    thr = 0.3
    lab, _, _ = oracle.ccl(vol, mask, thr, 26)
    ms, best = oracle.cluster_mass(vol, mask, thr, lab)
    act = (vol > thr) & (mask == 1)
    ref, n = ndimage.label(act, structure=ndimage.generate_binary_structure(3, 3))
    q = np.where(act, (vol * np.float32(1e4)).astype(np.uint32), 0).astype(np.int64)
    sums = ndimage.sum(q, ref, index=np.arange(1, n + 1)).astype(np.int64)
    assert n > 3 and best == sums.max()
    # every component's mass sits at its label
    for r in range(1, n + 1):
        first = np.flatnonzero(ref.ravel() == r)[0]
        assert ms[lab[first]] == sums[r - 1]


def _tfce_scipy(vol, mask, conn):
    # thresholds 0, 0.2846, ... accumulated in float32 up to the quantised max
    # (LIB:16756-16762, LIB:18470); scipy labels each level, sizes by bincount
    top = np.float32(np.trunc(np.float32(vol[mask == 1].max()) * np.float32(1e4)) / np.float32(1e4))
    acc = np.zeros(vol.shape, np.float32)
    struct = ndimage.generate_binary_structure(3, 1 if conn == 6 else 3)
    h = np.float32(0.0)
    while h <= top:
        lab, n = ndimage.label((vol > h) & (mask == 1), structure=struct)
        sizes = np.bincount(lab.ravel()).astype(np.float32)
        val = (np.sqrt(sizes[lab]) * h) * h
        acc = np.where(lab > 0, (acc + val).astype(np.float32), acc)
        h = np.float32(h + np.float32(0.2846))
    q = np.trunc(acc[mask == 1].max() * np.float32(1e4))
    return np.float32(np.float32(q) / np.float32(1e4)), acc


@pytest.mark.parametrize("conn", [26, 6])
def test_tfce_matches_scipy_levels(conn):
    rng = np.random.default_rng(40 + conn)
    vol = (2.5 * ndimage.gaussian_filter(rng.standard_normal((13, 15, 17)), 1.3)).astype(np.float32)
    mask = (rng.random(vol.shape) < 0.97).astype(np.float32)
    m, vals = oracle.tfce_max(vol, mask, conn, want_values=True)
    ref, acc = _tfce_scipy(vol, mask, conn)
    assert np.array_equal(vals.reshape(vol.shape)[mask == 1], acc[mask == 1])
    assert m == ref and m > 0


def test_tfce_uniform_cube_closed_form():
    # one cube of n voxels at value v: sum over h_j < v of sqrt(n) h_j^2
    vol = np.zeros((8, 8, 8), np.float32)
    vol[2:5, 2:6, 1:4] = 1.7
    n = 3 * 4 * 3
    mask = np.ones_like(vol)
    m = oracle.tfce_max(vol, mask, 26)
    hs = 0.2846 * np.arange(6)  # 0 .. 1.423 < 1.7 <= 1.7076
    expect = np.sqrt(n) * np.sum(hs ** 2)
    assert abs(m - expect) <= 2e-4


def test_ccl_diagonal_contacts_separate_6_and_26():
    vol = np.zeros((5, 5, 5), np.float32)
    vol[0, 0, 0] = vol[1, 1, 1] = vol[2, 2, 2] = 1.0  # corner-to-corner chain
    vol[4, 4, 0] = vol[4, 4, 1] = 1.0                  # face-connected pair
    mask = np.ones_like(vol)
    _, _, b26 = oracle.ccl(vol, mask, 0.5, 26)
    _, _, b6 = oracle.ccl(vol, mask, 0.5, 6)
    assert b26 == 3 and b6 == 2
    lab, _, _ = oracle.ccl(vol, mask, 0.5, 26)
    assert lab[31] == 0 and lab[62] == 0           # the chain carries its minimum index
    assert lab[120] == 120 and lab[121] == 120     # z=4, y=4, x=0/1
    assert lab[1] == 3 * vol.size                  # background


def test_pvalues_count():
    rng = np.random.default_rng(3)
    D, H, W = 4, 5, 6
    mp = rng.standard_normal((D, H, W)).astype(np.float32)
    mask = (rng.random((D, H, W)) < 0.8).astype(np.float32)
    md = np.sort(rng.standard_normal(50).astype(np.float32))
    P = np.zeros(D * H * W, np.float32)
    oracle.lib().orc_pvalues(0, W, H, D, oracle.P(mp), oracle.P(mask), None, None, ctypes.c_float(0.0), 50,
                             oracle.P(md), oracle.P(P))
    m = mask.ravel() == 1
    cnt = (mp.ravel()[m][:, None] > md[None, :]).sum(1)
    assert np.array_equal(P[m], cnt.astype(np.float32) / np.float32(50))
    assert np.all(P[~m] == 0)


@pytest.mark.parametrize("wl", [1, 2, 3])
def test_perm_max64_is_the_max_of_the_fp64_maps(wl):
    """orc_perm_max64 (the full-size accuracy check's FP64 maxima) equals the
    max over the in-mask voxels of orc_perm_stats' FP64 maps (pinned above
    against scipy, closed forms and brute-force least squares)."""
    from broccoli_amd import synth

    st = synth.study(wl, scale=0.12)
    K = 24
    perms = synth.permutations(st, K, seed=41)
    ref = oracle.perm_stats(st, perms, K)[:, st["mask"].ravel() == 1.0]
    got = oracle.perm_max64(st, perms, K)
    assert np.allclose(got, ref.max(axis=1), rtol=1e-12, atol=1e-12)


def test_vectorised_quantised_max_is_the_oracle_quantisation():
    """oracle.quantised_max equals orc_quantize (the C restatement of
    (int)(T * 10000.0f), KM:299) followed by / 10000.0f, LIB:10988, on
    values across the range, including products past 2^24 where FP32
    rounding of T * 1e4 matters; beyond the int range it saturates."""
    rng = np.random.default_rng(3)
    t = np.concatenate([rng.normal(0, 5, 2000), rng.uniform(-3e3, 3e3, 2000), rng.uniform(1e3, 2e5, 2000),
                        [0.0, 1e-5, -1e-5, 4.99995, 525.5731937, 214748.3]]).astype(np.float32)
    lib = oracle.lib()
    want = np.array([np.float32(lib.orc_quantize(float(x))) / np.float32(10000.0) for x in t], np.float32)
    assert np.array_equal(oracle.quantised_max(t), want)
    assert oracle.quantised_max(np.array([3e5]))[0] == np.float32(np.float32(2 ** 31 - 1) / np.float32(10000.0))
