"""Parity of the first-level permutation test on the GPU (through the C ABI)
with the FP64 oracle (oracle/fl_oracle.c).

Reference: LIB:16454-16558 (ApplyPermutationTestFirstLevel), with the
whitening of LIB:16363-16452 and the kernels of kernelWhitening.cpp and
kernelStatistics3/5.cpp.  Accuracy gate as for the second level:
|stat - stat_fp64| <= 1e-4 (1 + |stat|) for every permuted statistic of every
in-mask voxel; cluster sizes and masses bit-identical on the GPU's own
statistic values; the set-up (AR estimates, whitened volumes) to FP64
round-off.
"""
import numpy as np
import pytest

from broccoli_amd import capi, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-4


def fl_study(seed=0, **kw):
    st = synth.first_level_study(seed=seed, **kw)
    return st


def in_mask(st):
    return st["mask"] == 1.0


def oracle_null(st, perm, test=0, inference=0, contrast=0, want_stats=True):
    wh, ar = oracle.fl_prepare(st["Y"], st["mask"], st["X"], st.get("fwhm", 8.0), st["voxel"])
    return oracle.fl_null(test, inference, wh, ar, st["mask"], st["X"], st["C"], perm, contrast=contrast,
                          cdt=st.get("cdt", 0.0), connectivity=st.get("connectivity", 26), want_stats=want_stats)


def check_stats(stats_gpu, stats_ref, m):
    ref = stats_ref[:, m]
    err = np.abs(stats_gpu.astype(np.float64) - ref)
    assert np.all(err <= TOL * (1.0 + np.abs(ref))), f"max rel err {np.max(err / (1 + np.abs(ref)))}"
    return float(np.max(err / (1 + np.abs(ref))))


def check_max(md, stats_gpu):
    """The null distribution is the quantised max of the GPU's statistics."""
    want = np.trunc(stats_gpu.max(axis=1).astype(np.float32) * np.float32(1e4)).astype(np.float32) / np.float32(1e4)
    np.testing.assert_array_equal(md, want)


def test_setup_matches_the_oracle(ctx):
    st = fl_study(seed=1)
    print("first-level set-up, grid", st["Y"].shape, flush=True)
    wh, ar = oracle.fl_prepare(st["Y"], st["mask"], st["X"], 8.0, st["voxel"])
    m = in_mask(st)
    with capi.FlPlan(ctx, st) as plan:
        ar_g = plan.ar_estimates()
        wh_g = plan.whitened()
        assert plan.num_voxels == int(m.sum())
    np.testing.assert_allclose(ar_g[:, m], ar[:, m], rtol=1e-6, atol=1e-7)
    assert np.all(ar_g[:, ~m] == 0)
    np.testing.assert_allclose(wh_g[:, m], wh[:, m], rtol=1e-6, atol=1e-5)


PREC = {"split_f16": capi.PREC_SPLIT_F16, "split_bf16": capi.PREC_SPLIT_BF16, "fp32": capi.PREC_FP32,
        "fp64": capi.PREC_FP64}
ALL = list(PREC)


@pytest.mark.parametrize("prec", ALL)
@pytest.mark.parametrize("seed,T,p", [(2, 64, 5), (3, 120, 8), (4, 48, 1), (14, 242, 12)])
def test_permuted_t_statistics_match_fp64(ctx, seed, T, p, prec):
    st = fl_study(seed=seed, T=T, p=p, ar_gradient=0.2)
    st["precision"] = prec
    if p == 1:
        st["C"] = np.ones((1, 1), np.float32)
    K = 70
    perm = capi.gen_perm_first_level(T, K, seed=seed)
    print(f"first-level t seed {seed} T {T} p {p} K {K} {prec}", flush=True)
    md_ref, stats_ref = oracle_null(st, perm)
    with capi.FlPlan(ctx, st) as plan:
        md, stats = plan.permutation_phase(0, perm, want_stats=True)
        tm = plan.timing()
    assert tm.precision_used == PREC[prec]
    assert tm.unsafe_waves == 0
    check_stats(stats, stats_ref, in_mask(st))
    check_max(md, stats)
    np.testing.assert_allclose(md, md_ref, atol=2e-4, rtol=1e-5)


@pytest.mark.parametrize("prec", ALL)
def test_design_without_an_intercept(ctx, prec):
    # the residuals need not sum to zero: Welford's mean term is kept
    st = fl_study(seed=12, T=61, p=5)
    st["precision"] = prec
    st["X"] = np.ascontiguousarray(st["X"][:, 1:])
    st["C"] = np.ascontiguousarray(st["C"][:, 1:])
    K = 30
    perm = capi.gen_perm_first_level(61, K, seed=6)
    print("first-level design without intercept, T=61", flush=True)
    _, stats_ref = oracle_null(st, perm)
    with capi.FlPlan(ctx, st) as plan:
        md, stats = plan.permutation_phase(0, perm, want_stats=True)
    check_stats(stats, stats_ref, in_mask(st))
    check_max(md, stats)


@pytest.mark.parametrize("prec", ALL)
def test_f_statistics_and_second_contrast(ctx, prec):
    st = fl_study(seed=5, T=80, p=6)
    st["precision"] = prec
    C = np.zeros((2, 6), np.float32)
    C[0, 2] = 1.0
    C[1, 3] = 1.0
    C[1, 4] = -1.0
    st["C"] = C
    K = 40
    perm = capi.gen_perm_first_level(80, K, seed=9)
    m = in_mask(st)
    print("first-level F and t contrast 1", flush=True)
    # t of contrast 1
    _, stats_ref = oracle_null(st, perm, test=0, contrast=1)
    with capi.FlPlan(ctx, st) as plan:
        _, stats = plan.permutation_phase(1, perm, want_stats=True)
    check_stats(stats, stats_ref, m)
    # F over both rows
    st["test"] = capi.FTEST
    _, stats_ref = oracle_null(st, perm, test=1)
    with capi.FlPlan(ctx, st) as plan:
        md, stats = plan.permutation_phase(0, perm, want_stats=True)
    check_stats(stats, stats_ref, m)
    check_max(md, stats)


def test_identity_permutation_refits_residuals(ctx):
    st = fl_study(seed=6, T=60)
    ident = np.arange(60, dtype=np.uint16)[None, :]
    print("first-level identity permutation", flush=True)
    with capi.FlPlan(ctx, st) as plan:
        md, stats = plan.permutation_phase(0, ident, want_stats=True)
    # inverse whitening rebuilds the residuals, which are orthogonal to X
    assert np.max(np.abs(stats)) < 1e-3


def test_long_series_read_from_global_memory(ctx):
    # T above the LDS tile limit: the whitened volumes are read from HBM
    st = fl_study(seed=7, W=7, H=6, D=5, T=640, p=4)
    K = 24
    perm = capi.gen_perm_first_level(640, K, seed=2)
    print("first-level T=640", flush=True)
    _, stats_ref = oracle_null(st, perm)
    with capi.FlPlan(ctx, st) as plan:
        md, stats = plan.permutation_phase(0, perm, want_stats=True)
    check_stats(stats, stats_ref, in_mask(st))
    check_max(md, stats)


@pytest.mark.parametrize("p", [14, 25])
def test_wide_designs(ctx, p):
    st = fl_study(seed=8, T=90, p=p)
    K = 20
    perm = capi.gen_perm_first_level(90, K, seed=4)
    print(f"first-level p={p}", flush=True)
    _, stats_ref = oracle_null(st, perm)
    with capi.FlPlan(ctx, st) as plan:
        md, stats = plan.permutation_phase(0, perm, want_stats=True)
    check_stats(stats, stats_ref, in_mask(st))
    check_max(md, stats)


@pytest.mark.parametrize("inference,conn,prec", [(capi.CLUSTER_EXTENT, 26, "split_f16"),
                                                 (capi.CLUSTER_MASS, 6, "split_f16"),
                                                 (capi.CLUSTER_EXTENT, 26, "split_bf16"),
                                                 (capi.CLUSTER_MASS, 26, "split_bf16"),
                                                 (capi.CLUSTER_EXTENT, 26, "fp32"), (capi.CLUSTER_MASS, 26, "fp32"),
                                                 (capi.CLUSTER_EXTENT, 6, "fp64"), (capi.CLUSTER_MASS, 26, "fp64")])
def test_cluster_inference_on_the_gpu_statistics(ctx, inference, conn, prec):
    st = fl_study(seed=10, W=16, H=16, D=12, T=60)
    st["precision"] = prec
    st["inference"] = inference
    st["connectivity"] = conn
    st["cdt"] = 1.5
    K = 30
    perm = capi.gen_perm_first_level(60, K, seed=5)
    m = in_mask(st)
    print(f"first-level cluster inference {inference} conn {conn}", flush=True)
    with capi.FlPlan(ctx, st) as plan:
        md, stats = plan.permutation_phase(0, perm, want_stats=True)
    G = st["mask"].size
    for k in range(K):
        g = np.zeros(G, np.float32)
        g[np.flatnonzero(m.ravel())] = stats[k]
        g = g.reshape(st["mask"].shape)
        lab, _, best = oracle.ccl(g, st["mask"], st["cdt"], conn)
        if inference == capi.CLUSTER_EXTENT:
            assert md[k] == np.float32(best), k
        else:
            _, bm = oracle.cluster_mass(g, st["mask"], st["cdt"], lab)
            assert md[k] == np.float32(bm) / np.float32(1e4), k


def test_invalid_inputs_are_refused(ctx):
    st = fl_study(seed=11, T=40)
    st2 = dict(st, inference=capi.TFCE)
    with pytest.raises(capi.BpgError) as e:
        capi.FlPlan(ctx, st2)
    assert e.value.code == -5
    with capi.FlPlan(ctx, st) as plan:
        bad = np.tile(np.arange(40, dtype=np.uint16), (2, 1))
        bad[1, 3] = 40
        with pytest.raises(capi.BpgError):
            plan.permutation_phase(0, bad)
        # the plan still works afterwards
        md, _ = plan.permutation_phase(0, capi.gen_perm_first_level(40, 4, seed=1))
        assert np.all(np.isfinite(md))


@pytest.mark.parametrize("inference", [capi.VOXEL, capi.CLUSTER_MASS])
def test_empty_and_single_voxel_masks(ctx, inference):
    st = fl_study(seed=13, T=40)
    st["inference"] = inference
    st["cdt"] = 1.0
    perm = capi.gen_perm_first_level(40, 10, seed=8)
    print(f"first-level empty / single-voxel mask, inference {inference}", flush=True)
    empty = dict(st, mask=np.zeros_like(st["mask"]))
    with capi.FlPlan(ctx, empty) as plan:
        assert plan.num_voxels == 0
        md, _ = plan.permutation_phase(0, perm)
        assert np.all(plan.ar_estimates() == 0)
    if inference == capi.VOXEL:
        s = np.frombuffer(np.float32(-1e6).tobytes(), np.int32)[0]
        assert np.all(md == np.float32(np.float32(s) / np.float32(10000)))
    else:
        assert np.all(md == 0)
    one = np.zeros_like(st["mask"])
    one[5, 7, 6] = 1.0
    single = dict(st, mask=one)
    _, stats_ref = oracle_null(single, perm)
    with capi.FlPlan(ctx, single) as plan:
        assert plan.num_voxels == 1
        md, stats = plan.permutation_phase(0, perm, want_stats=True)
    check_stats(stats, stats_ref, one == 1.0)


@pytest.mark.parametrize("prec", ["split_f16", "split_bf16", "fp32"])
def test_ill_conditioned_permutations_repeat_in_fp64(ctx, prec):
    # 16 regressors on 20 time points: some samples keep less than 1/64 of
    # |x|^2 as residual, which the fast forms hand to the FP64 form
    st = fl_study(seed=15, T=20, p=16)
    st["precision"] = prec
    K = 24
    perm = capi.gen_perm_first_level(20, K, seed=7)
    print("first-level ill-conditioned, T=20 p=16", flush=True)
    _, stats_ref = oracle_null(st, perm)
    with capi.FlPlan(ctx, st) as plan:
        md, stats = plan.permutation_phase(0, perm, want_stats=True)
        tm = plan.timing()
    assert tm.precision_used == PREC[prec]
    assert tm.unsafe_waves > 0 and tm.unsafe_waves + tm.fp32_permutations == K
    check_stats(stats, stats_ref, in_mask(st))
    check_max(md, stats)


@pytest.mark.parametrize("prec", ["split_f16", "split_bf16", "fp32"])
def test_fast_forms_fall_back_to_fp64_outside_their_limits(ctx, prec):
    # the split-FP16 form covers 25 regressors (two blocks of 16 basis
    # columns, T <= 384); the other fast forms 16
    cases = [({"T": 90, "p": 25}, capi.PREC_SPLIT_F16 if prec == "split_f16" else capi.PREC_FP64),
             ({"W": 7, "H": 6, "D": 5, "T": 400, "p": 20}, capi.PREC_FP64),
             ({"W": 7, "H": 6, "D": 5, "T": 600, "p": 4}, capi.PREC_FP64)]
    for kw, want in cases:
        st = fl_study(seed=16, **kw)
        st["precision"] = prec
        perm = capi.gen_perm_first_level(kw["T"], 8, seed=1)
        print(f"first-level fallback {kw}", flush=True)
        with capi.FlPlan(ctx, st) as plan:
            plan.permutation_phase(0, perm)
            assert plan.timing().precision_used == want


def test_split_f16_hands_long_series_to_the_split_bf16_form(ctx):
    # 480 < T <= 560: the FP16 basis no longer fits in LDS beside the tile
    st = fl_study(seed=17, W=7, H=6, D=5, T=500, p=6)
    K = 16
    perm = capi.gen_perm_first_level(500, K, seed=3)
    print("first-level split_f16 at T=500", flush=True)
    _, stats_ref = oracle_null(st, perm)
    with capi.FlPlan(ctx, st) as plan:
        md, stats = plan.permutation_phase(0, perm, want_stats=True)
        assert plan.timing().precision_used == capi.PREC_SPLIT_BF16
    check_stats(stats, stats_ref, in_mask(st))
    check_max(md, stats)


@pytest.mark.parametrize("scale", [1e-7, 3e7])
@pytest.mark.parametrize("prec", ["split_f16", "split_bf16"])
def test_statistics_do_not_depend_on_the_data_scale(ctx, scale, prec):
    # the split-FP16 form rescales each series by a power of two; the
    # statistic is scale-free, so tiny and huge intensities give the same
    # maps as the oracle (strong AR(1): a wide dynamic range in x)
    st = fl_study(seed=18, T=96, p=6, ar=(0.85, 0.05, 0.0, 0.0))
    st["Y"] = (st["Y"].astype(np.float64) * scale).astype(np.float32)
    st["precision"] = prec
    K = 24
    perm = capi.gen_perm_first_level(96, K, seed=11)
    print(f"first-level data scale {scale} {prec}", flush=True)
    _, stats_ref = oracle_null(st, perm)
    with capi.FlPlan(ctx, st) as plan:
        md, stats = plan.permutation_phase(0, perm, want_stats=True)
        assert plan.timing().precision_used == PREC[prec]
    check_stats(stats, stats_ref, in_mask(st))
    check_max(md, stats)


def test_near_unit_root_series_match_fp64(ctx):
    # AR(1) near 1: the smoothed estimates come out above 1 (1.18..1.40 here),
    # so inverse whitening grows without bound; the samples whose |x|^2
    # leaves the FP16 range are flagged and every permutation is repeated in
    # FP64, which still matches the oracle
    st = fl_study(seed=19, T=128, p=4, ar=(0.97, 0.02, 0.0, 0.0))
    st["precision"] = "split_f16"
    K = 32
    perm = capi.gen_perm_first_level(128, K, seed=13)
    print("first-level near unit root, split_f16", flush=True)
    _, stats_ref = oracle_null(st, perm)
    with capi.FlPlan(ctx, st) as plan:
        md, stats = plan.permutation_phase(0, perm, want_stats=True)
        tm = plan.timing()
        ar = plan.ar_estimates()
    m = in_mask(st)
    print(f"AR1 estimates {ar[0][m].min():.3f}..{ar[0][m].max():.3f}, repeated in FP64: {tm.unsafe_waves}", flush=True)
    assert tm.unsafe_waves > 0 and tm.unsafe_waves + tm.fp32_permutations == K
    check_stats(stats, stats_ref, m)
    check_max(md, stats)


@pytest.mark.parametrize("test", [0, 1])
def test_split_f16_two_column_blocks(ctx, test):
    # 20 regressors without an intercept (the mean term), t and a 3-row F:
    # the split-FP16 form with two blocks of 16 basis columns
    st = fl_study(seed=20, T=120, p=21)
    st["X"] = np.ascontiguousarray(st["X"][:, 1:])
    p = st["X"].shape[1]
    C = np.zeros((3, p), np.float32)
    C[0, 1] = 1.0
    C[1, 2] = 1.0
    C[2, 17] = 1.0
    st["C"] = C[:1] if test == 0 else C
    st["test"] = capi.TTEST if test == 0 else capi.FTEST
    st["precision"] = "split_f16"
    K = 24
    perm = capi.gen_perm_first_level(120, K, seed=17)
    print(f"first-level split_f16 p={p} test {test}", flush=True)
    _, stats_ref = oracle_null(st, perm, test=test)
    with capi.FlPlan(ctx, st) as plan:
        md, stats = plan.permutation_phase(0, perm, want_stats=True)
        tm = plan.timing()
    assert tm.precision_used == capi.PREC_SPLIT_F16
    check_stats(stats, stats_ref, in_mask(st))
    check_max(md, stats)
