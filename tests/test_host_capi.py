"""CPU tests of the product's host side: the C ABI library loads and exports
every declared symbol, and its host pieces (generators, design detection,
permutation counts, model precompute, argument checks) agree with the
pinned oracle and the reference vectors.  No GPU calls."""
import os
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from broccoli_amd import capi
from oracle import oracle

ROOT = Path(__file__).resolve().parents[1]
GOLD = np.load(ROOT / "tests" / "golden" / "reference_vectors.npz")


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    declared = set(re.findall(r"\b(bpg_[a-z0-9_]+)\s*\(", (ROOT / "include" / "bpg.h").read_text()))
    assert declared, "no declarations parsed"
    assert declared == set(capi.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(capi.LIB_PATH)], capture_output=True, text=True).stdout
    for name in declared:
        assert re.search(rf"\bT {name}\b", out), name


def test_binding_argument_counts_match_declarations():
    """Each ctypes signature has as many arguments as the C declaration."""
    lib = capi.load()
    text = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "bpg.h").read_text(), flags=re.S)
    decls = re.findall(r"\b(bpg_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", text)
    assert len(decls) >= len(capi.EXPORTS)
    for name, params in decls:
        params = params.strip()
        n = 0 if params in ("", "void") else params.count(",") + 1
        assert len(getattr(lib, name).argtypes) == n, name


@pytest.mark.parametrize("cname,pyname", [("bpg_timing", "Timing"), ("bpg_study_desc", "StudyDesc")])
def test_binding_structures_mirror_declarations(cname, pyname):
    """The ctypes structures name the C structure's fields in its order (a
    field added in C and not in the binding would shift every later one)."""
    text = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "bpg.h").read_text(), flags=re.S)
    body = re.search(r"typedef struct\s*\w*\s*\{([^{}]*)\}\s*" + cname + r"\s*;", text).group(1)
    names = []
    for decl in body.split(";"):
        decl = decl.strip()
        if decl:
            # "type a, b", "type* p", "type a[3]": the names after the type
            for part in decl.split(","):
                names.append(re.sub(r"\[.*\]", "", part).strip().split()[-1].lstrip("*"))
    fields = [f[0] for f in getattr(capi, pyname)._fields_]
    assert fields == names


def test_dropin_class_symbols_exported():
    out = subprocess.run(["nm", "-DC", "--defined-only", str(capi.LIB_PATH)], capture_output=True,
                         text=True).stdout
    for m in ["BROCCOLI_LIB::PerformGLMTTestSecondLevelPermutationWrapper()",
              "BROCCOLI_LIB::PerformGLMFTestSecondLevelPermutationWrapper()",
              "BROCCOLI_LIB::PerformMeanSecondLevelPermutationWrapper()",
              "BROCCOLI_LIB::BROCCOLI_LIB(unsigned int, unsigned int, int, bool)",
              "BROCCOLI_LIB::SetDesignMatrix(float*, float*)",
              "BROCCOLI_LIB::SetPermutationMatrices(unsigned short**)",
              "BROCCOLI_LIB::GetOpenCLRunKernelErrors()",
              "BROCCOLI_LIB::PerformGLMTTestSecondLevelWrapper()",
              "BROCCOLI_LIB::PerformGLMFTestSecondLevelWrapper()",
              "BROCCOLI_LIB::SetOutputBetaVolumesMNI(float*)",
              "BROCCOLI_LIB::SetOutputResidualsMNI(float*)",
              "BROCCOLI_LIB::SetOutputResidualVariances(float*)"]:
        assert m in out, m


@pytest.mark.parametrize("key", [k for k in GOLD.files if k.startswith("sign_")])
def test_compat_sign_generator_is_the_reference_stream(key):
    N, K = [int(t[1:]) for t in key.split("_")[1:]]
    assert np.array_equal(capi.gen_sign_matrix(N, K, compat=True, seed=1), GOLD[key])


@pytest.mark.parametrize("key", [k for k in GOLD.files if k.startswith("twosample_")])
def test_compat_two_sample_generator_is_the_reference_stream(key):
    _, n, nn, k = key.split("_")
    N, n1, K = int(n[1:]), int(nn[1:]), int(k[1:])
    assert np.array_equal(capi.gen_perm_two_sample(N, n1, K, compat=True, seed=1), GOLD[key])


@pytest.mark.parametrize("key", [k for k in GOLD.files if k.startswith("correlation_")])
def test_compat_correlation_generator_is_the_reference_stream(key):
    N, K = [int(t[1:]) for t in key.split("_")[1:]]
    assert np.array_equal(capi.gen_perm_correlation(N, K, compat=True, seed=1), GOLD[key])


@pytest.mark.parametrize("key", [k for k in GOLD.files if k.startswith("firstlevel_")])
def test_compat_first_level_generator_is_the_reference_stream(key):
    T, K = [int(t[1:]) for t in key.split("_")[1:]]
    assert np.array_equal(capi.gen_perm_first_level(T, K, compat=True, seed=1), GOLD[key])
    assert np.array_equal(oracle.gen_perm_first_level(T, K, seed=1), GOLD[key])


def test_first_level_generator_invariants():
    a = capi.gen_perm_first_level(50, 400, compat=False, seed=3)
    assert np.array_equal(a, capi.gen_perm_first_level(50, 400, compat=False, seed=3))
    ident = np.arange(50)
    assert all(np.array_equal(np.sort(r), ident) for r in a)
    assert len({r.tobytes() for r in a}) == 400
    assert not any(np.array_equal(r, ident) for r in a)
    with pytest.raises(capi.BpgError):
        capi.gen_perm_first_level(3, 6, compat=False, seed=1)  # 3! - 1 = 5 non-identity rows


@pytest.mark.parametrize("seed", [1, 2, 12345, 2**31 - 1, 4000000000])
def test_private_rand_stream_is_glibc_rand(seed):
    """The compat stream restates glibc's rand(): compare long draws with the
    real rand() (through the oracle, which calls it) for several seeds."""
    a = capi.gen_perm_correlation(50, 300, compat=True, seed=seed)
    b = oracle.gen_perm_correlation(50, 300, seed=seed)
    assert np.array_equal(a, b)
    s1 = capi.gen_sign_matrix(30, 200, compat=True, seed=seed)
    s2 = oracle.gen_sign_matrix(30, 200, seed=seed)
    assert np.array_equal(s1, s2)


def test_compat_stream_continues_across_calls():
    """seed == 0 continues the stream, as successive contrasts do in the
    reference (the generator is called once per contrast)."""
    first = capi.gen_perm_correlation(20, 50, compat=True, seed=7)
    second = capi.gen_perm_correlation(20, 50, compat=True, seed=0)
    lib = oracle.lib()
    lib.orc_srand(7)
    o1 = np.zeros((50, 20), np.uint16)
    o2 = np.zeros((50, 20), np.uint16)
    import ctypes
    lib.orc_gen_perm_correlation(20, ctypes.c_long(50), oracle.P(o1))
    lib.orc_gen_perm_correlation(20, ctypes.c_long(50), oracle.P(o2))
    assert np.array_equal(first, o1) and np.array_equal(second, o2)


def test_seeded_generators_invariants_and_reproducibility():
    a = capi.gen_perm_correlation(40, 500, compat=False, seed=99)
    b = capi.gen_perm_correlation(40, 500, compat=False, seed=99)
    c = capi.gen_perm_correlation(40, 500, compat=False, seed=100)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert np.array_equal(a[0], np.arange(40))
    assert all(sorted(r.tolist()) == list(range(40)) for r in a)
    assert len({r.tobytes() for r in a}) == 500
    s = capi.gen_sign_matrix(10, 1024, compat=False, seed=5)  # all 2^10 sign vectors
    assert len({r.tobytes() for r in s}) == 1024 and np.all(s[0] == 1)
    t = capi.gen_perm_two_sample(12, 5, 792, compat=False, seed=1)  # all C(12,5)
    assert len({(r < 5).tobytes() for r in t}) == 792


def test_generators_refuse_more_permutations_than_exist():
    with pytest.raises(capi.BpgError):
        capi.gen_sign_matrix(8, 257, compat=False, seed=1)
    with pytest.raises(capi.BpgError):
        capi.gen_perm_two_sample(6, 3, 21, compat=False, seed=1)
    with pytest.raises(capi.BpgError):
        capi.gen_perm_correlation(4, 25, compat=False, seed=1)


def test_design_detection_matches_oracle():
    rng = np.random.default_rng(4)
    for trial in range(20):
        N = int(rng.integers(8, 60))
        p = int(rng.integers(2, 6))
        X = rng.standard_normal((N, p))
        X[:, 0] = 1
        X[:, 1] = rng.random(N) < 0.4
        if trial % 3 == 0:
            X[:, 2:] = np.round(X[:, 2:])
        C = rng.integers(-1, 2, (3, p)).astype(float)
        C[0] = 0
        C[0, 1] = 1
        for ftest in (False, True):
            a = capi.detect_designs(X, C, ftest)
            b = oracle.detect_designs(X, C, ftest)
            for x, y in zip(a, b):
                assert np.array_equal(x, y)


def test_max_permutations_matches_cli():
    lib = oracle.lib()
    for test, design, N, n2 in [(2, 0, 20, 0), (0, 0, 60, 30), (0, 0, 12, 5), (0, 1, 10, 0), (0, 1, 20, 0),
                                (1, 0, 100, 25), (0, 2, 10, 0)]:
        got = capi.max_permutations(test, design, N, n2)
        ref = lib.orc_max_permutations(test, design, N, n2)
        assert got == pytest.approx(float(ref), rel=1e-12) or (ref >= 9e18 and got >= 9e18)


def test_model_precompute_matches_oracle():
    X = GOLD["tt_design"].astype(np.float32)
    for C, ftest in [(GOLD["tt_contrast"][None, :], False), (np.eye(9)[[1, 2, 4]], True)]:
        a = capi.model_precompute(X, C.astype(np.float32), ftest)
        b = oracle.model_precompute(X, C.astype(np.float32), ftest)
        for x, y in zip(a, b):
            assert np.allclose(x, y, rtol=2e-6, atol=1e-7)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is visible")
def test_create_without_a_device_fails_cleanly():
    import ctypes

    lib = capi.load()
    h = ctypes.c_void_p()
    assert lib.bpg_create(0, ctypes.byref(h)) == -6
    assert lib.bpg_last_error(None) == b"no context"


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is visible")
def test_dropin_reports_initialisation_failure(tmp_path):
    """The drop-in keeps the reference's failure reporting: no device ->
    GetOpenCLInitiated() is false with a message (LIB:1234-1238)."""
    drv = ROOT / "tests" / "cpp" / "dropin_driver"
    (tmp_path / "study.txt").write_text("4 4 4 6 1 1 2 0 8 0 0 1\n")
    for name, n in [("Y.bin", 6 * 64), ("mask.bin", 64), ("X.bin", 6), ("C.bin", 1)]:
        np.ones(n, np.float32).tofile(tmp_path / name)
    res = subprocess.run([str(drv), str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert res.returncode == 4
    assert "Initialization error" in res.stdout


def _small_study(rng):
    from broccoli_amd import synth
    st = synth.study(2, scale=0.1)
    st["inference"] = capi.VOXEL
    return st


def test_validate_study_accepts_the_benchmark_designs():
    from broccoli_amd import synth
    for wl in (1, 2, 3):
        st = synth.study(wl, scale=0.1)
        rc, msg = capi.validate_study(st)
        assert rc == capi.OK, (wl, msg)


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_validate_study_refuses_non_finite_data_inside_the_mask(bad):
    """A NaN statistic's quantised value (int)(T * 1e4) (KM:299) depends on
    the device, so non-finite data inside the mask are refused; outside the
    mask they are never read."""
    st = _small_study(None)
    Y = st["Y"].copy()
    inside = np.argwhere(st["mask"] == 1.0)[5]
    Y[3][tuple(inside)] = bad
    st["Y"] = Y
    rc, msg = capi.validate_study(st)
    assert rc == -1 and "NaN or infinite" in msg
    Y = st["Y"].copy()
    Y[3][tuple(inside)] = 0.5
    outside = np.argwhere(st["mask"] != 1.0)[0]
    Y[3][tuple(outside)] = bad
    st["Y"] = Y
    assert capi.validate_study(st)[0] == capi.OK


def test_validate_study_refuses_bad_designs():
    st = _small_study(None)
    X = st["X"].copy()
    X[:, 3] = X[:, 2]  # rank deficient
    st["X"] = X
    rc, msg = capi.validate_study(st)
    assert rc == -4 and "rank deficient" in msg
    st = _small_study(None)
    st["N"] = st["p"]  # N <= p
    assert capi.validate_study(st)[0] == -1
    st = _small_study(None)
    assert capi.validate_study(st, precision=7)[0] == -1


def test_max_t_contraction_kernels_do_not_spill():
    """The build's kernel report (broccoli_amd/_build/kernel_resources.tsv,
    the compiler's resource remarks): no split-BF16 contraction kernel of the
    max-T paths (output modes 0 and 3, and the persistent form) spills
    vector registers to scratch.  A launch bound that capped a 13-k-step
    kernel at 256 registers once made it spill and run 3x slower; this
    catches that at build time.  (The bitmask and statistics modes are not
    held to it: several of their instantiations spill a little by design,
    e.g. the rolling schedule's bits-only kernels at 8 k-steps, held to 256
    registers for two waves per SIMD.)"""
    report = ROOT / "broccoli_amd" / "_build" / "kernel_resources.tsv"
    if not report.exists():
        pytest.skip("no build report (run the build first)")
    rows = [line.split("\t") for line in report.read_text().splitlines()[1:]]
    names = subprocess.run(["c++filt"], input="\n".join(r[1] for r in rows), capture_output=True,
                           text=True).stdout.splitlines()
    checked, spilling = 0, []
    for r, name in zip(rows, names):
        m = re.search(r"k_contract_bf16(?:_roll)?<(\d+), \d+, \d+, \w+, \w+, (\d+), \w+>", name)
        p = re.search(r"k_contract_bf16_persist<\d+, \d+, \d+, \w+, \w+, (\d+)>", name)
        if m:
            hot = int(m.group(2)) in (0, 3)
        elif p:
            hot = True
        else:
            continue
        if hot:
            checked += 1
            if int(r[5]) > 0:
                spilling.append(name)
    assert checked > 100
    assert not spilling, spilling[:5]
