"""RandomiseGroupLevel end to end on the GPU: NIfTI and text inputs written
here, outputs read back and checked against the oracle on the same inputs
(observed maps against FP64, the permutation values written by
-writepermutationvalues against the FP64 null within the reported-tie rule,
p-values counted exactly against that null, output names and header
fields as the reference writes them: RGL:1588-1703, HelpFunctions.cpp:325-406)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from broccoli_amd import capi, synth
from oracle import oracle
from tests import nifti_util

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
CLI = ROOT / "broccoli_amd" / "RandomiseGroupLevel"


def write_study(tmp, st, perms, gz=True):
    D, H, W = st["mask"].shape
    N = st["N"]
    vol = st["Y"].reshape(N, D, H, W)
    name = tmp / ("vols.nii.gz" if gz else "vols.nii")
    nifti_util.write(name, vol)
    nifti_util.write(tmp / "mask.nii", st["mask"].astype(np.uint8))
    X, C = st["X"], st["C"]
    (tmp / "design.mat").write_text(f"NumRegressors {X.shape[1]}\nNumSubjects {N}\n" +
                                    "\n".join(" ".join(repr(float(v)) for v in row) for row in X) + "\n")
    (tmp / "design.con").write_text(f"NumRegressors {C.shape[1]}\nNumContrasts {C.shape[0]}\n" +
                                    "\n".join(" ".join(repr(float(v)) for v in row) for row in C) + "\n")
    if "perm" in perms:
        rows = perms["perm"].astype(np.int64) + 1
    else:
        rows = perms["sign"]
    (tmp / "perms.txt").write_text("\n".join(" ".join(str(int(v)) if "perm" in perms else f"{v:g}" for v in r)
                                             for r in rows) + "\n")
    return name


def run_cli(*args):
    print("CLI", " ".join(map(str, args)), flush=True)
    r = subprocess.run([str(CLI), *map(str, args)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:])
    return r


@pytest.mark.parametrize("case", ["ttest_voxel", "groupmean_voxel", "ftest_cluster"])
def test_cli_end_to_end_matches_oracle(tmp_path, case):
    if case == "ttest_voxel":
        st = synth.study(2, scale=0.15)
        st["inference"] = capi.VOXEL
        K, extra = 64, ["-inferencemode", "0"]
    elif case == "groupmean_voxel":
        st = synth.study(1, scale=0.2)
        st["inference"] = capi.VOXEL
        K, extra = 128, ["-inferencemode", "0", "-groupmean"]
    else:
        st = synth.study(3, scale=0.2)
        st["inference"] = capi.CLUSTER_EXTENT
        st["cdt"] = 3.0
        K, extra = 40, ["-inferencemode", "1", "-teststatistics", "1", "-cdt", "3"]
    perms = synth.permutations(st, K, seed=31)
    vols = write_study(tmp_path, st, perms, gz=(case != "groupmean_voxel"))
    args = [vols, "-mask", tmp_path / "mask.nii", "-permutations", K, "-permutationfile", tmp_path / "perms.txt",
            "-writepermutationvalues", tmp_path / "null.txt", "-writepermutations", tmp_path / "used.txt", *extra]
    if case != "groupmean_voxel":
        args[1:1] = ["-design", tmp_path / "design.mat", "-contrasts", tmp_path / "design.con"]
    r = run_cli(*args)
    assert r.returncode == 0, r.stdout

    ftest = case == "ftest_cluster"
    stat_name = "vols_perm_fvalues" if ftest else "vols_perm_tvalues"
    ext = ".nii" if case == "groupmean_voxel" else ".nii.gz"
    maps, hdr = nifti_util.read(tmp_path / (stat_name + ext))
    pvals, _ = nifti_util.read(tmp_path / ("vols_perm_pvalues" + ext))
    assert hdr["datatype"] == 16 and hdr["magic"] == b"n+1\0"
    assert hdr["cal_max"] == pytest.approx(float(maps.max()))

    maps64, P64, md64 = oracle.randomise(st, perms, K, precision=oracle.F64)
    got = maps.reshape(-1)
    err = np.abs(got - maps64[0]) / (1.0 + np.abs(maps64[0]))
    assert err.max() <= 1e-5
    # the null written as text (fixed, 6 decimals), one value per line: the
    # same numbers as the C ABI gives for these permutations (the C-ABI
    # null is checked against the oracle in test_gpu_parity)
    null = np.array([float(x) for x in (tmp_path / "null1.txt").read_text().split()])
    assert null.shape == (K,)
    with capi.Context(0) as ctx, capi.Plan(ctx, st, "split_bf16") as plan:
        md_c, _ = plan.permutation_phase(0, **perms)
    assert np.array_equal(null.astype(np.float32), md_c)
    if case != "ftest_cluster":
        assert np.all(np.abs(np.round(null * 1e4) - np.round(md64[0].astype(np.float64) * 1e4)) <= 1)
    else:
        # cluster sizes equal the FP64 path's except for permutations whose
        # map has an FP64 statistic within the error band of the threshold
        ref = oracle.perm_stats(st, perms, K)[:, st["mask"].ravel() == 1.0]
        near = (np.abs(ref - st["cdt"]) <= 1e-4 * (1.0 + abs(st["cdt"]))).any(axis=1)
        diff = null.astype(np.float32) != md64[0]
        assert not np.any(diff & ~near), np.flatnonzero(diff & ~near)
    # p-values: exact counts against the null the run produced
    mask = st["mask"].ravel() == 1.0
    D, H, W = st["mask"].shape
    Pref = np.zeros(W * H * D, np.float32)
    if case == "ftest_cluster":
        lab, sz, _ = oracle.ccl(got.reshape(st["mask"].shape), st["mask"], st["cdt"], 26)
        oracle.lib().orc_pvalues(1, W, H, D, oracle.P(got), oracle.P(st["mask"]), oracle.P(lab), oracle.P(sz),
                                 st["cdt"], K, oracle.P(null.astype(np.float32)), oracle.P(Pref))
    else:
        oracle.lib().orc_pvalues(0, W, H, D, oracle.P(got), oracle.P(st["mask"]), None, None, 0.0, K,
                                 oracle.P(null.astype(np.float32)), oracle.P(Pref))
    assert np.array_equal(pvals.reshape(-1)[mask], Pref[mask])
    # the permutations used are the file's (1-based, or the signs)
    used = np.array([[float(x) for x in line.split()] for line in (tmp_path / "used.txt").read_text().splitlines()])
    expect = perms["sign"] if case == "groupmean_voxel" else perms["perm"] + 1
    assert np.array_equal(used, expect.astype(np.float64))


def test_cli_generated_permutations_and_output_name(tmp_path):
    """Without a permutation file the library draws them (compat stream) and
    -output names the results (suffix inserted before the first dot)."""
    st = synth.study(2, scale=0.12)
    st["inference"] = capi.VOXEL
    perms = synth.permutations(st, 16, seed=1)
    vols = write_study(tmp_path, st, perms)
    out = tmp_path / "res"
    r = run_cli(vols, "-design", tmp_path / "design.mat", "-contrasts", tmp_path / "design.con", "-mask",
                tmp_path / "mask.nii", "-permutations", 50, "-inferencemode", "0", "-output", out,
                "-writepermutations", tmp_path / "used.txt")
    assert r.returncode == 0
    assert "Two sample t-test design detected for t-contrast 1, 30 subjects in group 1 and 30 subjects in group 2" \
        in r.stdout
    assert (tmp_path / "res_perm_tvalues.nii").exists() and (tmp_path / "res_perm_pvalues.nii").exists()
    used = np.array([[int(float(x)) for x in line.split()] for line in (tmp_path / "used.txt").read_text().splitlines()])
    ref = capi.gen_perm_two_sample(st["N"], 30, 50, compat=True, seed=1) + 1
    assert np.array_equal(used, ref)
