"""CPU tests of the RandomiseGroupLevel front end: usage, option and file
checks with the reference's messages (code/Bash_Wrapper/
RandomiseGroupLevel.cpp:200-741), and the NIfTI reader (dimensions and
volume counts of .nii / .nii.gz, several voxel types, both byte orders),
all of which run before a device is needed."""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from tests import nifti_util

ROOT = Path(__file__).resolve().parents[1]
CLI = ROOT / "broccoli_amd" / "RandomiseGroupLevel"
NO_GPU = not os.path.exists("/dev/kfd")


def run(*args, cwd=None):
    r = subprocess.run([str(CLI), *map(str, args)], capture_output=True, text=True, timeout=120, cwd=cwd)
    return r.returncode, r.stdout


def design_files(tmp, N, X, C):
    d = tmp / "design.mat"
    d.write_text(f"NumRegressors {X.shape[1]}\nNumSubjects {N}\n\n" +
                 "\n".join(" ".join(f"{v:.6f}" for v in row) for row in X) + "\n")
    c = tmp / "design.con"
    c.write_text(f"NumRegressors {C.shape[1]}\nNumContrasts {C.shape[0]}\n\n" +
                 "\n".join(" ".join(f"{v:g}" for v in row) for row in C) + "\n")
    return d, c


@pytest.fixture
def small(tmp_path):
    rng = np.random.default_rng(0)
    N, D, H, W = 8, 3, 4, 5
    vol = rng.standard_normal((N, D, H, W)).astype(np.float32)
    nifti_util.write(tmp_path / "vols.nii.gz", vol)
    X = np.c_[np.ones(N), rng.standard_normal(N)]
    C = np.array([[0.0, 1.0]])
    d, c = design_files(tmp_path, N, X, C)
    return tmp_path, N, (W, H, D), d, c


def test_usage_without_arguments():
    rc, out = run()
    assert rc == 0 and "RandomiseGroupLevel volumes.nii -design design.mat" in out


def test_extension_and_missing_file(tmp_path):
    rc, out = run(tmp_path / "a.img")
    assert rc == 1 and "File extension is not .nii or .nii.gz, .img is not allowed!" in out
    rc, out = run(tmp_path / "missing.nii")
    assert rc == 1 and "Could not open file" in out


def test_option_checks(small):
    tmp, N, _, d, c = small
    v = tmp / "vols.nii.gz"
    assert run(v, "-design", d)[1].startswith("No contrasts file detected")
    assert "No design file detected" in run(v, "-contrasts", c)[1]
    assert "Number of permutations must be an integer! You provided 12x" in run(v, "-permutations", "12x")[1]
    assert "Number of permutations must be > 0!" in run(v, "-permutations", "0")[1]
    assert "Test statistics must be 0 or 1!" in run(v, "-teststatistics", "2")[1]
    assert "TFCE is currently turned off!" in run(v, "-inferencemode", "3")[1]
    assert "Significance level must be between" in run(v, "-significance", "1.5")[1]
    assert "Unrecognized option! -bogus" in run(v, "-bogus")[1]
    assert "Unable to read value after -cdt" in run(v, "-cdt")[1]
    assert "Cannot use F-test for group mean" in run(v, "-groupmean", "-teststatistics", "1")[1]


def test_design_file_checks(small):
    tmp, N, _, d, c = small
    v = tmp / "vols.nii.gz"
    bad = tmp / "bad.mat"
    bad.write_text("NumRegressor 2\nNumSubjects 8\n")
    out = run(v, "-design", bad, "-contrasts", c)[1]
    assert "First element of the design file should be the string 'NumRegressors', but it is NumRegressor" in out
    bad.write_text("NumRegressors 2\nNumSubjects 9\n")
    out = run(v, "-design", bad, "-contrasts", c)[1]
    assert "Input data contains 8 volumes, while the design file says 9 subjects" in out
    bad.write_text("NumRegressors 26\nNumSubjects 8\n")
    assert "Number of regressors must be <= 25" in run(v, "-design", bad, "-contrasts", c)[1]
    badc = tmp / "bad.con"
    badc.write_text("NumRegressors 3\nNumContrasts 1\n0 1 0\n")
    assert "while contrast file says there are 3 regressors" in run(v, "-design", d, "-contrasts", badc)[1]
    short = tmp / "short.mat"
    short.write_text("NumRegressors 2\nNumSubjects 8\n1 0\n1 1\n")
    assert "Could not read all values of the design file" in run(v, "-design", short, "-contrasts", c)[1]
    rank = tmp / "rank.con"
    rank.write_text("NumRegressors 2\nNumContrasts 2\n0 1\n0 2\n")
    out = run(v, "-design", d, "-contrasts", rank, "-teststatistics", "1")[1]
    assert "Contrast matrix does not have full rank" in out


@pytest.mark.parametrize("dtype,name,big", [(np.float32, "v.nii", False), (np.int16, "v.nii.gz", False),
                                            (np.uint8, "v.nii", False), (np.float32, "v.nii.gz", True),
                                            (np.uint16, "v.nii", True), (np.float64, "v.nii", False)])
def test_reader_dimensions_types_byte_orders(tmp_path, dtype, name, big):
    N, D, H, W = 6, 2, 3, 7
    vol = (np.arange(N * D * H * W) % 50).reshape(N, D, H, W).astype(dtype)
    nifti_util.write(tmp_path / name, vol, big_endian=big)
    mask = np.ones((D + 1, H, W), np.uint8)
    nifti_util.write(tmp_path / "m.nii", mask)
    rc, out = run(tmp_path / name, "-groupmean", "-mask", tmp_path / "m.nii")
    assert rc == 1
    assert f"Input data has the dimensions {W} x {H} {D}, while the mask volume has the dimensions {W} x {H} x {D + 1}" in out
    bad = tmp_path / "d.mat"
    bad.write_text(f"NumRegressors 1\nNumSubjects {N + 1}\n")
    c = tmp_path / "d.con"
    c.write_text("NumRegressors 1\nNumContrasts 1\n1\n")
    out = run(tmp_path / name, "-design", bad, "-contrasts", c)[1]
    assert f"Input data contains {N} volumes, while the design file says {N + 1} subjects" in out


def test_single_volume_and_unknown_type(tmp_path):
    nifti_util.write(tmp_path / "one.nii", np.zeros((1, 2, 2, 2), np.float32))
    assert "Input data is a single volume, nothing to permute!" in run(tmp_path / "one.nii", "-groupmean")[1]
    raw = bytearray(open(tmp_path / "one.nii", "rb").read())
    raw[70:72] = (128).to_bytes(2, "little")  # RGB24
    (tmp_path / "rgb.nii").write_bytes(bytes(raw))
    assert "Could not open volumes!" in run(tmp_path / "rgb.nii", "-groupmean")[1]


@pytest.mark.skipif(not NO_GPU, reason="a GPU is visible")
def test_reaches_the_device_after_all_host_checks(small):
    """With valid inputs the front end gets as far as the library's
    initialisation, which fails cleanly without a device (LIB:1234-1238)."""
    tmp, N, (W, H, D), d, c = small
    rc, out = run(tmp / "vols.nii.gz", "-design", d, "-contrasts", c, "-permutations", "20")
    assert f"Data size: {W} x {H} x {D} x {N}" in out
    assert "Correlation design detected for t-contrast 1" in out
    # the CLI's factorial is Stirling's approximation (RGL:38-42): 8! -> 39890
    assert "Max number of permutations for contrast 1 is 39890" in out
    assert "Initialization error" in out and rc == 1
