"""Writes tests/golden/reference_vectors.npz.

Run in the development container, where the reference tree is mounted and
oracle/_ref/libbroccoli_ref.so (the reference's own broccoli_lib.cpp compiled
unmodified, see oracle/Makefile) can be built.  The fixture holds:

- the inputs the reference's MATLAB test holds
  (code/Matlab_Wrapper/test_GLMTTest_Second_Level_Permutation.m:125-175):
  the 10x25 design (first 9 columns used), the contrast and the 6x10
  permutation matrix (1-based there, stored 0-based here);
- outputs of the reference's own routines on given inputs:
  GenerateSignMatrixSecondLevel, GeneratePermutationMatrixSecondLevel
  {TwoSample,Correlation} and GeneratePermutationMatrixFirstLevel (after
  srand(1)), the host Clusterize (26-connected
  BFS, labels 1..n), and pinv.

Nothing here runs on a GPU; the tests only read the .npz.
"""
from __future__ import annotations

import ctypes
import re
import sys
from ctypes import c_float, c_int, c_long, c_void_p
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import oracle  # noqa: E402

MFILE = Path("/root/reference/code/Matlab_Wrapper/test_GLMTTest_Second_Level_Permutation.m")
OUT = Path(__file__).resolve().with_name("reference_vectors.npz")


def held_inputs():
    lines = MFILE.read_text().splitlines()
    # X_GLM = [ ... ]; at lines 125-137 (1-based)
    rows = []
    for ln in lines[126:136]:
        vals = [float(v) for v in ln.split()]
        if vals:
            rows.append(vals)
    X = np.array(rows)
    assert X.shape == (10, 25), X.shape
    con = [float(v) for v in re.findall(r"-?\d+\.\d+", lines[150])]
    assert len(con) == 9, con
    perms = []
    for ln in lines[169:175]:
        m = re.search(r"=\s*\[([^\]]*)\]", ln)
        perms.append([int(v) - 1 for v in m.group(1).split()])
    P = np.array(perms, np.uint16)
    assert P.shape == (6, 10)
    return X[:, :9], np.array(con), P


def main():
    oracle.build(ref=True)
    ref = oracle.ref_lib()
    assert ref is not None, "oracle/_ref/libbroccoli_ref.so was not built"
    P = oracle.P
    out = {}
    X9, con, tperm = held_inputs()
    out["tt_design"] = X9
    out["tt_contrast"] = con
    out["tt_permutations"] = tperm
    # pinv (LIB:16561-16586) of the held design
    Xf = np.asfortranarray(X9).ravel(order="F").copy()
    pv = np.zeros(9 * 10)
    ref.ref_pinv(10, 9, P(Xf), P(pv))
    out["tt_pinv"] = pv.reshape(10, 9).T  # result is 9 x 10 column-major
    # generators after srand(1)
    for N, K in [(6, 64), (8, 100), (20, 500)]:
        a = np.zeros(N * K, np.float32)
        oracle.lib().orc_srand(1)
        ref.ref_gen_sign_matrix(N, c_long(K), P(a))
        out[f"sign_N{N}_K{K}"] = a.reshape(K, N)
    for N, n1, K in [(10, 4, 100), (12, 6, 400), (60, 30, 1000)]:
        a = np.zeros(N * K, np.uint16)
        oracle.lib().orc_srand(1)
        ref.ref_gen_perm_two_sample(N, n1, N - n1, c_long(K), P(a))
        out[f"twosample_N{N}_n{n1}_K{K}"] = a.reshape(K, N)
    for N, K in [(6, 200), (10, 300), (25, 500)]:
        a = np.zeros(N * K, np.uint16)
        oracle.lib().orc_srand(1)
        ref.ref_gen_perm_correlation(N, c_long(K), P(a))
        out[f"correlation_N{N}_K{K}"] = a.reshape(K, N)
    # first-level time-point permutations (LIB:16877-16922)
    for T, K in [(12, 100), (40, 300), (160, 200)]:
        a = np.zeros(T * K, np.uint16)
        oracle.lib().orc_srand(1)
        ref.ref_gen_perm_first_level(T, c_long(K), P(a))
        out[f"firstlevel_T{T}_K{K}"] = a.reshape(K, T)
    # host Clusterize (LIB:18163-18297) on smoothed noise volumes
    rng = np.random.default_rng(20260009)
    from scipy.ndimage import gaussian_filter

    for i, (W, H, D, thr) in enumerate([(12, 10, 8, 0.15), (20, 18, 16, 0.2), (31, 27, 23, 0.25)]):
        vol = gaussian_filter(rng.standard_normal((D, H, W)), 1.0).astype(np.float32)
        mask = (rng.random((D, H, W)) < 0.9).astype(np.float32)
        lab = np.zeros(W * H * D, np.int32)
        mx = ctypes.c_int()
        nclu = ctypes.c_int()
        ref.ref_clusterize(W, H, D, P(vol), c_float(thr), P(mask), P(lab), ctypes.byref(mx), ctypes.byref(nclu))
        out[f"clu{i}_data"] = vol
        out[f"clu{i}_mask"] = mask
        out[f"clu{i}_thr"] = np.float32(thr)
        out[f"clu{i}_labels"] = lab.reshape(D, H, W)
        out[f"clu{i}_max"] = np.int32(mx.value)
        out[f"clu{i}_n"] = np.int32(nclu.value)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
