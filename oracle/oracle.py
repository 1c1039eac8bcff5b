"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() (to
check the GPU results) and bench.py's cpu_baseline leg.  The product package never imports
it.  The functions restate the reference path (see bpg_oracle.c for the
reference lines of each).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_float, c_int, c_long, c_uint, c_void_p
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libbroccoli_ref.so"

REF32, F64 = 0, 1


def build(ref: bool = False) -> None:
    """Compiles liboracle.so (and the reference harness when asked and the
    reference tree is present)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "all"], check=True)
    if ref and Path("/root/reference/code/BROCCOLI_LIB/broccoli_lib.cpp").exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)


class Study(ctypes.Structure):
    _fields_ = [("test", c_int), ("inference", c_int), ("W", c_int), ("H", c_int), ("D", c_int),
                ("N", c_int), ("p", c_int), ("nc", c_int), ("Y", c_void_p), ("mask", c_void_p),
                ("X", c_void_p), ("xtxxt", c_void_p), ("C", c_void_p), ("ctxtxc", c_void_p),
                ("K", c_void_p), ("perm", c_void_p), ("sign", c_void_p), ("cdt", c_float),
                ("connectivity", c_int), ("nthreads", c_int)]


_lib = None
_ref = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = ctypes.CDLL(str(LIB))
        _lib.orc_srand.argtypes = [c_uint]
        _lib.orc_max_quantized.restype = c_float
        _lib.orc_max_quantized.argtypes = [c_int, c_int, c_int, c_void_p, c_void_p]
        _lib.orc_quantize.argtypes = [c_float]
        _lib.orc_ccl.restype = c_uint
        _lib.orc_ccl.argtypes = [c_int, c_int, c_int, c_void_p, c_void_p, c_float, c_int, c_void_p, c_void_p]
        _lib.orc_tfce_max.restype = c_float
        _lib.orc_tfce_max.argtypes = [c_int, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p]
        _lib.orc_cluster_mass.restype = c_uint
        _lib.orc_cluster_mass.argtypes = [c_int, c_int, c_int, c_void_p, c_void_p, c_float, c_int, c_void_p,
                                          c_void_p]
        _lib.orc_stirling_factorial.restype = c_double
        _lib.orc_max_permutations.restype = c_long
        _lib.orc_pvalues.argtypes = [c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_float, c_long, c_void_p, c_void_p]
    return _lib


def ref_lib():
    """The reference's own CPU routines (oracle/_ref), or None when not built."""
    global _ref
    if _ref is None and REF_LIB.exists():
        _ref = ctypes.CDLL(str(REF_LIB))
        _ref.ref_clusterize.argtypes = [c_int, c_int, c_int, c_void_p, c_float, c_void_p, c_void_p,
                                        c_void_p, c_void_p]
    return _ref


def P(a):
    return None if a is None else a.ctypes.data_as(c_void_p)


# ------------------------------------------------------------- generators

def gen_sign_matrix(N, K, seed=1):
    out = np.zeros((K, N), np.float32)
    lib().orc_srand(seed)
    lib().orc_gen_sign_matrix(N, c_long(K), P(out))
    return out


def gen_perm_two_sample(N, n1, K, seed=1):
    out = np.zeros((K, N), np.uint16)
    lib().orc_srand(seed)
    lib().orc_gen_perm_two_sample(N, n1, N - n1, c_long(K), P(out))
    return out


def gen_perm_correlation(N, K, seed=1):
    out = np.zeros((K, N), np.uint16)
    lib().orc_srand(seed)
    lib().orc_gen_perm_correlation(N, c_long(K), P(out))
    return out


def detect_designs(X, C, ftest):
    N, p = X.shape
    nc = C.shape[0]
    Xf = np.asfortranarray(X, np.float32).ravel(order="F").copy()
    Cc = np.ascontiguousarray(C, np.float32).ravel()
    d, n1, n2 = (np.zeros(nc, np.int32) for _ in range(3))
    lib().orc_detect_designs(N, p, nc, P(Xf), P(Cc), int(ftest), P(d), P(n1), P(n2))
    g = 1 if ftest else nc
    return d[:g], n1[:g], n2[:g]


def model_precompute(X, C, ftest):
    N, p = X.shape
    nc = C.shape[0]
    Xf = np.asfortranarray(X, np.float32).ravel(order="F").copy()
    Cc = np.ascontiguousarray(C, np.float32).ravel()
    xtxxt = np.zeros(N * p, np.float32)
    ctxtxc = np.zeros(nc * nc if ftest else nc, np.float32)
    rc = lib().orc_model_precompute(N, p, nc, int(ftest), P(Xf), P(Cc), P(xtxxt), P(ctxtxc))
    assert rc == 0
    return xtxxt, ctxtxc


def pinv64(X):
    N, p = X.shape
    Xf = np.asfortranarray(X, np.float64).ravel(order="F").copy()
    out = np.zeros(N * p)
    lib().orc_pinv64(N, p, P(Xf), P(out))
    return out.reshape(p, N)  # out[s + r*N] -> [r][s]


def nuisance_projector64(X, Crows):
    N, p = X.shape
    r = Crows.shape[0]
    Xf = np.asfortranarray(X, np.float64).ravel(order="F").copy()
    Cr = np.ascontiguousarray(Crows, np.float64).ravel()
    R = np.zeros(N * N)
    lib().orc_nuisance_projector64(N, p, P(Xf), r, P(Cr), P(R))
    return R.reshape(N, N).T  # column-major -> [i][j]


def ccl(stat, mask, thr, connectivity=26):
    D, H, W = mask.shape
    lab = np.zeros(mask.size, np.uint32)
    sz = np.zeros(mask.size, np.uint32)
    st = np.ascontiguousarray(stat, np.float32)
    mk = np.ascontiguousarray(mask, np.float32)
    best = lib().orc_ccl(W, H, D, P(st), P(mk), c_float(thr), connectivity, P(lab), P(sz))
    return lab, sz, int(best)


def cluster_mass(stat, mask, thr, labels):
    """Per-label mass sum of (uint)(T*1e4) over supra-threshold voxels and
    the largest one (KC:432-458)."""
    D, H, W = mask.shape
    ms = np.zeros(mask.size, np.uint32)
    st = np.ascontiguousarray(stat, np.float32)
    mk = np.ascontiguousarray(mask, np.float32)
    best = lib().orc_cluster_mass(W, H, D, P(st), P(mk), c_float(thr), 26, P(labels), P(ms))
    return ms, int(best)


def tfce_max(stat, mask, connectivity=26, want_values=False):
    """Quantised max TFCE of one map (LIB:18464-18508, KC:482-510), and the
    per-voxel sums when asked."""
    D, H, W = mask.shape
    st = np.ascontiguousarray(stat, np.float32)
    mk = np.ascontiguousarray(mask, np.float32)
    vals = np.zeros(mask.size, np.float32) if want_values else None
    m = lib().orc_tfce_max(W, H, D, P(st), P(mk), connectivity, P(vals))
    return (m, vals) if want_values else m


def max_quantized(stat, mask):
    D, H, W = mask.shape
    return lib().orc_max_quantized(W, H, D, P(np.ascontiguousarray(stat, np.float32)),
                                   P(np.ascontiguousarray(mask, np.float32)))


# ------------------------------------------------------------- studies

class _Bound:
    """Keeps the arrays of an orc_study alive."""

    def __init__(self, st: dict, perms: dict, K, nthreads=0):
        N, p, nc = st["N"], st["p"], st["nc"]
        ftest = st["test"] == 1
        self.Y = np.ascontiguousarray(st["Y"], np.float32)
        self.mask = np.ascontiguousarray(st["mask"], np.float32)
        self.X = np.asfortranarray(st["X"], np.float32).ravel(order="F").copy()
        self.C = np.ascontiguousarray(st["C"], np.float32).ravel()
        if "xtxxt" in st:
            self.xtxxt, self.ctxtxc = st["xtxxt"], st["ctxtxc"]
        else:
            self.xtxxt, self.ctxtxc = model_precompute(st["X"], st["C"], ftest)
        nmaps = nc if st["test"] == 0 else 1
        Ks = K if isinstance(K, (list, tuple, np.ndarray)) else [K] * nmaps
        self.K = np.array(Ks, np.int64)
        self.perm_arrs = []
        self.sign = None
        if st["test"] == 2:
            self.sign = np.ascontiguousarray(perms["sign"], np.float32)
            self.permp = None
        else:
            pl = perms["perm"] if isinstance(perms["perm"], list) else [perms["perm"]] * nmaps
            self.perm_arrs = [np.ascontiguousarray(a, np.uint16) for a in pl]
            self.permp = (c_void_p * nmaps)(*[a.ctypes.data for a in self.perm_arrs])
        s = Study()
        s.test, s.inference = st["test"], st.get("inference", 0)
        s.W, s.H, s.D, s.N, s.p, s.nc = st["W"], st["H"], st["D"], N, p, nc
        s.Y, s.mask, s.X, s.xtxxt = P(self.Y), P(self.mask), P(self.X), P(self.xtxxt)
        s.C, s.ctxtxc, s.K = P(self.C), P(self.ctxtxc), P(self.K)
        s.perm = ctypes.cast(self.permp, c_void_p) if self.permp is not None else None
        s.sign = P(self.sign)
        s.cdt = st.get("cdt", 0.0)
        s.connectivity = st.get("connectivity", 26)
        s.nthreads = nthreads
        self.s = s
        self.nmaps = nmaps


def perm_stats(st: dict, perms: dict, K: int, c: int = 0, precision: int = F64) -> np.ndarray:
    """Statistic maps of the first K permutations of map c: (K, D*H*W)."""
    b = _Bound(st, perms, K)
    out = np.zeros((K, st["W"] * st["H"] * st["D"]))
    rc = lib().orc_perm_stats(ctypes.byref(b.s), c, precision, c_long(K), P(out))
    assert rc == 0
    return out


def quantised_max(m) -> np.ndarray:
    """The null value the reference stores for a maximum statistic m
    (KM:299 + LIB:10988), vectorised: (float)(int)(float(m) * 10000.0f) /
    10000.0f in FP32 arithmetic, the int conversion saturating as the GPU's
    v_cvt_i32_f32 does (the library's choice for |t| > 214748.3647, where
    C leaves the conversion undefined).  Monotone in m."""
    mf = np.asarray(m, np.float64).astype(np.float32)
    prod = (mf * np.float32(10000.0)).astype(np.float64)
    q = np.clip(np.trunc(prod), -2.0 ** 31, 2.0 ** 31 - 1)
    return (q.astype(np.float32) / np.float32(10000.0)).astype(np.float32)


def perm_max64(st: dict, perms: dict, K: int, c: int = 0, nthreads: int = 0) -> np.ndarray:
    """FP64 maximum statistic (unquantised) of the first K permutations of
    map c over the in-mask voxels: (K,)."""
    b = _Bound(st, perms, K, nthreads)
    out = np.zeros(K)
    rc = lib().orc_perm_max64(ctypes.byref(b.s), c, c_long(K), P(out))
    assert rc == 0
    return out


def randomise(st: dict, perms: dict, K, precision: int = F64):
    """The whole path: (maps, pvalues, maxdist list)."""
    b = _Bound(st, perms, K)
    G = st["W"] * st["H"] * st["D"]
    maps = np.zeros(b.nmaps * G, np.float32)
    pv = np.zeros(b.nmaps * G, np.float32)
    md = [np.zeros(int(k), np.float32) for k in b.K]
    mdp = (c_void_p * b.nmaps)(*[m.ctypes.data for m in md])
    rc = lib().orc_randomise(ctypes.byref(b.s), precision, P(maps), P(pv), ctypes.cast(mdp, c_void_p))
    assert rc == 0
    return maps.reshape(b.nmaps, G), pv.reshape(b.nmaps, G), md


def permutation_phase32(st: dict, perms: dict, K: int, Ystar: np.ndarray, nthreads: int = 0):
    """Times-able CPU restatement of the reference's K loop (voxel mode) on
    already-transformed data Ystar (grid layout)."""
    st2 = dict(st)
    st2["Y"] = Ystar
    b = _Bound(st2, perms, K, nthreads)
    md = np.zeros(K, np.float32)
    rc = lib().orc_permutation_phase32(ctypes.byref(b.s), 0, c_long(K), P(md))
    assert rc == 0
    return md


def transformed32(st: dict, c: int = 0) -> np.ndarray:
    """The FP32 TransformData output of map c (grid layout, float32)."""
    N = st["N"]
    ftest = st["test"] == 1
    p, nc = st["p"], st["nc"]
    Y = np.ascontiguousarray(st["Y"], np.float32).copy()
    transform = (st["test"] == 0 and p > 1) or (ftest and p > nc)
    if transform:
        rows = st["C"][c:c + 1] if st["test"] == 0 else st["C"]
        R = nuisance_projector64(st["X"].astype(np.float64), rows.astype(np.float64))
        Rf = np.ascontiguousarray(R.T, np.float32)  # column-major
        lib().orc_transform32(st["W"], st["H"], st["D"], N, P(Rf), P(np.ascontiguousarray(st["mask"], np.float32)),
                              P(Y))
    return Y


# ------------------------------------------------------------- first level
# Volumes are numpy arrays of shape (T, D, H, W) (index x + y*W + z*W*H +
# t*W*H*D), masks (D, H, W), designs (T, p), contrasts (nc, p), AR fields
# (4, D, H, W), permutations (K, T).

def _f64(a):
    return np.ascontiguousarray(a, np.float64)


def _xcol(X):
    return np.asfortranarray(X, np.float64).ravel(order="F").copy()


def fl_filters(fwhm, vx, vy, vz, size=9):
    f = [np.zeros(size, np.float32) for _ in range(3)]
    lib().orc_fl_filters(size, c_float(fwhm), c_float(vx), c_float(vy), c_float(vz), P(f[0]), P(f[1]), P(f[2]))
    return f


def fl_smooth(vol, filters, certainty=None, scert=None):
    D, H, W = vol.shape
    out = np.zeros(vol.shape)
    fx, fy, fz = filters
    cert = None if certainty is None else np.ascontiguousarray(certainty, np.float32)
    sc = None if scert is None else _f64(scert)
    lib().orc_fl_smooth(W, H, D, P(_f64(vol)), P(cert), P(sc), P(fx), P(fy), P(fz), P(out))
    return out


def fl_regress(Y, mask, X):
    T, D, H, W = Y.shape
    out = np.zeros(Y.shape)
    lib().orc_fl_regress(W, H, D, T, X.shape[1], P(_f64(Y)), P(np.ascontiguousarray(mask, np.float32)),
                         P(_xcol(X)), P(out))
    return out


def fl_ar4(data, mask, invalid=0):
    T, D, H, W = data.shape
    out = np.zeros((4, D, H, W))
    lib().orc_fl_ar4(W, H, D, T, P(_f64(data)), P(np.ascontiguousarray(mask, np.float32)), invalid, P(out))
    return out


def fl_whiten(data, ar, mask):
    T, D, H, W = data.shape
    out = np.zeros(data.shape)
    lib().orc_fl_whiten(W, H, D, T, P(_f64(data)), P(_f64(ar)), P(np.ascontiguousarray(mask, np.float32)), P(out))
    return out


def fl_color(white, ar, mask, perm):
    T, D, H, W = white.shape
    out = np.zeros(white.shape)
    lib().orc_fl_color(W, H, D, T, P(_f64(white)), P(_f64(ar)), P(np.ascontiguousarray(mask, np.float32)),
                       P(np.ascontiguousarray(perm, np.uint16)), P(out))
    return out


def fl_prepare(Y, mask, X, fwhm=8.0, voxel=(3.0, 3.0, 3.0)):
    """(whitened volumes, AR estimates) before the permutations."""
    T, D, H, W = Y.shape
    wh = np.zeros(Y.shape)
    ar = np.zeros((4, D, H, W))
    Xf = np.asfortranarray(X, np.float32).ravel(order="F").copy()
    lib().orc_fl_prepare(W, H, D, T, X.shape[1], P(np.ascontiguousarray(Y, np.float32)),
                         P(np.ascontiguousarray(mask, np.float32)), P(Xf), c_float(fwhm), c_float(voxel[0]),
                         c_float(voxel[1]), c_float(voxel[2]), P(wh), P(ar))
    return wh, ar


def fl_stat(test, x, mask, X, C, contrast=0):
    T, D, H, W = x.shape
    C = np.atleast_2d(C)
    out = np.zeros((D, H, W))
    lib().orc_fl_stat(test, W, H, D, T, X.shape[1], C.shape[0], P(_f64(x)), P(np.ascontiguousarray(mask, np.float32)),
                      P(_xcol(X)), P(_f64(C).ravel()), contrast, P(out))
    return out


def gen_perm_first_level(T, K, seed=1):
    out = np.zeros((K, T), np.uint16)
    lib().orc_srand(seed)
    rc = lib().orc_gen_perm_first_level(T, c_long(K), P(out))
    assert rc == 0
    return out


def fl_null(test, inference, white, ar, mask, X, C, perm, contrast=0, cdt=0.0, connectivity=26,
            want_stats=False):
    """Null distribution of one map (K values) and, when asked, the
    statistic maps (K, D, H, W)."""
    T, D, H, W = white.shape
    C = np.atleast_2d(C)
    K = perm.shape[0]
    md = np.zeros(K, np.float32)
    stats = np.zeros((K, D, H, W)) if want_stats else None
    Xf = np.asfortranarray(X, np.float32).ravel(order="F").copy()
    Cf = np.ascontiguousarray(C, np.float32).ravel()
    rc = lib().orc_fl_null(test, inference, W, H, D, T, X.shape[1], C.shape[0], P(_f64(white)), P(_f64(ar)),
                           P(np.ascontiguousarray(mask, np.float32)), P(Xf), P(Cf), contrast, c_long(K),
                           P(np.ascontiguousarray(perm, np.uint16)), c_float(cdt), connectivity, P(md), P(stats))
    assert rc == 0
    return (md, stats) if want_stats else md


def ref_gen_perm_first_level(T, K, seed=1):
    """The reference's own generator (oracle/_ref), or None when not built."""
    r = ref_lib()
    if r is None:
        return None
    out = np.zeros((K, T), np.uint16)
    lib().orc_srand(seed)  # one process-global rand() stream
    r.ref_gen_perm_first_level(T, c_long(K), P(out))
    return out
