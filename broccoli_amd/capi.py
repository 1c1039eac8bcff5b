"""ctypes binding of the C ABI in include/bpg.h (libbpg.so).

This is the host-side mirror used by the tests, smoke() and bench.py.  It
loads the in-tree library and fails loudly when it is missing: there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_float, c_int, c_long, c_uint, c_ulonglong, c_void_p
from pathlib import Path

import numpy as np

import os

# BPG_LIB selects another build of the same library (e.g. a tuning variant).
LIB_PATH = Path(os.environ.get("BPG_LIB", Path(__file__).resolve().with_name("libbpg.so")))

TTEST, FTEST, GROUP_MEAN = 0, 1, 2
VOXEL, CLUSTER_EXTENT, CLUSTER_MASS, TFCE = 0, 1, 2, 3
TWOSAMPLE, CORRELATION, MEAN_DESIGN = 0, 1, 2
PREC_FP32, PREC_SPLIT_BF16, PREC_FP64, PREC_SPLIT_F16, PREC_FP32_VALU, PREC_SPLIT_FP8 = 0, 1, 2, 3, 4, 5
FL_PRECISIONS = {"split_f16": PREC_SPLIT_F16, "split_bf16": PREC_SPLIT_BF16, "fp32": PREC_FP32, "fp64": PREC_FP64}
PRECISIONS = {"fp32": PREC_FP32, "split_bf16": PREC_SPLIT_BF16, "fp64": PREC_FP64, "fp32_valu": PREC_FP32_VALU,
              "split_fp8": PREC_SPLIT_FP8}

OK = 0
ERRORS = {-1: "BPG_ERR_ARG", -2: "BPG_ERR_HIP", -3: "BPG_ERR_NOMEM", -4: "BPG_ERR_DESIGN",
          -5: "BPG_ERR_UNSUPPORTED", -6: "BPG_ERR_NODEVICE"}

# every symbol include/bpg.h declares
EXPORTS = [
    "bpg_create", "bpg_destroy", "bpg_last_error", "bpg_device_name",
    "bpg_plan_create", "bpg_validate_study", "bpg_plan_destroy", "bpg_plan_num_voxels",
    "bpg_plan_permutation_phase", "bpg_plan_permutation_phase_noalloc", "bpg_plans_permutation_phase", "bpg_plan_last_timing", "bpg_plan_observed_maps",
    "bpg_plan_glm",
    "bpg_plan_pvalues", "bpg_plan_pvalues_ties", "bpg_plan_last_tie_flags", "bpg_gen_sign_matrix", "bpg_gen_perm_two_sample",
    "bpg_gen_perm_correlation", "bpg_detect_designs", "bpg_max_permutations",
    "bpg_model_precompute",
    "bpg_gen_perm_first_level", "bpg_fl_plan_create", "bpg_fl_plan_destroy", "bpg_fl_plan_num_voxels",
    "bpg_fl_plan_ar_estimates", "bpg_fl_plan_whitened", "bpg_fl_plan_permutation_phase",
    "bpg_fl_plan_last_timing",
    "bpg_plan_tune", "bpg_tune_save", "bpg_tune_load", "bpg_tune_entries",
    "bpg_batch_create", "bpg_batch_run", "bpg_batch_last_timing", "bpg_batch_destroy",
]


class StudyDesc(ctypes.Structure):
    _fields_ = [("test", c_int), ("inference", c_int), ("precision", c_int),
                ("connectivity", c_int), ("W", c_int), ("H", c_int), ("D", c_int),
                ("N", c_int), ("p", c_int), ("nc", c_int), ("Y", c_void_p),
                ("mask", c_void_p), ("X", c_void_p), ("C", c_void_p), ("cdt", c_float)]


class FlDesc(ctypes.Structure):
    _fields_ = [("test", c_int), ("inference", c_int), ("precision", c_int), ("connectivity", c_int),
                ("W", c_int), ("H", c_int),
                ("D", c_int), ("T", c_int), ("p", c_int), ("nc", c_int), ("volumes", c_void_p),
                ("mask", c_void_p), ("X", c_void_p), ("C", c_void_p), ("voxel_size", c_float * 3),
                ("ar_smoothing_fwhm", c_float), ("cdt", c_float)]


class Timing(ctypes.Structure):
    _fields_ = [("phase_ms", c_double), ("transform_ms", c_double), ("operand_ms", c_double),
                ("contract_ms", c_double), ("cluster_ms", c_double),
                ("contract_launches", c_int), ("perm_voxels", c_long),
                ("fit_columns", c_int), ("slots_per_perm", c_int), ("precision_used", c_int),
                ("unsafe_waves", c_int), ("fp32_permutations", c_long), ("fp64_permutations", c_long),
                ("cdt_tie_maps", c_long), ("pvalue_ties", c_long), ("fp64_samples", c_long)]


class TuneResult(ctypes.Structure):
    _fields_ = [("variant", c_int), ("rows_per_wave", c_int), ("chunk_pow", c_int), ("vt_per_chunk", c_int),
                ("default_ms", c_double), ("best_ms", c_double), ("candidates", c_int)]


class BpgError(RuntimeError):
    def __init__(self, code: int, msg: str = ""):
        self.code = code
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")


_lib = None


def load() -> ctypes.CDLL:
    """Loads libbpg.so from the package directory (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() or "
                           "python broccoli_amd/build.py")
    lib = ctypes.CDLL(str(LIB_PATH))
    vp = c_void_p
    sig = {
        "bpg_create": (c_int, [c_int, POINTER(c_void_p)]),
        "bpg_destroy": (None, [vp]),
        "bpg_last_error": (ctypes.c_char_p, [vp]),
        "bpg_device_name": (ctypes.c_char_p, [vp]),
        "bpg_plan_create": (c_int, [vp, POINTER(StudyDesc), POINTER(c_void_p)]),
        "bpg_plan_destroy": (None, [vp]),
        "bpg_validate_study": (c_int, [POINTER(StudyDesc), ctypes.c_char_p, c_int]),
        "bpg_plan_num_voxels": (c_long, [vp]),
        "bpg_plan_permutation_phase": (c_int, [vp, c_int, vp, vp, c_long, vp, vp]),
        "bpg_plans_permutation_phase": (c_int, [vp, c_int, vp, vp, vp, vp, vp]),
        "bpg_plan_permutation_phase_noalloc": (c_int, [vp, c_int, vp, vp, c_long, vp]),
        "bpg_plan_glm": (c_int, [vp, vp, vp, vp, vp]),
        "bpg_plan_last_timing": (c_int, [vp, POINTER(Timing)]),
        "bpg_plan_observed_maps": (c_int, [vp, vp]),
        "bpg_plan_pvalues": (c_int, [vp, vp, vp, vp, vp]),
        "bpg_plan_pvalues_ties": (c_int, [vp, vp, vp, vp, vp, vp]),
        "bpg_plan_last_tie_flags": (c_int, [vp, vp]),
        "bpg_gen_sign_matrix": (c_int, [c_int, c_long, c_int, c_ulonglong, vp]),
        "bpg_gen_perm_two_sample": (c_int, [c_int, c_int, c_int, c_long, c_int, c_ulonglong, vp]),
        "bpg_gen_perm_correlation": (c_int, [c_int, c_long, c_int, c_ulonglong, vp]),
        "bpg_detect_designs": (c_int, [c_int, c_int, c_int, vp, vp, c_int, vp, vp, vp]),
        "bpg_max_permutations": (c_double, [c_int, c_int, c_int, c_int]),
        "bpg_model_precompute": (c_int, [c_int, c_int, c_int, c_int, vp, vp, vp, vp]),
        "bpg_gen_perm_first_level": (c_int, [c_int, c_long, c_int, c_ulonglong, vp]),
        "bpg_fl_plan_create": (c_int, [vp, POINTER(FlDesc), POINTER(c_void_p)]),
        "bpg_fl_plan_destroy": (None, [vp]),
        "bpg_fl_plan_num_voxels": (c_long, [vp]),
        "bpg_fl_plan_ar_estimates": (c_int, [vp, vp]),
        "bpg_fl_plan_whitened": (c_int, [vp, vp]),
        "bpg_fl_plan_permutation_phase": (c_int, [vp, c_int, vp, c_long, vp, vp]),
        "bpg_fl_plan_last_timing": (c_int, [vp, POINTER(Timing)]),
        "bpg_plan_tune": (c_int, [vp, c_int, c_long, POINTER(TuneResult)]),
        "bpg_tune_save": (c_int, [vp, ctypes.c_char_p]),
        "bpg_tune_load": (c_int, [vp, ctypes.c_char_p]),
        "bpg_tune_entries": (c_int, [vp]),
        "bpg_batch_create": (c_int, [vp, c_int, vp, vp, POINTER(c_void_p)]),
        "bpg_batch_run": (c_int, [vp, vp, vp, vp]),
        "bpg_batch_last_timing": (c_int, [vp, POINTER(c_double), POINTER(c_double), POINTER(c_int), POINTER(c_int)]),
        "bpg_batch_destroy": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(c_void_p)


def _check(rc: int, ctx=None):
    if rc != OK:
        msg = ""
        if ctx is not None:
            m = load().bpg_last_error(ctx)
            msg = m.decode() if m else ""
        raise BpgError(rc, msg)


# ---------------------------------------------------------------- generators

def gen_sign_matrix(N: int, K: int, compat: bool = True, seed: int = 0) -> np.ndarray:
    """compat: the reference stream (seed != 0 reseeds it like srand(seed))."""
    out = np.zeros((K, N), np.float32)
    _check(load().bpg_gen_sign_matrix(N, K, int(compat), seed, _ptr(out)))
    return out


def gen_perm_two_sample(N: int, n1: int, K: int, compat: bool = True, seed: int = 0) -> np.ndarray:
    out = np.zeros((K, N), np.uint16)
    _check(load().bpg_gen_perm_two_sample(N, n1, N - n1, K, int(compat), seed, _ptr(out)))
    return out


def gen_perm_correlation(N: int, K: int, compat: bool = True, seed: int = 0) -> np.ndarray:
    out = np.zeros((K, N), np.uint16)
    _check(load().bpg_gen_perm_correlation(N, K, int(compat), seed, _ptr(out)))
    return out


def detect_designs(X: np.ndarray, C: np.ndarray, ftest: bool):
    """X: (N, p) design, C: (nc, p) contrasts -> (designs, n1, n2)."""
    N, p = X.shape
    nc = C.shape[0]
    Xc = np.asfortranarray(X, np.float32).ravel(order="F").copy()
    Cc = np.ascontiguousarray(C, np.float32).ravel()
    d = np.zeros(nc, np.int32)
    n1 = np.zeros(nc, np.int32)
    n2 = np.zeros(nc, np.int32)
    _check(load().bpg_detect_designs(N, p, nc, _ptr(Xc), _ptr(Cc), int(ftest), _ptr(d), _ptr(n1), _ptr(n2)))
    g = 1 if ftest else nc
    return d[:g], n1[:g], n2[:g]


def max_permutations(test: int, design: int, N: int, n2: int) -> float:
    return load().bpg_max_permutations(test, design, N, n2)


def model_precompute(X: np.ndarray, C: np.ndarray, ftest: bool):
    N, p = X.shape
    nc = C.shape[0]
    Xc = np.asfortranarray(X, np.float32).ravel(order="F").copy()
    Cc = np.ascontiguousarray(C, np.float32).ravel()
    xtxxt = np.zeros(N * p, np.float32)
    ctxtxc = np.zeros(nc * nc if ftest else nc, np.float32)
    _check(load().bpg_model_precompute(N, p, nc, int(ftest), _ptr(Xc), _ptr(Cc), _ptr(xtxxt), _ptr(ctxtxc)))
    return xtxxt, ctxtxc


# ---------------------------------------------------------------- device API

class Context:
    def __init__(self, device: int = 0):
        self._lib = load()
        h = c_void_p()
        rc = self._lib.bpg_create(device, ctypes.byref(h))
        if rc != OK:
            raise BpgError(rc, f"bpg_create(device={device})")
        self.handle = h

    @property
    def device_name(self) -> str:
        return self._lib.bpg_device_name(self.handle).decode()

    def tune_save(self, path) -> None:
        _check(self._lib.bpg_tune_save(self.handle, str(path).encode()), self.handle)

    def tune_load(self, path) -> int:
        n = self._lib.bpg_tune_load(self.handle, str(path).encode())
        if n < 0:
            _check(n, self.handle)
        return n

    @property
    def tune_entries(self) -> int:
        return self._lib.bpg_tune_entries(self.handle)

    def close(self):
        if self.handle:
            self._lib.bpg_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _study_desc(s: dict, precision):
    """bpg_study_desc of a study dict, and the arrays it points into."""
    keep = (np.ascontiguousarray(s["Y"], np.float32), np.ascontiguousarray(s["mask"], np.float32),
            np.asfortranarray(s["X"], np.float32).ravel(order="F").copy(),
            np.ascontiguousarray(s["C"], np.float32).ravel())
    d = StudyDesc()
    d.test = s["test"]
    d.inference = s.get("inference", VOXEL)
    d.precision = PRECISIONS[precision] if isinstance(precision, str) else precision
    d.connectivity = s.get("connectivity", 26)
    d.W, d.H, d.D = s["W"], s["H"], s["D"]
    d.N, d.p, d.nc = s["N"], s["p"], s["nc"]
    d.Y, d.mask, d.X, d.C = (_ptr(a) for a in keep)
    d.cdt = s.get("cdt", 0.0)
    return d, keep


def validate_study(study: dict, precision: str = "split_bf16"):
    """bpg_validate_study: (status, message); no device needed."""
    d, keep = _study_desc(study, precision)
    buf = ctypes.create_string_buffer(256)
    rc = load().bpg_validate_study(ctypes.byref(d), buf, 256)
    return rc, buf.value.decode()


class Plan:
    """A device-resident study (bpg_plan).

    study: dict with W, H, D, N, p, nc, test, Y (N, D, H, W) float32, mask
    (D, H, W) float32, X (N, p), C (nc, p); optional inference, cdt,
    connectivity.
    """

    def __init__(self, ctx: Context, study: dict, precision: str = "split_bf16"):
        self.ctx = ctx
        self._lib = ctx._lib
        self.study = study
        s = study
        d, self._keep = _study_desc(study, precision)
        self.desc = d
        h = c_void_p()
        _check(self._lib.bpg_plan_create(ctx.handle, ctypes.byref(d), ctypes.byref(h)), ctx.handle)
        self.handle = h
        self.nmaps = s["nc"] if s["test"] == TTEST else 1
        self.G = s["W"] * s["H"] * s["D"]

    @property
    def num_voxels(self) -> int:
        return self._lib.bpg_plan_num_voxels(self.handle)

    def permutation_phase(self, map_index: int, perm=None, sign=None, want_stats: bool = False):
        """Runs all permutations of one statistical map.

        Returns (maxdist float32[K], stats float32[K, V_mask] or None)."""
        if sign is not None:
            sign = np.ascontiguousarray(sign, np.float32)
            K = sign.shape[0]
        else:
            perm = np.ascontiguousarray(perm, np.uint16)
            K = perm.shape[0]
        maxdist = np.zeros(K, np.float32)
        stats = np.zeros((K, self.num_voxels), np.float32) if want_stats else None
        _check(self._lib.bpg_plan_permutation_phase(self.handle, map_index, _ptr(perm), _ptr(sign), K,
                                                    _ptr(maxdist), _ptr(stats)), self.ctx.handle)
        return maxdist, stats

    def permutation_phase_noalloc(self, map_index: int, d_perm, d_sign, K: int, d_maxdist):
        """bpg_plan_permutation_phase_noalloc on device pointers (ints or
        c_void_p): permutations or signs in, K null values out."""
        _check(self._lib.bpg_plan_permutation_phase_noalloc(self.handle, map_index, d_perm, d_sign, K, d_maxdist),
               self.ctx.handle)

    def timing(self) -> Timing:
        t = Timing()
        _check(self._lib.bpg_plan_last_timing(self.handle, ctypes.byref(t)))
        return t

    def tune(self, map_index: int, K: int) -> TuneResult:
        """bpg_plan_tune: searches the split-BF16 contraction's shape for
        this map and K on the device and keeps the fastest in the context."""
        r = TuneResult()
        _check(self._lib.bpg_plan_tune(self.handle, map_index, K, ctypes.byref(r)), self.ctx.handle)
        return r

    def observed_maps(self) -> np.ndarray:
        out = np.zeros(self.nmaps * self.G, np.float32)
        _check(self._lib.bpg_plan_observed_maps(self.handle, _ptr(out)), self.ctx.handle)
        return out.reshape(self.nmaps, self.G)

    def glm(self):
        """Betas (p, G), maps (nmaps, G), residuals (N, G), residual variances (G)."""
        p = 1 if self.study["test"] == GROUP_MEAN else self.study["p"]
        N = self.study["N"]
        b = np.zeros(p * self.G, np.float32)
        m = np.zeros(self.nmaps * self.G, np.float32)
        r = np.zeros(N * self.G, np.float32)
        v = np.zeros(self.G, np.float32)
        _check(self._lib.bpg_plan_glm(self.handle, _ptr(b), _ptr(m), _ptr(r), _ptr(v)), self.ctx.handle)
        return b.reshape(p, self.G), m.reshape(self.nmaps, self.G), r.reshape(N, self.G), v

    def pvalues(self, maxdists, want_labels: bool = False, want_ties: bool = False):
        """P (nmaps, G); with want_labels also the cluster labels and with
        want_ties the per-voxel tie counts (bpg_plan_pvalues_ties), in that
        order after P."""
        Ks = np.array([len(m) for m in maxdists], np.int64)
        arrs = [np.ascontiguousarray(m, np.float32) for m in maxdists]
        ptrs = (c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        P = np.zeros(self.nmaps * self.G, np.float32)
        labels = np.zeros(self.nmaps * self.G, np.uint32) if want_labels else None
        ties = np.zeros(self.nmaps * self.G, np.uint32) if want_ties else None
        if want_ties:
            _check(self._lib.bpg_plan_pvalues_ties(self.handle, _ptr(Ks), ctypes.cast(ptrs, c_void_p), _ptr(P),
                                                   _ptr(labels), _ptr(ties)), self.ctx.handle)
        else:
            _check(self._lib.bpg_plan_pvalues(self.handle, _ptr(Ks), ctypes.cast(ptrs, c_void_p), _ptr(P),
                                              _ptr(labels)), self.ctx.handle)
        out = [P.reshape(self.nmaps, self.G)]
        if want_labels:
            out.append(labels.reshape(self.nmaps, self.G))
        if want_ties:
            out.append(ties.reshape(self.nmaps, self.G))
        return tuple(out) if len(out) > 1 else out[0]

    def last_tie_flags(self, K: int) -> np.ndarray:
        """Per permutation of the last phase: 1 when its map had a statistic
        within the error band of the cluster-defining threshold."""
        out = np.zeros(K, np.uint8)
        _check(self._lib.bpg_plan_last_tie_flags(self.handle, _ptr(out)))
        return out

    def close(self):
        if self.handle:
            self._lib.bpg_plan_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def permutation_phases(plans, maps, perms):
    """The permutation phases of several studies in one call
    (bpg_plans_permutation_phase): plans[i] with map maps[i] and perms[i], a
    dict with "perm" or "sign".  Returns the list of maxdist arrays."""
    n = len(plans)
    lib = plans[0]._lib
    P = (c_void_p * n)(*[p.handle for p in plans])
    M = (c_int * n)(*maps)
    keep = []
    pp = (c_void_p * n)()
    ss = (c_void_p * n)()
    Ks = (c_long * n)()
    outs = []
    for i, d in enumerate(perms):
        if "sign" in d:
            a = np.ascontiguousarray(d["sign"], np.float32)
            ss[i] = a.ctypes.data
        else:
            a = np.ascontiguousarray(d["perm"], np.uint16)
            pp[i] = a.ctypes.data
        keep.append(a)
        Ks[i] = a.shape[0]
        outs.append(np.zeros(a.shape[0], np.float32))
    O = (c_void_p * n)(*[o.ctypes.data for o in outs])
    _check(lib.bpg_plans_permutation_phase(P, n, M, pp, ss, Ks, O), plans[0].ctx.handle)
    return outs


class Batch:
    """bpg_batch: the permutation phases of several studies captured once as
    a hipGraph (persistent contraction per class); run() replays it."""

    def __init__(self, plans, maps, Ks):
        self.plans = list(plans)
        self._lib = self.plans[0]._lib
        n = len(self.plans)
        self.Ks = [int(k) for k in Ks]
        P = (c_void_p * n)(*[p.handle for p in self.plans])
        M = (c_int * n)(*maps)
        Kc = (c_long * n)(*self.Ks)
        h = c_void_p()
        _check(self._lib.bpg_batch_create(P, n, M, Kc, ctypes.byref(h)), self.plans[0].ctx.handle)
        self.handle = h

    def run(self, perms):
        """perms[i]: dict with "perm" or "sign" of K[i] rows; returns the
        list of maxdist arrays."""
        n = len(self.plans)
        pp = (c_void_p * n)()
        ss = (c_void_p * n)()
        keep, outs = [], []
        for i, d in enumerate(perms):
            if "sign" in d:
                a = np.ascontiguousarray(d["sign"], np.float32)
                ss[i] = a.ctypes.data
            else:
                a = np.ascontiguousarray(d["perm"], np.uint16)
                pp[i] = a.ctypes.data
            assert a.shape[0] == self.Ks[i]
            keep.append(a)
            outs.append(np.zeros(self.Ks[i], np.float32))
        O = (c_void_p * n)(*[o.ctypes.data for o in outs])
        _check(self._lib.bpg_batch_run(self.handle, pp, ss, O), self.plans[0].ctx.handle)
        return outs

    def last_timing(self):
        """(run ms, graph ms, studies with repeated permutations, persistent
        launches) of the last run."""
        ms, gms, fb, nc = c_double(), c_double(), c_int(), c_int()
        _check(self._lib.bpg_batch_last_timing(self.handle, ctypes.byref(ms), ctypes.byref(gms), ctypes.byref(fb),
                                               ctypes.byref(nc)))
        return ms.value, gms.value, fb.value, nc.value

    def close(self):
        if self.handle:
            self._lib.bpg_batch_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ------------------------------------------------------------- first level

def gen_perm_first_level(T: int, K: int, compat: bool = True, seed: int = 0) -> np.ndarray:
    """Time-point permutations of the first-level test, (K, T) uint16."""
    out = np.zeros((K, T), np.uint16)
    _check(load().bpg_gen_perm_first_level(T, K, int(compat), seed, _ptr(out)))
    return out


class FlPlan:
    """A first-level study on the device (bpg_fl_plan): regression and AR(4)
    whitening run at creation.

    study: dict with Y (T, D, H, W) float32, mask (D, H, W), X (T, p), C
    (nc, p), voxel (3 sizes in mm); optional test (TTEST), inference
    (VOXEL), precision ("split_f16" (the default), "split_bf16" or "fp32":
    FP32 recursion with the projections on the matrix cores and FP64
    repeats; "fp64"), cdt, connectivity (26), fwhm (AR smoothing, 8 mm).
    """

    def __init__(self, ctx: Context, study: dict):
        self.ctx = ctx
        self._lib = ctx._lib
        s = study
        T, D, H, W = s["Y"].shape
        self._Y = np.ascontiguousarray(s["Y"], np.float32)
        self._mask = np.ascontiguousarray(s["mask"], np.float32)
        self._X = np.asfortranarray(s["X"], np.float32).ravel(order="F").copy()
        C = np.atleast_2d(np.asarray(s["C"], np.float32))
        self._C = np.ascontiguousarray(C).ravel()
        d = FlDesc()
        d.test = s.get("test", TTEST)
        d.inference = s.get("inference", VOXEL)
        prec = s.get("precision", "split_f16")
        d.precision = FL_PRECISIONS[prec] if isinstance(prec, str) else prec
        d.connectivity = s.get("connectivity", 26)
        d.W, d.H, d.D, d.T = W, H, D, T
        d.p, d.nc = s["X"].shape[1], C.shape[0]
        d.volumes, d.mask, d.X, d.C = _ptr(self._Y), _ptr(self._mask), _ptr(self._X), _ptr(self._C)
        for i in range(3):
            d.voxel_size[i] = float(s["voxel"][i])
        d.ar_smoothing_fwhm = float(s.get("fwhm", 8.0))
        d.cdt = float(s.get("cdt", 0.0))
        self.desc = d
        h = c_void_p()
        _check(self._lib.bpg_fl_plan_create(ctx.handle, ctypes.byref(d), ctypes.byref(h)), ctx.handle)
        self.handle = h
        self.shape = (T, D, H, W)
        self.G = W * H * D
        self.nmaps = C.shape[0] if d.test == TTEST else 1

    @property
    def num_voxels(self) -> int:
        return self._lib.bpg_fl_plan_num_voxels(self.handle)

    def ar_estimates(self) -> np.ndarray:
        out = np.zeros(4 * self.G, np.float32)
        _check(self._lib.bpg_fl_plan_ar_estimates(self.handle, _ptr(out)), self.ctx.handle)
        return out.reshape((4,) + self.shape[1:])

    def whitened(self) -> np.ndarray:
        out = np.zeros(self.shape[0] * self.G, np.float32)
        _check(self._lib.bpg_fl_plan_whitened(self.handle, _ptr(out)), self.ctx.handle)
        return out.reshape(self.shape)

    def permutation_phase(self, map_index: int, perm, want_stats: bool = False):
        """(maxdist float32[K], stats float32[K, V_mask] or None)."""
        perm = np.ascontiguousarray(perm, np.uint16)
        K = perm.shape[0]
        maxdist = np.zeros(K, np.float32)
        stats = np.zeros((K, self.num_voxels), np.float32) if want_stats else None
        _check(self._lib.bpg_fl_plan_permutation_phase(self.handle, map_index, _ptr(perm), K, _ptr(maxdist),
                                                       _ptr(stats)), self.ctx.handle)
        return maxdist, stats

    def timing(self) -> Timing:
        t = Timing()
        _check(self._lib.bpg_fl_plan_last_timing(self.handle, ctypes.byref(t)))
        return t

    def close(self):
        if self.handle:
            self._lib.bpg_fl_plan_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
