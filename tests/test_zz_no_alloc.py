"""The permutation phase without device allocation
(bpg_plan_permutation_phase_noalloc) against the allocating entry point
(bpg_plan_permutation_phase) on the same inputs: the null values must be
identical (the same kernels), on two calls in a row, the second smaller."""
import ctypes
from ctypes import byref, c_size_t, c_void_p

import numpy as np
import pytest

from broccoli_amd import capi, synth

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(c_void_p), c_size_t]
        _hip.hipMemcpy.argtypes = [c_void_p, c_void_p, c_size_t, ctypes.c_int]
        _hip.hipFree.argtypes = [c_void_p]
    return _hip


class DevBuf:
    """A device buffer of the caller (the test allocates it; the entry point
    under test must not allocate anything)."""

    def __init__(self, host: np.ndarray):
        self.nbytes = host.nbytes
        self.ptr = c_void_p()
        assert hip().hipMalloc(byref(self.ptr), max(self.nbytes, 4)) == 0
        assert hip().hipMemcpy(self.ptr, host.ctypes.data_as(c_void_p), self.nbytes, 1) == 0  # host -> device

    def get(self, out: np.ndarray) -> np.ndarray:
        assert hip().hipMemcpy(out.ctypes.data_as(c_void_p), self.ptr, out.nbytes, 2) == 0  # device -> host
        return out

    def free(self):
        hip().hipFree(self.ptr)


def run_both(ctx, st, perms, prec, K):
    with capi.Plan(ctx, st, prec) as plan:
        want, _ = plan.permutation_phase(0, **{k: v[:K] for k, v in perms.items()})
        key = "sign" if "sign" in perms else "perm"
        src = np.ascontiguousarray(perms[key][:K])
        d_in = DevBuf(src)
        d_out = DevBuf(np.full(K, np.nan, np.float32))
        try:
            args = (d_in.ptr, None) if key == "perm" else (None, d_in.ptr)
            plan.permutation_phase_noalloc(0, args[0], args[1], K, d_out.ptr)
            got = d_out.get(np.empty(K, np.float32))
            tm = plan.timing()
        finally:
            d_in.free()
            d_out.free()
    return want, got, tm


@pytest.mark.parametrize("wl,prec", [(2, "split_bf16"), (1, "split_bf16"), (2, "fp32"), (3, "split_bf16")])
def test_noalloc_phase_equals_the_allocating_phase(ctx, wl, prec):
    st = synth.study(wl, scale=0.2)
    st["inference"] = capi.VOXEL
    perms = synth.permutations(st, 300, seed=77)
    for K in (300, 57):  # two calls in a row, the second smaller
        print(f"no-alloc phase: workload {wl} {prec} K={K} V={st['V']}", flush=True)
        want, got, tm = run_both(ctx, st, perms, prec, K)
        assert tm.unsafe_waves == 0
        assert np.array_equal(got, want), np.flatnonzero(got != want)[:10]


def test_noalloc_phase_refuses_what_it_cannot_do(ctx):
    st = synth.study(2, scale=0.1)
    st["inference"] = capi.VOXEL
    perms = synth.permutations(st, 16, seed=5)
    bad = perms["perm"].copy()
    bad[3, 2] = st["N"]
    with capi.Plan(ctx, st, "split_bf16") as plan:
        d_in = DevBuf(np.ascontiguousarray(bad))
        d_out = DevBuf(np.zeros(16, np.float32))
        try:
            with pytest.raises(capi.BpgError, match="out of range"):
                plan.permutation_phase_noalloc(0, d_in.ptr, None, 16, d_out.ptr)
            # a K whose scratch exceeds the static array
            with pytest.raises(capi.BpgError, match="scratch"):
                plan.permutation_phase_noalloc(0, d_in.ptr, None, 1 << 24, d_out.ptr)
        finally:
            d_in.free()
            d_out.free()
        # the plan stays usable
        md, _ = plan.permutation_phase(0, **perms)
        assert md.shape == (16,)
    st["inference"] = capi.CLUSTER_EXTENT
    with capi.Plan(ctx, st, "split_bf16") as plan:
        with pytest.raises(capi.BpgError, match="voxel inference"):
            plan.permutation_phase_noalloc(0, 1, None, 16, 1)
