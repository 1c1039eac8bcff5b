"""Benchmark of the permutation phase (perm*voxels/s, GLM max-T) on one MI355X.

Workload: the two-sample GLM with covariates -- 91x109x91 grid with an
ellipsoid mask (~228k voxels), N=60 subjects (30/30), p=4 regressors
(intercept, group, demeaned age, sex), t contrast [0 1 0 0], 10,000 row
permutations -- on synthetic data of that shape.

One step = the whole permutation phase of the contrast: nuisance
residualisation and operand split of the HBM-resident data, permutation
operand build, the fused contraction + t + max-T kernel over all 10,000
permutations, and the max-T array copied to the host.  The study is
uploaded once before timing.

Printed JSON line: metric/value for the split-BF16 MFMA form (the headline),
the plain FP32 form measured in the same run, accuracy against the FP64
oracle (every statistic of sampled permutations; the quantised max-T of all
K permutations and the p-value counts of the observed map, with their ties),
the roofline of the dominant kernel (its launch time from HIP events on the
plan's stream), and the reference algorithm restated on the host cores
(cpu_baseline).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from broccoli_amd import capi, synth  # noqa: E402

METRIC = "perm·voxels/s (GLM max-T), split-BF16 MFMA vs plain FP32, t err≤1e-4(1+|t|)"
PEAK_BF16_TFLOPS = 2500.0  # MI355X dense BF16 MFMA
PEAK_FP32_TFLOPS = 157.3


def run_batch(args):
    """Workload 4: 64 studies, one batched call per step."""
    studies = synth.studies_batch(64)
    ctx = capi.Context(0)
    plans, perms, Ks = [], [], []
    for i, st in enumerate(studies):
        mx = capi.max_permutations(2 if st["test"] == capi.GROUP_MEAN else 0,
                                   capi.CORRELATION, st["N"], 0)
        K = int(min(st["K"], mx))
        Ks.append(K)
        perms.append(synth.permutations(st, K, seed=20260400 + i))
        plans.append(capi.Plan(ctx, st, "split_bf16"))
    # the studies' shapes from the autotuner (set-up, outside the timed region)
    t0 = time.perf_counter()
    tuned = [p.tune(0, K) for p, K in zip(plans, Ks)] if not args.no_tune else []
    tune_s = time.perf_counter() - t0
    pv = sum(K * st["V"] for K, st in zip(Ks, studies))
    flops = sum(2.0 * st["N"] * st["p"] * K * st["V"] for K, st in zip(Ks, studies))
    print(f"# 64 studies ready: {pv:.3e} perm*voxels per step", file=sys.stderr, flush=True)
    maps = [0] * len(plans)
    for _ in range(args.warmup):
        capi.permutation_phases(plans, maps, perms)
    hip_sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        capi.permutation_phases(plans, maps, perms)
    hip_sync()
    dt = (time.perf_counter() - t0) / args.steps
    kern = sum(p.timing().contract_ms for p in plans)
    n_fp32 = sum(p.timing().precision_used == capi.PREC_FP32 for p in plans)
    p_fp32 = sum(p.timing().fp32_permutations for p in plans)
    t0 = time.perf_counter()
    md_single = [p.permutation_phase(0, **d)[0] for p, d in zip(plans, perms)]
    seq = time.perf_counter() - t0
    print(f"# stream batch: {dt*1e3:.3f} ms/step (one study after another: {seq*1e3:.3f} ms)", file=sys.stderr,
          flush=True)
    # the batch as one hipGraph with a persistent contraction per class
    stream_dt = dt
    batch = capi.Batch(plans, maps, Ks)
    for _ in range(args.warmup):
        md_graph = batch.run(perms)
    hip_sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        md_graph = batch.run(perms)
    hip_sync()
    dt = (time.perf_counter() - t0) / args.steps
    _, graph_ms, fallback, classes = batch.last_timing()
    same = sum(int(np.array_equal(a, b)) for a, b in zip(md_graph, md_single))
    batch.close()
    print(f"# graph batch: {dt*1e3:.3f} ms/step (graph {graph_ms:.3f} ms; {classes} persistent launches, "
          f"{fallback} studies with repeated permutations)",
          file=sys.stderr, flush=True)
    achieved = flops / dt / 1e12
    res = {
        "metric": METRIC, "value": pv / dt, "unit": "perm·voxels/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16", "accumulate": "fp32", "data": "synthetic",
        "config": {"workload": WORKLOADS[4], "studies": len(plans), "perm_voxels": pv,
                   "V_mean": float(np.mean([st["V"] for st in studies])), "precision": "split_bf16",
                   "form": "one hipGraph per step: persistent contraction per class of studies"},
        "stream_batch_ms": stream_dt * 1e3,
        "one_after_another_ms": seq * 1e3,
        "graph": {"persistent_launches": classes, "graph_ms_last_step": graph_ms,
                  "studies_with_repeated_permutations": fallback,
                  "nulls_equal_to_single_study_calls": same},
        "autotuner": {"buckets": ctx.tune_entries, "seconds": tune_s,
                      "studies_off_default": int(sum(1 for r in tuned if r.variant or r.chunk_pow)),
                      "sum_default_ms": float(sum(r.default_ms for r in tuned)),
                      "sum_tuned_ms": float(sum(r.best_ms for r in tuned))} if tuned else None,
        "studies_rerun_in_fp32": int(n_fp32),
        "permutations_repeated_in_fp32": int(p_fp32), "permutations_total": int(sum(Ks)),
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_BF16_TFLOPS, "traffic": None, "kernel": "k_contract_bf16",
                     "algorithmic_flop_per_step": flops, "sum_of_contraction_ms": kern,
                     "note": "achieved over the wall time of the batch (the studies overlap)"},
    }
    for p in plans:
        p.close()
    ctx.close()
    print(json.dumps(res))


WORKLOADS = {
    1: "one-sample t by sign flipping, 91x109x91 grid, ellipsoid mask, N=20, p=1, 1000 sign flips, voxel max-T",
    4: "64 studies on the 91x109x91 grid, own ellipsoid masks (50k-230k voxels), N~U{12..120}, p~U{1..6}, "
       "t on the last covariate (sign flips when p=1), 5000 permutations each, voxel max-T, one hipGraph per step",
    2: "two-sample GLM t-test with covariates, 91x109x91 grid, ellipsoid mask, N=60 (30/30), p=4, "
       "contrast [0 1 0 0], 10000 row permutations, voxel max-T",
    3: "F test over 3 group dummies, 91x109x91 grid, N=100, p=6, smooth noise (FWHM 3), 10000 permutations, "
       "cluster-extent max (F>4, 26-connectivity)",
    5: "t test on a covariate, 182x218x182 grid (~1.83M voxels), N=200, p=8, 10000 row permutations, voxel max-T",
    6: "first-level permutation test: 64x64x33 EPI grid (3.75x3.75x4 mm), ellipsoid mask, T=240, p=12 "
       "(intercept, trend, block task, sinusoid, 8 nuisance), AR(4) noise, t on the task, AR(4) whitening "
       "(8 mm smoothing), 10000 time-point permutations, voxel max-T",
}
WORKLOAD_5F = ("2-row F test over group dummies 1 and 2, 182x218x182 grid (~1.83M voxels), N=200, p=8, "
               "10000 row permutations, voxel max-T")
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector
PEAK_FP32_MATRIX_TFLOPS = 157.3  # MI355X FP32 matrix


# rocprofv3 summaries of the dominant kernel per workload: HBM traffic per
# launch from the FETCH_SIZE / WRITE_SIZE passes, scaled by a copy of known
# size (the counters cannot be read from inside this process)
PROFILES = {2: ROOT / "profiles" / "k_contract_bf16_two_sample_glm_10k.md",
            6: ROOT / "profiles" / "k_fl_perm_hx_first_level_10k.md"}


def profiled_traffic(workload):
    f = PROFILES.get(workload)
    if f is None or not f.exists():
        return None, None
    m = re.search(r"traffic_bytes_per_launch: (\d+)", f.read_text())
    return (int(m.group(1)), str(f.relative_to(ROOT))) if m else (None, None)


def run_first_level(args):
    """Workload 6: the first-level permutation test of one subject."""
    st = synth.first_level_workload()
    T, p = st["T"], st["p"]
    K = args.K
    perm = capi.gen_perm_first_level(T, K, compat=False, seed=20260006)
    ctx = capi.Context(0)

    def timed(precision, steps, warmup):
        t0 = time.perf_counter()
        plan = capi.FlPlan(ctx, dict(st, precision=precision))
        setup_wall = time.perf_counter() - t0
        for _ in range(warmup):
            plan.permutation_phase(0, perm)
        hip_sync()
        t0 = time.perf_counter()
        kern = []
        for _ in range(steps):
            md, _ = plan.permutation_phase(0, perm)
            kern.append(plan.timing().contract_ms)
        hip_sync()
        return plan, md, (time.perf_counter() - t0) / steps, float(np.mean(kern)), setup_wall

    plan64, md64, dt64, kern64, _ = timed("fp64", max(2, min(10, args.steps // 2)), 1)
    plan64.close()
    plan32, md32, dt32, kern32, _ = timed("fp32", max(2, min(10, args.steps // 2)), 1)
    plan32.close()
    planb, mdb, dtb, kernb, _ = timed("split_bf16", max(2, min(10, args.steps // 2)), 1)
    planb.close()
    plan, md, dt, kern_ms, setup_wall = timed("split_f16", args.steps, args.warmup)
    V = plan.num_voxels
    tm = plan.timing()
    print(f"# first-level study: V={V} T={T} p={p} K={K} set-up {tm.transform_ms:.1f} ms (wall {setup_wall*1e3:.0f} ms); "
          f"split-FP16 form {dt*1e3:.2f} ms/step (kernel {kern_ms:.2f}), split-BF16 form {dtb*1e3:.2f} (kernel "
          f"{kernb:.2f}), fp32 form {dt32*1e3:.2f} (kernel {kern32:.2f}), fp64 form {dt64*1e3:.2f} ms/step "
          f"(kernel {kern64:.2f})", file=sys.stderr, flush=True)
    # per permutation, voxel and time point: the AR(4) recursion (4 FMA),
    # the projections on the p design columns (p FMA), |x|^2 (1 FMA), and
    # sum x when the constant is outside col(X)
    X = st["X"].astype(np.float64)
    ones = np.ones(T)
    const_in_span = np.linalg.norm(ones - X @ np.linalg.lstsq(X, ones, rcond=None)[0]) < 1e-6 * np.sqrt(T)
    flops = float(T) * (2 * 4 + 2 * p + 2 + (0 if const_in_span else 1)) * K * V
    traffic, traffic_src = profiled_traffic(6)
    achieved = flops / (kern_ms * 1e-3) / 1e12
    res = {
        "metric": "perm·voxels/s (first-level max-T)", "value": K * V / dt, "unit": "perm·voxels/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f16", "accumulate": "fp32", "data": "synthetic",
        "config": {"workload": WORKLOADS[6], "V_mask": V, "T": T, "p": p, "K": K, "precision": "split_f16"},
        "setup_ms": tm.transform_ms,
        "phase_breakdown_ms": {"permutation_kernel": kern_ms, "library_call": tm.phase_ms},
        "permutations_repeated_in_fp64": tm.unsafe_waves,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic, "traffic_source": traffic_src,
                     "algorithmic_bytes_per_launch": 4.0 * T * V + 2.0 * K * T + 4.0 * K,
                     "kernel": "k_fl_perm_hx", "algorithmic_flop_per_launch": flops, "avg_launch_ms": kern_ms,
                     "note": "the recursion, |x|^2 and the FP16 splits on the vector ALUs (the bound), the "
                             "projections on the FP16 matrix cores; algorithmic flop over the measured launch "
                             "time against the FP32 vector peak"},
        "split_bf16_form": {"value": K * V / dtb, "ms_per_step": dtb * 1e3, "kernel_ms": kernb,
                            "maxT_equal_to_split_f16_form": int(np.sum(mdb == md))},
        "fp32_form": {"value": K * V / dt32, "ms_per_step": dt32 * 1e3, "kernel_ms": kern32,
                      "maxT_equal_to_split_f16_form": int(np.sum(md32 == md))},
        "fp64_form": {"value": K * V / dt64, "ms_per_step": dt64 * 1e3, "kernel_ms": kern64,
                      "achieved_tflops": flops / (kern64 * 1e-3) / 1e12,
                      "frac_of_fp64_peak": flops / (kern64 * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
                      "maxT_equal_to_split_f16_form": int(np.sum(md64 == md))},
        "speedup_vs_fp64": dt64 / dt,
    }
    plan.close()
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_first_level(st, perm, args.cpu_seconds)
    ctx.close()
    if args.copy_mib:
        copy_calibration(args.copy_mib)
    print(json.dumps(res))


def host_cores():
    """Host cores this process may use: its affinity set, capped by
    OMP_NUM_THREADS when that is set (the GPU box sets it to the CPU share
    of one GPU)."""
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(omp)) if omp.isdigit() and int(omp) > 0 else n


def cpu_baseline_first_level(st, perm, target_s):
    """The oracle's FP64 restatement of the first-level loop (inverse
    whitening of the whole grid, then the statistic map, OpenMP over the
    host cores) on a bounded sample of the permutations."""
    from oracle import oracle

    ncores = host_cores()
    os.environ.setdefault("OMP_NUM_THREADS", str(ncores))
    wh, ar = oracle.fl_prepare(st["Y"], st["mask"], st["X"], st.get("fwhm", 8.0), st["voxel"])
    V = int((st["mask"] == 1).sum())
    t0 = time.perf_counter()
    oracle.fl_null(0, 0, wh, ar, st["mask"], st["X"], st["C"], perm[:1])
    per = time.perf_counter() - t0
    kc = int(max(1, min(2000, target_s / max(per, 1e-6))))
    t0 = time.perf_counter()
    oracle.fl_null(0, 0, wh, ar, st["mask"], st["X"], st["C"], perm[:kc])
    dt = time.perf_counter() - t0
    return dict(value=kc * V / dt, unit="perm·voxels/s", cores=ncores, kind="port",
                sample=f"{kc} of the {perm.shape[0]} permutations, all {V} voxels, {dt:.1f} s, FP64 "
                       "inverse whitening + statistic map + quantised max per permutation")


def hip_sync():
    try:
        ctypes.CDLL("libamdhip64.so").hipDeviceSynchronize()
    except OSError:
        pass


def run_form(ctx, st, perms, precision, steps, warmup, tune=False):
    plan = capi.Plan(ctx, st, precision)
    if tune:
        # the autotuner's shape for this bucket (set-up, outside the timing)
        key = "sign" if "sign" in perms else "perm"
        r = plan.tune(0, perms[key].shape[0])
        run_form.tuned = {"rows_per_wave": r.rows_per_wave, "chunk_pow": r.chunk_pow,
                          "vt_per_chunk": r.vt_per_chunk, "default_ms": r.default_ms, "tuned_ms": r.best_ms}
    for _ in range(warmup):
        plan.permutation_phase(0, **perms)
    hip_sync()
    kern = []
    t0 = time.perf_counter()
    for _ in range(steps):
        md, _ = plan.permutation_phase(0, **perms)
        kern.append(plan.timing().contract_ms)
    hip_sync()
    dt = (time.perf_counter() - t0) / steps
    tm = plan.timing()
    return plan, md, dt, float(np.mean(kern)), tm


def mfma_busy(st, K, kern_ms, slots, clock_hz=2.4e9, simds=1024):
    """Share of the MFMA pipes' cycles the launch keeps busy: 32x32x16 BF16
    MFMAs issued (32 cycles each; 2 split products for sign operands, 3
    otherwise; subjects padded to 16, 16 // slots permutations per lane
    and half-subtile) over the launch time at the peak clock."""
    N, V = st["N"], st["V"]
    kk = (N + 15) // 16
    per_sub = 2 * (16 // slots)
    n_sub = (K + per_sub - 1) // per_sub
    n_vt = (V + 255) // 256 * 8
    prods = 2 if st["test"] == capi.GROUP_MEAN else 3
    return n_sub * n_vt * kk * prods * 32 / (kern_ms * 1e-3 * clock_hz * simds)


def accuracy(ctx, st, perms, nsample, forms=("split_bf16", "fp32")):
    """Sampled permutations of each form against the FP64 oracle."""
    from oracle import oracle

    sub = {k: v[:nsample] for k, v in perms.items() if k in ("perm", "sign")}
    out = {}
    for prec in forms:
        with capi.Plan(ctx, st, prec) as plan:
            md, stats = plan.permutation_phase(0, want_stats=True, **sub)
        ref = oracle.perm_stats(st, sub, nsample)[:, st["mask"].ravel() == 1.0]
        err = np.abs(stats - ref) / (1.0 + np.abs(ref))
        q64 = oracle.quantised_max(ref.max(axis=1))
        qg = md
        out[prec] = dict(max_err=float(err.max()), sampled_perms=nsample,
                         maxT_quanta_differing=int((q64 != qg).sum()))
    return out


def accuracy_cluster(ctx, st, perms, nsample, forms=("split_bf16", "fp32")):
    """Cluster inference: sampled permutations' statistics of each form
    against the FP64 oracle at the full grid, and their largest clusters
    against the oracle's labelling of the FP64 statistics; a size that
    differs counts as a tie when some FP64 statistic lies within
    1e-4 (1 + |cdt|) of the threshold."""
    from oracle import oracle

    sub = {k: v[:nsample] for k, v in perms.items() if k in ("perm", "sign")}
    inmask = st["mask"].ravel() == 1.0
    ref = oracle.perm_stats(st, sub, nsample)[:, inmask]
    band = 1e-4 * (1.0 + abs(st["cdt"]))
    best64, near = [], []
    for k in range(nsample):
        g = np.zeros(inmask.size, np.float32)
        g[inmask] = ref[k]
        best64.append(oracle.ccl(g.reshape(st["mask"].shape), st["mask"], st["cdt"], st.get("connectivity", 26))[2])
        near.append(bool((np.abs(ref[k] - st["cdt"]) <= band).any()))
    out = {}
    for prec in forms:
        with capi.Plan(ctx, st, prec) as plan:
            md, stats = plan.permutation_phase(0, want_stats=True, **sub)
        err = np.abs(stats - ref) / (1.0 + np.abs(ref))
        differ = [k for k in range(nsample) if md[k] != float(best64[k])]
        out[prec] = dict(max_err=float(err.max()), sampled_perms=nsample, cluster_sizes_differing=len(differ),
                         differing_outside_ties=sum(1 for k in differ if not near[k]))
    return out


def full_k_check(st, perms, K, md, md32, obs, P, ties, ncores):
    """Every permutation's quantised max-T against the FP64 oracle's, and the
    p-value counts of the observed map against those of the FP64 null.

    A quantum may differ only inside the accuracy gate's band: the GPU
    maximum is within 1e-4 (1 + |m|) of the FP64 one, so its quantum lies in
    [trunc((m64 - b) 1e4), trunc((m64 + b) 1e4)].  A count may differ only
    by comparisons the library reports as ties (null values within
    1e-4 (2 + |T|) of T)."""
    from oracle import oracle

    t0 = time.perf_counter()
    m64 = oracle.perm_max64(st, perms, K, nthreads=ncores)
    dt = time.perf_counter() - t0
    # the reference's stored null value of each FP64 maximum, and of the
    # edges of the accuracy gate's band around it (a monotone map)
    q64 = oracle.quantised_max(m64)
    band = 1e-4 * (1.0 + np.abs(m64))
    lo = oracle.quantised_max(m64 - band)
    hi = oracle.quantised_max(m64 + band)
    out = {"permutations": int(K), "oracle_seconds": dt, "oracle_cores": ncores}
    for name, m in (("split_bf16", md), ("fp32", md32)):
        if m is None:
            continue
        d = m != q64
        out[name] = {"maxT_quanta_differing": int(d.sum()),
                     "maxT_outside_band": int((d & ((m < lo) | (m > hi))).sum()),
                     "max_abs_diff": float(np.abs(m.astype(np.float64) - q64).max())}
    mask = st["mask"].ravel() == 1.0
    T = obs[0][mask]
    cg = np.round(P[0][mask].astype(np.float64) * K).astype(np.int64)
    c64 = np.searchsorted(np.sort(q64), T, side="left")
    tv = ties[0][mask].astype(np.int64)
    dc = np.abs(cg - c64)
    out["pvalues"] = {"voxels": int(mask.sum()), "counts_differing": int((dc > 0).sum()),
                      "counts_differing_outside_ties": int((dc > tv).sum()),
                      "ties": int(tv.sum()), "voxels_with_ties": int((tv > 0).sum())}
    return out


def cpu_baseline(st, perms, target_s):
    """The reference's per-permutation loop restated in C (oracle, ref32
    precision, OpenMP over the host cores) on a bounded sample of the same
    workload's permutations."""
    from oracle import oracle

    ncores = host_cores()
    Ystar = oracle.transformed32(st, 0)
    sub = {k: v for k, v in perms.items() if k in ("perm", "sign")}
    one = {k: v[:2] for k, v in sub.items()}
    t0 = time.perf_counter()
    oracle.permutation_phase32(st, one, 2, Ystar, nthreads=ncores)
    per = (time.perf_counter() - t0) / 2
    kc = int(max(2, min(2000, target_s / max(per, 1e-6))))
    s2 = {k: v[:kc] for k, v in sub.items()}
    t0 = time.perf_counter()
    oracle.permutation_phase32(st, s2, kc, Ystar, nthreads=ncores)
    dt = time.perf_counter() - t0
    return dict(value=kc * st["V"] / dt, unit="perm·voxels/s", cores=ncores, kind="port",
                sample=f"{kc} of the 10000 permutations, all {st['V']} voxels, "
                       f"{dt:.1f} s, one full-grid pass + quantised max per permutation")


def copy_calibration(mib):
    """A known-size read + write (torch's elementwise negation: reads and
    writes mib MiB once each) for scaling the traffic counters."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.neg(a, out=b)
    torch.cuda.synchronize()
    print(f"# copy calibration: {mib} MiB read + {mib} MiB written per launch", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--K", type=int, default=10000)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="use the planner's default shapes")
    ap.add_argument("--workload", type=int, default=2, choices=[1, 2, 3, 4, 5, 6],
                    help="2 (default, the metric's workload); 1: sign-flip group mean, K=1000; "
                         "3: F test with cluster-extent inference; 4: 64 studies in one batched call; "
                         "5: 182x218x182 grid, N=200, p=8; 6: first-level permutation test")
    ap.add_argument("--scale", type=float, default=1.0, help="grid scale (1.0 = full size)")
    ap.add_argument("--contrast", default="t", choices=["t", "F"],
                    help="workload 5: the t contrast or the 2-row F contrast over group dummies 1 and 2")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="after the run, launch a plain elementwise copy of this many MiB three times "
                         "(calibrates the FETCH_SIZE / WRITE_SIZE counters against known bytes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the permutation phase is measured on a single "
              "MI355X (across GPUs the permutations would be split, each GPU holding a copy of the "
              "study, with one gather of the K maxima at the end)", file=sys.stderr)
        sys.exit(2)

    if args.workload == 4:
        run_batch(args)
        return
    if args.workload == 6:
        run_first_level(args)
        return
    st = synth.study(args.workload, scale=args.scale, contrast=args.contrast)
    K = args.K if args.workload == 2 or args.K != 10000 else st["K"]
    perms = synth.permutations(st, K, seed=20260000 + args.workload)
    if args.workload != 2:
        # extra data points: no CPU baseline or oracle sample at these sizes
        args.no_cpu = True
    V = st["V"]
    N, p = st["N"], st["p"]
    ctx = capi.Context(0)
    print(f"# study ready: V={V} N={N} p={p} K={K} device={ctx.device_name}", file=sys.stderr, flush=True)

    run_form.tuned = None
    plan, md, dt, kern_ms, tm = run_form(ctx, st, perms, "split_bf16", args.steps, args.warmup, tune=not args.no_tune)
    # the observed map and its p-values (with ties) against this null
    t0 = time.perf_counter()
    obs = plan.observed_maps()
    Pv, pties = plan.pvalues([md], want_ties=True)
    pval_ms = (time.perf_counter() - t0) * 1e3
    plan.close()
    n_products = 2 if ("sign" in perms and st["N"] > 64) else 3
    print(f"# split_bf16: {dt*1e3:.3f} ms/step, contraction {kern_ms:.3f} ms", file=sys.stderr, flush=True)
    value = K * V / dt
    flops = 2.0 * N * p * K * V
    achieved = flops / (kern_ms * 1e-3) / 1e12
    traffic, traffic_src = profiled_traffic(args.workload) if args.scale == 1.0 else (None, None)
    # the data pass read once (4 bytes per subject and voxel), the FP64 norms,
    # the max-T array
    algo_bytes = 4.0 * N * V + 8.0 * V + 4.0 * K
    res = {
        "metric": METRIC, "value": value, "unit": "perm·voxels/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16", "accumulate": "fp32", "data": "synthetic",
        "config": {"workload": (WORKLOADS[args.workload] if args.contrast == "t" else WORKLOAD_5F) +
                   ("" if args.scale == 1.0 else f" (grid scale {args.scale})"),
                   "V_mask": V, "N": N, "p": p, "K": K, "precision": "split_bf16"},
        "phase_breakdown_ms": {"transform": tm.transform_ms, "operand": tm.operand_ms,
                               "contract": tm.contract_ms, "cluster": tm.cluster_ms,
                               "library_call": tm.phase_ms},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_BF16_TFLOPS, "traffic": traffic, "traffic_source": traffic_src,
                     "algorithmic_bytes_per_launch": algo_bytes,
                     "kernel": "k_contract_bf16", "algorithmic_flop_per_launch": flops,
                     "avg_launch_ms": kern_ms,
                     "mfma_pipe_busy_est": mfma_busy(st, K, kern_ms, tm.slots_per_perm),
                     # the split form issues 3 BF16 products per FP32-accurate
                     # product (2 for +-1 operands above 64 subjects): its own
                     # ceiling
                     "split_form_peak": PEAK_BF16_TFLOPS / n_products,
                     "frac_of_split_form_peak": achieved / (PEAK_BF16_TFLOPS / n_products)},
        "fit_columns": tm.fit_columns,
        "precision_used": {capi.PREC_SPLIT_BF16: "split_bf16", capi.PREC_FP32: "fp32",
                           capi.PREC_FP64: "fp64"}[tm.precision_used],
        "ill_conditioned_waves": tm.unsafe_waves,
        "permutations_repeated_in_fp32": tm.fp32_permutations,
        "permutations_repeated_in_fp64": tm.fp64_permutations,
        "observed_and_pvalues_ms": pval_ms,
        "autotuner": run_form.tuned,
    }
    md32 = None
    if not args.no_fp32:
        nc = max(2, min(3 if args.workload == 5 else 10, args.steps // 2))  # steps of the comparison forms
        plan, md32, dt32, kern32, tm32 = run_form(ctx, st, perms, "fp32", nc, 1)
        plan.close()
        print(f"# fp32: {dt32*1e3:.3f} ms/step, contraction {kern32:.3f} ms", file=sys.stderr, flush=True)
        a32 = flops / (kern32 * 1e-3) / 1e12
        res["fp32_form"] = {"value": K * V / dt32, "ms_per_step": dt32 * 1e3, "contract_ms": kern32,
                            "achieved_tflops": a32, "frac_of_fp32_peak": a32 / PEAK_FP32_TFLOPS}
        res["fp32_form"]["kernel"] = ("k_contract_f32mx (FP32 matrix cores)" if tm32.fit_columns <= 16
                                      else "k_contract_f32 (vector ALUs)")
        res["speedup_vs_fp32"] = dt32 / dt
        res["maxT_equal_fp32_vs_bf16"] = int(np.sum(md32 == md))
        # the FP32 form on the vector ALUs as well (the plain per-voxel loop)
        plan, mdv, dtv, kernv, _ = run_form(ctx, st, perms, "fp32_valu", nc, 1)
        plan.close()
        print(f"# fp32 (vector ALUs): {dtv*1e3:.3f} ms/step, contraction {kernv:.3f} ms", file=sys.stderr, flush=True)
        av = flops / (kernv * 1e-3) / 1e12
        res["fp32_valu_form"] = {"value": K * V / dtv, "ms_per_step": dtv * 1e3, "contract_ms": kernv,
                                 "achieved_tflops": av, "frac_of_fp32_peak": av / PEAK_FP32_TFLOPS,
                                 "maxT_equal_to_fp32_form": int(np.sum(mdv == md32))}
        res["speedup_vs_fp32_valu"] = dtv / dt
        if args.workload == 5:
            # the split-FP8 form (e4m3 terms, E8M0 scales per 32 subjects)
            plan, md8, dt8, kern8, tm8 = run_form(ctx, st, perms, "split_fp8", nc, 1)
            plan.close()
            print(f"# split_fp8: {dt8*1e3:.3f} ms/step, contraction {kern8:.3f} ms", file=sys.stderr, flush=True)
            a8 = flops / (kern8 * 1e-3) / 1e12
            res["fp8_form"] = {"value": K * V / dt8, "ms_per_step": dt8 * 1e3, "contract_ms": kern8,
                               "achieved_tflops": a8, "frac_of_fp8_peak": a8 / 5000.0,
                               "permutations_repeated_in_fp32": tm8.fp32_permutations,
                               "maxT_equal_to_split_bf16_form": int(np.sum(md8 == md))}
            res["speedup_vs_fp8"] = dt8 / dt
    if not args.no_accuracy:
        forms = ("split_bf16", "fp32", "split_fp8") if args.workload == 5 else ("split_bf16", "fp32")
        if st.get("inference", capi.VOXEL) in (capi.CLUSTER_EXTENT, capi.CLUSTER_MASS):
            res["accuracy_vs_fp64"] = accuracy_cluster(ctx, st, perms, 8, forms)
        else:
            res["accuracy_vs_fp64"] = accuracy(ctx, st, perms, 8, forms)
        if args.workload in (1, 2):
            full = full_k_check(st, perms, K, md, md32, obs, Pv, pties, host_cores())
            res["accuracy_vs_fp64"]["full_K"] = full
            res["maxT_quanta_differing"] = full["split_bf16"]["maxT_quanta_differing"]
            res["maxT_outside_band"] = full["split_bf16"]["maxT_outside_band"]
            res["pvalue_ties"] = full["pvalues"]["ties"]
            res["pvalue_counts_differing_outside_ties"] = full["pvalues"]["counts_differing_outside_ties"]
            print(f"# full K vs FP64: {full}", file=sys.stderr, flush=True)
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(st, perms, args.cpu_seconds)
    ctx.close()
    if args.copy_mib:
        copy_calibration(args.copy_mib)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
