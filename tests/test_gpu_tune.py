"""The autotuner (bpg_plan_tune, bpg_tune_save / bpg_tune_load): it picks
for each bucket the fastest of its candidate shapes on the device (so never
slower than the planner's default, measured in the same search), the chosen
shape changes nothing in the results, and the cache file is read back --
explicitly and through BPG_TUNE_CACHE at context creation."""
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from broccoli_amd import capi, synth

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.parametrize("wl,K", [(1, 256), (2, 300), (3, 200)])
def test_tuned_shape_is_no_slower_and_changes_no_result(wl, K, tmp_path):
    st = synth.study(wl, scale=0.3)
    perms = synth.permutations(st, K, seed=5)
    print(f"tuning workload {wl} V={st['V']} K={K}", flush=True)
    with capi.Context(0) as c0, capi.Plan(c0, st, "split_bf16") as plan:
        md0, _ = plan.permutation_phase(0, **perms)  # the planner's default shape
    with capi.Context(0) as ctx:
        with capi.Plan(ctx, st, "split_bf16") as plan:
            r = plan.tune(0, K)
            md, _ = plan.permutation_phase(0, **perms)
        print(f"  candidates {r.candidates}, chosen R={r.rows_per_wave} chunk x2^{r.chunk_pow} "
              f"({r.vt_per_chunk} tiles): {r.best_ms:.3f} ms vs default {r.default_ms:.3f} ms")
        assert r.candidates == (3 if wl == 3 else 6)  # cluster inference: no alternative subtile count
        assert 0 < r.best_ms <= r.default_ms
        assert np.array_equal(md, md0)
        assert ctx.tune_entries == 1
        ctx.tune_save(tmp_path / "tune.txt")
    text = (tmp_path / "tune.txt").read_text()
    assert text.startswith("#") and len(text.splitlines()) == 2
    with capi.Context(0) as c2:
        assert c2.tune_entries == 0
        assert c2.tune_load(tmp_path / "tune.txt") == 1
        assert c2.tune_entries == 1
        with capi.Plan(c2, st, "split_bf16") as plan:
            md2, _ = plan.permutation_phase(0, **perms)
        assert np.array_equal(md2, md0)


def test_cache_named_by_the_environment_is_read_at_context_creation(tmp_path):
    st = synth.study(2, scale=0.2)
    with capi.Context(0) as ctx, capi.Plan(ctx, st, "split_bf16") as plan:
        plan.tune(0, 128)
        ctx.tune_save(tmp_path / "cache.txt")
    code = ("from broccoli_amd import capi; c = capi.Context(0); print('entries', c.tune_entries); c.close()")
    env = dict(os.environ, BPG_TUNE_CACHE=str(tmp_path / "cache.txt"))
    res = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True, text=True,
                         timeout=120)
    print(res.stdout, res.stderr)
    assert "entries 1" in res.stdout
