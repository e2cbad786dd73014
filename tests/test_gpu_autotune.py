"""Autotuner over group size x walk kernel variant: the choice is one of the candidates, is recorded
with the distribution statistics it is keyed by, is cached; every group size gives the same fields
to within the walk's truncation error (checked against FP64 direct summation) and every kernel
variant the same fields (bit for bit for the occupancy variant, to FP32 rounding for the shorter
interaction lists); group_size="auto" in the frontend runs the chosen configuration."""

import numpy as np
import pytest

from gpu_cases import case
from pytreegrav_amd import Accel, engine
from pytreegrav_amd._lib import PTG_ACC, PTG_POT

pytestmark = pytest.mark.gpu


def _rms_rel(a, b):
    return float(np.sqrt(np.mean(np.sum((a - b) ** 2, 1)) / np.mean(np.sum(b * b, 1))))


def test_autotune_choice_recorded_and_results_invariant():
    x, m, h = case("plummer_1e5_h")
    t = engine.DeviceTree(x, m, h, quadrupole=True)
    rec = engine.autotune_group_size(t, x, h, t.order, PTG_POT | PTG_ACC, quadrupole=True, theta=0.7,
                                     precision="float32", force=True)
    print(f"[autotune] {rec}", flush=True)
    assert rec["choice"] in engine.TUNE_SIZES and 0 <= rec["variant"] < len(engine.WALK_VARIANTS)
    # every group size, and every kernel variant for the one-target-per-lane sizes
    want = {f"{gs}/{v}" for gs in engine.TUNE_SIZES
            for v in (range(len(engine.WALK_VARIANTS)) if gs <= 64 else (0,))}
    assert set(rec["ms"]) == want and all(v > 0 for v in rec["ms"].values())
    assert set(rec["distribution"]) == {"max_level", "nodes_per_particle_x8", "target_spread_log2"}
    assert rec in engine.tuning_log().values()
    again = engine.autotune_group_size(t, x, h, t.order, PTG_POT | PTG_ACC, quadrupole=True, theta=0.7,
                                       precision="float32")
    assert again is rec  # cached
    idx = np.random.default_rng(0).choice(len(x), 2000, replace=False)
    d = engine.direct(x[idx], h[idx], x, m, h, PTG_ACC).cpu().numpy()
    # r == 0 pairs are skipped by the target form, so the sampled targets' own mass is excluded as in
    # the self walk
    for gs in engine.TUNE_SIZES:
        a = engine.walk(t, x, h, t.order, PTG_ACC, quadrupole=True, group_size=gs, theta=0.7,
                        precision="float32").cpu().numpy()
        assert _rms_rel(a[idx], d) < 2e-3, gs
    # the kernel variants change occupancy and list lengths, never the sums: the occupancy variant is
    # bit-identical; shorter lists interleave the node and particle list evaluations differently, so
    # the FP32 additions come in another order (differences at the rounding level)
    for gs in (32, 64):
        for expansion in (None, 1e-4):
            outs = [engine.walk(t, x, h, t.order, PTG_POT | PTG_ACC, group_size=gs, theta=0.7, precision="float32",
                                expansion=expansion, variant=v).cpu().numpy() for v in range(len(engine.WALK_VARIANTS))]
            assert np.array_equal(outs[1], outs[0]), (gs, expansion)
            rel = np.abs(outs[2] - outs[0]).max(0) / np.abs(outs[0]).max(0)
            assert rel.max() < 1e-5, (gs, expansion, rel)


def test_frontend_auto_group_size_runs_the_choice():
    x, m, h = case("uniform_1e4")
    a = Accel(x, m, h, theta=0.7, device="cuda", group_size="auto", method="tree")
    key_rec = [r for r in engine.tuning_log().values()]
    assert key_rec, "autotuning left no record"
    t = engine.DeviceTree(x, m, h)
    rec = engine.autotune_group_size(t, x, h, t.order, PTG_ACC, theta=0.7, precision="float32")
    b = engine.walk(t, x, h, t.order, PTG_ACC, group_size=rec["choice"], theta=0.7, precision="float32",
                    variant=rec["variant"]).cpu().numpy()
    assert np.array_equal(a, b)
