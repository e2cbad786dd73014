"""Group expansion of the FP32 walk (no reference counterpart): far accepted elements folded into a
third-order Taylor expansion about each group's centre.

* A point mass far from one group pins the kernel's expansion to the closed-form field: the
  difference from FP64 direct summation stays within the truncation bound of the dropped
  fourth- and higher-order terms ((M/D^2) x^3 (4 - 3x) / (1 - x)^2 for the acceleration,
  (M/D) x^4 / (1 - x) for the potential, x = R/D), well below what a missing or wrong third-order
  term would leave ((R/D)^2 relative).
* A tolerance too small to fold anything (groups of nonzero radius) gives the plain walk bit for bit.
* On Plummer spheres the folded walk stays within the tolerance's accuracy against FP64 direct
  summation, with the fold counter reporting the folded element-group pairs."""

import numpy as np
import pytest
import torch

from gpu_cases import case
from pytreegrav_amd import ConstructTree, engine
from pytreegrav_amd._lib import PTG_ACC, PTG_POT, PtgError
from pytreegrav_amd.sampling import plummer_truncated

pytestmark = pytest.mark.gpu


def _direct64(xt, xs, ms):
    d = xs[None, :, :] - xt[:, None, :]
    r = np.sqrt(np.sum(d * d, -1))
    pot = -np.sum(ms[None, :] / r, 1)
    acc = np.sum(ms[None, :, None] * d / r[..., None] ** 3, 1)
    return pot, acc


@pytest.mark.parametrize("D", [200.0, 60.0])
def test_far_point_mass_matches_closed_form_within_bound(D):
    rng = np.random.default_rng(3)
    ns = 500
    centre = np.array([0.6, 0.8, 0.0]) * D
    u = rng.normal(size=(ns, 3))
    xs = centre + 1e-3 * u / np.linalg.norm(u, axis=1, keepdims=True) * rng.uniform(0, 1, (ns, 1))
    ms = np.full(ns, 1.0 / ns)
    hs = np.full(ns, 0.01)
    xt = rng.uniform(-1, 1, (64, 3))
    ht = np.full(64, 0.01)
    print(f"[expansion] far point mass D={D}", flush=True)
    t = ConstructTree(xs, ms, hs)
    perm = engine.morton_order(xt)
    ctr = torch.zeros(4, dtype=torch.int64, device=engine.device())
    out = engine.walk(t.dev, xt, ht, perm, PTG_POT | PTG_ACC, group_size=64, theta=0.7, precision="float32",
                      grouping="fixed", counters=ctr, expansion=1e-3).cpu().numpy()
    c = ctr.cpu().numpy()
    assert c[3] == 1 and c[0] == 0, c  # the root is accepted and folded; nothing else is evaluated
    pot, acc = _direct64(xt, xs, ms)
    R = np.sqrt(np.sum(((xt.max(0) - xt.min(0)) / 2) ** 2))
    Dg = np.linalg.norm(centre - (xt.max(0) + xt.min(0)) / 2)
    xr = R / Dg
    bound_a = xr ** 3 * (4 - 3 * xr) / (1 - xr) ** 2 / Dg ** 2
    bound_p = xr ** 4 / (1 - xr) / Dg
    fp32 = 1e-6 / Dg ** 2  # rounding of the FP32 node-relative coordinates and sums
    da = np.sqrt(np.sum((out[:, 1:4] - acc) ** 2, 1))
    assert da.max() < bound_a + fp32, (da.max(), bound_a)
    assert np.abs(out[:, 0] - pot).max() < bound_p + 1e-6 / Dg, (np.abs(out[:, 0] - pot).max(), bound_p)
    # the third-order terms matter at this distance: a second-order expansion would miss by more
    second = (R / Dg) ** 2 / Dg ** 2
    assert bound_a < 0.25 * second


def test_tiny_tolerance_equals_plain_walk():
    x, m, h = case("plummer_1e5_h")
    t = ConstructTree(x, m, h)
    ctr = torch.zeros(4, dtype=torch.int64, device=engine.device())
    # runs of 64 distinct targets: every group has a nonzero radius (a one-target group would fold
    # exactly, at its own position, but with other rounding)
    kw = dict(group_size=64, theta=0.7, precision="float32", grouping="fixed")
    a = engine.walk(t.dev, x, h, t.dev.order, PTG_POT | PTG_ACC, **kw)
    b = engine.walk(t.dev, x, h, t.dev.order, PTG_POT | PTG_ACC, counters=ctr, expansion=1e-30, **kw)
    assert int(ctr[3]) == 0
    assert torch.equal(a, b)


def _rms(a, d):
    return np.sqrt(np.mean(np.sum((a - d) ** 2, 1)) / np.mean(np.sum(d * d, 1)))


@pytest.mark.parametrize("criterion,gs", [("theta", 64), ("learned", 64), ("learned", 128), ("theta", 256)])
def test_plummer_accuracy_with_expansion(criterion, gs):
    N = 200_000
    x, m = plummer_truncated(N, seed=5)
    h = np.full(N, 0.01)
    print(f"[expansion] plummer N={N} criterion={criterion} gs={gs}", flush=True)
    t = ConstructTree(x, m, h, quadrupole=True)
    rng = np.random.default_rng(0)
    sel = np.sort(rng.choice(N, 4000, replace=False))
    d = engine.direct(x[sel], h[sel], x, m, h, PTG_ACC).cpu().numpy()
    kw = dict(theta=0.7 if criterion == "theta" else 2.0, criterion=criterion, alpha=1.2e-3 if criterion != "theta"
              else None, group_size=gs)
    ctr0 = torch.zeros(4, dtype=torch.int64, device=engine.device())
    ctr1 = torch.zeros(4, dtype=torch.int64, device=engine.device())
    a0 = engine.self_fields(t.dev, x, h, PTG_POT | PTG_ACC, counters=ctr0, **kw).cpu().numpy()
    a1 = engine.self_fields(t.dev, x, h, PTG_POT | PTG_ACC, counters=ctr1, expansion=3e-4, **kw).cpu().numpy()
    e0, e1 = _rms(a0[sel, 1:4], d), _rms(a1[sel, 1:4], d)
    c0, c1 = ctr0.cpu().numpy(), ctr1.cpu().numpy()
    print(f"  rms {e0:.3e} -> {e1:.3e}; interactions/particle {c0[0] / N:.0f} -> {c1[0] / N:.0f}, "
          f"folds/group {c1[3] / (N / gs):.0f}", flush=True)
    assert c1[3] > 0 and c1[0] < 0.8 * c0[0]
    # the same accepted elements (counted per target or folded)
    assert c1[1] == c0[1]
    assert e1 < max(1.3 * e0, e0 + 5e-4), (e0, e1)


def test_expansion_argument_checks():
    x, m, h = case("uniform_1e4")
    t = ConstructTree(x, m, h, quadrupole=True)
    with pytest.raises(ValueError):
        engine.walk(t.dev, x, h, t.dev.order, PTG_ACC, precision="float64", expansion=1e-3)
    with pytest.raises(ValueError):
        engine.walk(t.dev, x, h, t.dev.order, PTG_ACC, precision="float32", quadrupole=True, expansion=1e-3)
    with pytest.raises(ValueError):
        engine.walk(t.dev, x, h, t.dev.order, 7, precision="float32", expansion=1e-3)
    with pytest.raises(ValueError):
        engine.walk(t.dev, x, h, t.dev.order, PTG_ACC, precision="float32", expansion=0.0)
    # the library checks too
    from pytreegrav_amd import _lib

    L = _lib.load()
    crit = _lib.Criterion(_lib.PTG_CRIT_THETA, 1, 0.0, 0.0, None)
    p = t.dev.pack(engine.PTG_F32, False)
    out = torch.empty((t.dev.n, 3), dtype=torch.float64, device=engine.device())
    ws = engine.workspace(L.ptg_walk_workspace_bytes(t.dev.n, 64))
    pos = engine.to_dev(x, (-1, 3))
    soft = engine.to_dev(h, (t.dev.n,))
    with pytest.raises(PtgError):
        engine.check(L.ptg_walk(_lib.ctypes.byref(p["wt"]), None, _lib.ptr(pos), _lib.ptr(soft),
                                _lib.ptr(t.dev.order), 0, t.dev.n, PTG_ACC, 0, 64, None, None, 0.7,
                                _lib.ctypes.byref(crit), 1.0, 0, _lib.ptr(out), None, _lib.ptr(ws), ws.numel(),
                                engine.stream()), "ptg_walk")


def _forest_systems():
    from test_gpu_forest import _concat, _systems

    return _concat(_systems()[:8]), _systems()[:8]


def test_forest_expansion_equals_single_walks():
    """Each system of a batched walk with the group expansion equals that system's own walk with the
    same tolerance (expansion x the system's RMS field from the forest's sample walk), bit for bit."""
    from pytreegrav_amd import Forest

    (x, m, h, off), systems = _forest_systems()
    F = Forest(x, m, h, off, precision="float32")
    r = F.evaluate(potential=True, accel=True, expansion=3e-4)
    fa, fp = r["accel"].cpu().numpy(), r["potential"].cpu().numpy()
    scale = F._exp_scale
    for s, (xs, ms, hs) in enumerate(systems):
        u = engine.DeviceTree(xs, ms, hs)
        o = engine.walk(u, xs, hs, u.order, PTG_POT | PTG_ACC, group_size=64, theta=0.7, precision="float32",
                        expansion=3e-4, eps_scale=scale[s:s + 1].contiguous()).cpu().numpy()
        assert np.array_equal(fp[off[s]:off[s + 1]], o[:, 0]), s
        assert np.array_equal(fa[off[s]:off[s + 1]], o[:, 1:4]), s


def test_forest_learned_expansion_accuracy_and_capture():
    from pytreegrav_amd import Forest

    (x, m, h, off), systems = _forest_systems()
    F = Forest(x, m, h, off, criterion="learned")
    ctr = torch.zeros(4, dtype=torch.int64, device=engine.device())
    a = F.evaluate(potential=True, accel=True, alpha=1.2e-3, expansion=1e-3, counters=ctr)
    assert int(ctr[3]) > 0
    acc = a["accel"].cpu().numpy()
    for s, (xs, ms, hs) in enumerate(systems[2:], start=2):
        d = engine.direct(xs, hs, xs, ms, hs, PTG_ACC, self_mode=True).cpu().numpy()
        assert _rms(acc[off[s]:off[s + 1]], d) < 4e-3, s
    replay, out = F.capture(potential=True, accel=True, alpha=1.2e-3, expansion=1e-3)
    ref = torch.cat([a["potential"][:, None], a["accel"]], 1).clone()
    for _ in range(3):
        out.zero_()
        replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref)


def test_frontend_expansion():
    from pytreegrav_amd import Accel, Potential

    N = 50_000
    x, m = plummer_truncated(N, seed=9)
    h = np.full(N, 0.01)
    dd = engine.direct(x, h, x, m, h, PTG_POT | PTG_ACC, self_mode=True).cpu().numpy()
    d, dp = dd[:, 1:4], dd[:, 0]
    a0 = Accel(x, m, h, device="cuda", precision="float32")
    a1 = Accel(x, m, h, device="cuda", precision="float32", expansion=3e-4)
    a2 = Accel(x, m, h, device="cuda", criterion="learned", alpha=1.2e-3, expansion=1e-3)
    e0, e1, e2 = _rms(a0, d), _rms(a1, d), _rms(a2, d)
    assert e1 < max(1.3 * e0, e0 + 5e-4) and e2 < 3e-3, (e0, e1, e2)
    # potential: error spread relative to the potential's spread, as the benchmark reports it
    p = Potential(x, m, h, device="cuda", precision="float32", expansion=3e-4)
    p0 = Potential(x, m, h, device="cuda", precision="float32")
    q0, q1 = np.std(p0 - dp) / np.std(dp), np.std(p - dp) / np.std(dp)
    print(f"  accel rms {e0:.2e} {e1:.2e} {e2:.2e}; potential rms {q0:.2e} -> {q1:.2e}", flush=True)
    assert q1 < max(1.3 * q0, q0 + 3e-4), (q0, q1)
    with pytest.raises(ValueError):
        Accel(x, m, h, device="cuda", precision="float64", expansion=3e-4)
    with pytest.raises(ValueError):
        Accel(x, m, h, device="cuda", criterion="learned", alpha=1.2e-3, expansion=1e-3, expansion_rel=-1.0)
    # expansion_rel=0 (no cap) folds at least as much as the default cap
    c_cap = torch.zeros(4, dtype=torch.int64, device=engine.device())
    c_free = torch.zeros(4, dtype=torch.int64, device=engine.device())
    t = engine.DeviceTree(x, m, h)
    engine.walk(t, x, h, t.order, PTG_ACC, group_size=64, theta=0.7, precision="float32", counters=c_cap,
                expansion=1e-3, expansion_rel=3.0)
    engine.walk(t, x, h, t.order, PTG_ACC, group_size=64, theta=0.7, precision="float32", counters=c_free,
                expansion=1e-3)
    assert 0 < int(c_cap[3]) <= int(c_free[3])


def test_autotune_with_expansion():
    x, m, h = case("plummer_1e5_h")
    t = ConstructTree(x, m, h)
    out = engine.walk(t.dev, x, h, t.dev.order, PTG_POT | PTG_ACC, group_size="auto", theta=0.7, precision="float32",
                      expansion=1e-4)
    log = [v for k, v in engine.tuning_log().items() if k[-2] == 1e-4]
    assert log and log[-1]["choice"] in engine.TUNE_SIZES
    d = engine.direct(x[:2000], h[:2000], x, m, h, PTG_ACC).cpu().numpy()
    assert _rms(out.cpu().numpy()[:2000, 1:4], d) < 3e-3
