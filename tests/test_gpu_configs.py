"""The benchmark workloads at their stated sizes, each against the oracle or the FP64 direct sum.

C1  N=1e5 uniform in the unit cube, m=1/N, h=0.01, FP64 Accel+Potential through the public API,
    theta=0.7, monopole and quadrupole: equal to the oracle's grouped walk (group_size 8, the
    reference default, frontend.py:432) to 1e-10 with equal interaction counts; error against the
    FP64 direct sum on 1e4 targets.
C2  Plummer N=1e6 (truncated at 100 a), h=0.01, theta in {0.4, 0.7, 1.0}: the FP64 walk of every
    target in one launch equals the oracle on group-aligned Morton slices (1e-10, equal counters);
    the FP32 walk tracks the FP64 walk within the reference's FP32 tolerances
    (tests/test_cuda_gravity.py:83-95); the tidal tensor (grouped_treewalk.py:332-347) likewise; and
    the errors against the FP64 direct sum on 1e4 targets fall with theta.
C3  Plummer N=1e7 sources, 1e6 targets uniform in [-2, 2]^3: AccelTarget / PotentialTarget in FP32
    with group_size="auto" (frontend.py:563-685, grouped_treewalk.py:269-275) judged by the FP64
    direct sum on 1e4 targets; the target Morton order equals the oracle's, and the FP64 external-
    target walk equals the oracle on Morton slices.
C4  512 Plummer clusters (N log-uniform in [1e3, 1e5]) in one forest: per-system invariants, 16
    systems against the FP64 direct sum, three systems bit-equal to their own single-system walk,
    and HIP-graph replays (the walk stage, and the whole step from the tree build on) equal to the
    eager result.
C5  Learned criterion + group expansion against the plain theta=0.7 FP32 walk at N=1e7 Plummer,
    both judged by the FP64 direct sum on 1e4 targets.

The FP64 direct sum (k_direct) is itself pinned against the oracle's brute force in
test_gpu_direct.py."""

import numpy as np
import pytest
import torch

import oracle
from pytreegrav_amd import engine
from pytreegrav_amd._lib import PTG_ACC, PTG_POT, PTG_TIDAL
from pytreegrav_amd.sampling import plummer_clusters, plummer_truncated

pytestmark = pytest.mark.gpu


def _relmax(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def _acc_err(a, ref):
    """RMS and max |da| relative to the RMS field (examples/error_benchmark.py:75-76, 95-98)."""
    da = np.sqrt(np.sum((a - ref) ** 2, 1))
    s = np.sqrt(np.mean(np.sum(ref ** 2, 1)))
    return float(np.sqrt(np.mean(da ** 2)) / s), float(da.max() / s)


def _pot_err(p, ref):
    """std(dphi) / std(phi) and the largest deviation of dphi from its mean (error_benchmark.py)."""
    d = p - ref
    return float(np.std(d) / np.std(ref)), float(np.abs(d - d.mean()).max() / np.std(ref))


def _tidal_err(t, ref):
    d = np.sqrt(np.sum((t - ref).reshape(len(t), -1) ** 2, 1))
    return float(np.sqrt(np.mean(d ** 2)) / np.sqrt(np.mean(np.sum(ref.reshape(len(t), -1) ** 2, 1))))


def _direct(xt, ht, xs, ms, hs, mask):
    return engine.direct(xt, ht, xs, ms, hs, mask, self_mode=False, precision="float64").cpu().numpy()


def _slices(n, width, count, align):
    """count evenly spaced [start, start + width) ranges of [0, n), starts aligned to align."""
    return [((k * (n - width)) // max(count - 1, 1)) // align * align for k in range(count)]


# ------------------------------------------------------------------------------------------------
# C1
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("quadrupole", [False, True])
def test_c1_uniform_1e5_fp64_api(quadrupole):
    from pytreegrav_amd import Accel, Potential

    n = 100_000
    rng = np.random.default_rng(2024)
    x = rng.random((n, 3))
    m = np.full(n, 1.0 / n)
    h = np.full(n, 0.01)
    print(f"[C1] uniform n={n} quad={quadrupole}", flush=True)
    a = Accel(x, m, h, theta=0.7, quadrupole=quadrupole)
    p = Potential(x, m, h, theta=0.7, quadrupole=quadrupole)
    ref, rctr = oracle.self_fields_tree(x, m, h, potential=True, accel=True, quadrupole=quadrupole, group_size=8,
                                        theta=0.7, return_counters=True)
    assert _relmax(a, ref["accel"]) < 1e-10 and _relmax(p, ref["potential"]) < 1e-10
    # the same walk through the engine, with the kernel's own counters
    from pytreegrav_amd import ConstructTree

    t = ConstructTree(x, m, h, quadrupole=quadrupole)
    ctr = torch.zeros(4, dtype=torch.int64, device=engine.device())
    engine.walk(t.dev, x, h, t.dev.order, PTG_POT | PTG_ACC, quadrupole=quadrupole, group_size=8, theta=0.7,
                precision="float64", counters=ctr)
    assert int(ctr[0]) == rctr["interactions"] and int(ctr[1]) == rctr["visits"]
    # accuracy on 1e4 targets against the FP64 direct sum (the self term is excluded by r == 0)
    sel = np.sort(rng.choice(n, 10_000, replace=False))
    d = _direct(x[sel], h[sel], x, m, h, PTG_POT | PTG_ACC)
    ea, _ = _acc_err(a[sel], d[:, 1:4])
    ep, _ = _pot_err(p[sel], d[:, 0])
    print(f"[C1] quad={quadrupole} accel rms {ea:.3e} pot rms {ep:.3e}", flush=True)
    # monopole theta=0.7 on this input: accel 2.0e-3, potential 1.1e-3 of the RMS (std) field
    assert ea < (1.5e-3 if quadrupole else 3e-3) and ep < (1e-3 if quadrupole else 2e-3)


# ------------------------------------------------------------------------------------------------
# C2
# ------------------------------------------------------------------------------------------------
_C2 = {}


def _c2_inputs():
    if "x" not in _C2:
        x, m = plummer_truncated(1_000_000, seed=22)
        h = np.full(len(m), 0.01)
        _C2.update(x=x, m=m, h=h)
        _C2["otree"] = oracle.build_tree(x, m, h)
        rng = np.random.default_rng(23)
        _C2["sel"] = np.sort(rng.choice(len(x), 10_000, replace=False))
        _C2["direct"] = _direct(x[_C2["sel"]], h[_C2["sel"]], x, m, h, PTG_POT | PTG_ACC | PTG_TIDAL)
    return _C2


@pytest.mark.parametrize("theta", [0.4, 0.7, 1.0])
def test_c2_plummer_1e6_theta_sweep(theta):
    from pytreegrav_amd import ConstructTree

    c = _c2_inputs()
    x, m, h, ot = c["x"], c["m"], c["h"], c["otree"]
    n = len(x)
    mask = PTG_POT | PTG_ACC | PTG_TIDAL
    print(f"[C2] plummer n={n} theta={theta}", flush=True)
    t = ConstructTree(x, m, h)
    idx = t.dev.order.cpu().numpy()
    assert np.array_equal(idx, ot.TreewalkIndices)
    # FP64: every target in one launch (groups of 8 consecutive Morton targets) ...
    w64 = engine.walk(t.dev, x, h, t.dev.order, mask, group_size=8, theta=theta, precision="float64")
    w64 = w64.cpu().numpy()
    xs, hs = np.ascontiguousarray(x[idx]), np.ascontiguousarray(h[idx])
    dev_ctr = torch.zeros(4, dtype=torch.int64, device=engine.device())
    for s in _slices(n, 4096, 4, 8):
        sl = slice(s, s + 4096)
        ref, rctr = oracle.fields_grouped(ot, xs[sl], hs[sl], potential=True, accel=True, tidal=True,
                                          group_size=8, theta=theta, return_counters=True)
        got = w64[idx[sl]]
        assert _relmax(got[:, 0], ref["potential"]) < 1e-10, s
        assert _relmax(got[:, 1:4], ref["accel"]) < 1e-10, s
        assert _relmax(got[:, 4:].reshape(-1, 3, 3), ref["tidal"]) < 1e-10, s
        # ... and the same slice walked alone: identical groups, so identical counters
        dev_ctr.zero_()
        engine.walk(t.dev, xs[sl], hs[sl], torch.arange(4096, device=engine.device()), mask, group_size=8,
                    theta=theta, precision="float64", counters=dev_ctr)
        assert int(dev_ctr[0]) == rctr["interactions"] and int(dev_ctr[1]) == rctr["visits"], s
    # FP32 against FP64 (same groups of 64): the reference's FP32 tolerances
    g64 = engine.walk(t.dev, x, h, t.dev.order, mask, group_size=64, theta=theta, precision="float64",
                      grouping="fixed").cpu().numpy()
    g32 = engine.walk(t.dev, x, h, t.dev.order, mask, group_size=64, theta=theta, precision="float32",
                      grouping="fixed").cpu().numpy()
    wp = g64[:, 0]
    assert np.sqrt(np.mean((g32[:, 0] - wp) ** 2)) / wp.std() < 1e-5
    rel = np.sqrt(((g32[:, 1:4] - g64[:, 1:4]) ** 2).sum(1)) / np.sqrt((g64[:, 1:4] ** 2).sum(1))
    assert np.median(rel) < 1e-5 and rel.max() < 1e-2
    tr = np.sqrt(((g32[:, 4:] - g64[:, 4:]) ** 2).sum(1)) / np.sqrt((g64[:, 4:] ** 2).sum(1))
    assert np.median(tr) < 1e-4
    # accuracy against the FP64 direct sum on 1e4 targets
    sel, d = c["sel"], c["direct"]
    errs = {}
    for name, w in (("fp64_g8", w64), ("fp64_g64", g64), ("fp32_g64", g32)):
        ea, ema = _acc_err(w[sel, 1:4], d[:, 1:4])
        ep, _ = _pot_err(w[sel, 0], d[:, 0])
        et = _tidal_err(w[sel, 4:], d[:, 4:])
        errs[name] = (ea, ema, ep, et)
        print(f"[C2] theta={theta} {name}: accel rms {ea:.3e} max {ema:.3e} pot {ep:.3e} tidal {et:.3e}", flush=True)
    _C2.setdefault("errs", {})[theta] = errs
    # the published accuracy of the reference's grouped walk (README.md:122-125, N=1e5 untruncated,
    # h=0): RMS 4.8e-4, 1.8e-3, 4.8e-3 at theta 0.4 / 0.7 / 1.0; this N=1e6 truncated, softened sphere
    # stays within 25 % of it (measured 2.7e-4, 1.3e-3, 5.1e-3 with groups of 8)
    bound = 1.25 * {0.4: 4.8e-4, 0.7: 1.8e-3, 1.0: 4.8e-3}[theta]
    for name in errs:
        assert errs[name][0] < bound, (name, errs[name])
        assert errs[name][3] < 10 * bound, (name, errs[name])
    # symmetric tidal tensor, as the walk writes it
    T = w64.reshape(-1, 1 + 3 + 9)[:, 4:].reshape(-1, 3, 3)
    assert np.array_equal(T, np.transpose(T, (0, 2, 1)))


def test_c2_errors_fall_with_theta():
    c = _c2_inputs()
    errs = c.get("errs", {})
    if len(errs) < 3:
        pytest.skip("needs the three theta cases of test_c2_plummer_1e6_theta_sweep in this session")
    for name in errs[0.7]:
        e = [errs[t][name] for t in (0.4, 0.7, 1.0)]
        assert e[0][0] < e[1][0] < e[2][0], (name, e)
        assert e[0][3] < e[1][3] < e[2][3], (name, e)


# ------------------------------------------------------------------------------------------------
# C3
# ------------------------------------------------------------------------------------------------
_C3 = {}


def _c3_inputs():
    if "x" not in _C3:
        x, m = plummer_truncated(10_000_000, seed=31)
        h = np.full(len(m), 0.01)
        rng = np.random.default_rng(32)
        xt = rng.uniform(-2.0, 2.0, size=(1_000_000, 3))
        _C3.update(x=x, m=m, h=h, xt=xt)
    return _C3


def test_c3_targets_1e7_sources_1e6_targets_fp32_auto():
    from pytreegrav_amd import AccelTarget, ConstructTree, PotentialTarget

    c = _c3_inputs()
    x, m, h, xt = c["x"], c["m"], c["h"], c["xt"]
    print(f"[C3] sources {len(x)} targets {len(xt)}", flush=True)
    dev = engine.device()
    xd, md, hd, xtd = (torch.from_numpy(a).to(dev) for a in (x, m, h, xt))
    tree = ConstructTree(xd, md, hd)
    engine._tuned.clear()
    a = AccelTarget(xtd, None, None, theta=0.7, tree=tree, group_size="auto", precision="float32")
    p = PotentialTarget(xtd, None, None, theta=0.7, tree=tree, group_size="auto", precision="float32")
    assert isinstance(a, torch.Tensor) and a.is_cuda and tuple(a.shape) == (len(xt), 3)
    log = engine.tuning_log()
    assert len(log) >= 1
    for rec in log.values():
        assert rec["choice"] in engine.TUNE_SIZES and 0 <= rec["variant"] < len(engine.WALK_VARIANTS)
        assert {int(k.split("/")[0]) for k in rec["ms"]} == set(engine.TUNE_SIZES)
    print(f"[C3] autotune {[(r['choice'], r['variant_name'], r['ms']) for r in log.values()]}", flush=True)
    rng = np.random.default_rng(33)
    sel = np.sort(rng.choice(len(xt), 10_000, replace=False))
    d = _direct(xt[sel], np.zeros(len(sel)), x, m, h, PTG_POT | PTG_ACC)
    ea, ema = _acc_err(a.cpu().numpy()[sel], d[:, 1:4])
    ep, epm = _pot_err(p.cpu().numpy()[sel], d[:, 0])
    print(f"[C3] accel rms {ea:.3e} max {ema:.3e}; pot rms {ep:.3e} max {epm:.3e}", flush=True)
    assert ea < 1.8e-3 and ep < 1.8e-3
    # the group size does not change the answer beyond the truncation error
    a64 = AccelTarget(xtd, None, None, theta=0.7, tree=tree, group_size=64, precision="float32").cpu().numpy()
    ea64, _ = _acc_err(a64[sel], d[:, 1:4])
    assert abs(ea - ea64) < 0.5 * max(ea, ea64)


def test_c3_target_order_and_fp64_slices_match_oracle():
    from pytreegrav_amd import ConstructTree

    c = _c3_inputs()
    x, m, h, xt = c["x"], c["m"], c["h"], c["xt"]
    print("[C3] fp64 slices", flush=True)
    order = engine.morton_order(xt).cpu().numpy()
    assert np.array_equal(order, oracle.morton_order(xt))
    ot = oracle.build_tree(x, m, h)
    t = ConstructTree(x, m, h)
    assert np.array_equal(t.dev.order.cpu().numpy(), ot.TreewalkIndices)
    xs = np.ascontiguousarray(xt[order])
    hs = np.zeros(len(xs))
    ctr = torch.zeros(4, dtype=torch.int64, device=engine.device())
    for s in _slices(len(xs), 2048, 3, 8):
        sl = slice(s, s + 2048)
        ctr.zero_()
        got = engine.walk(t.dev, xs[sl], hs[sl], torch.arange(2048, device=engine.device()), PTG_POT | PTG_ACC,
                          group_size=8, theta=0.7, precision="float64", counters=ctr).cpu().numpy()
        ref, rctr = oracle.fields_grouped(ot, xs[sl], hs[sl], potential=True, accel=True, group_size=8, theta=0.7,
                                          return_counters=True)
        assert int(ctr[0]) == rctr["interactions"] and int(ctr[1]) == rctr["visits"], s
        assert _relmax(got[:, 0], ref["potential"]) < 1e-10 and _relmax(got[:, 1:4], ref["accel"]) < 1e-10, s


# ------------------------------------------------------------------------------------------------
# C4
# ------------------------------------------------------------------------------------------------
def test_c4_512_clusters_forest():
    from pytreegrav_amd import Forest

    systems = plummer_clusters(512, seed=0)
    x = np.vstack([s[0] for s in systems])
    m = np.concatenate([s[1] for s in systems])
    h = np.concatenate([s[2] for s in systems])
    off = np.concatenate([[0], np.cumsum([len(s[0]) for s in systems])])
    n = len(x)
    print(f"[C4] systems {len(systems)} n {n}", flush=True)
    assert np.diff(off).min() >= 1000 and np.diff(off).max() <= 100_000
    kw = dict(alpha=1.6e-3, expansion=1e-3)
    F = Forest(x, m, h, off, precision="float32", criterion="learned")
    r = F.evaluate(potential=True, accel=True, **kw)
    acc, pot = r["accel"].cpu().numpy(), r["potential"].cpu().numpy()
    assert np.isfinite(acc).all() and np.isfinite(pot).all()
    # per-system invariants: one root per system, systems' particle ranges hold their own Morton order
    t = F.tree
    roots = t.roots.cpu().numpy()
    assert np.all(np.diff(roots) > 0) and roots[0] == n
    order = t.order.cpu().numpy()
    for s in range(len(systems)):
        o = order[off[s]:off[s + 1]]
        assert o.min() == off[s] and o.max() == off[s + 1] - 1
    masses = t.masses.cpu().numpy()
    msum = np.add.reduceat(m, off[:-1])
    assert np.allclose(masses[roots], msum, rtol=1e-12)
    # the potential is negative everywhere (positive masses) and the net force of each system is small
    assert np.all(pot < 0)
    # 16 systems against the FP64 direct sum within the system
    rng = np.random.default_rng(7)
    errs = []
    for s in rng.choice(len(systems), 16, replace=False):
        xs_, ms_, hs_ = systems[s]
        a, b = off[s], off[s + 1]
        idx = np.sort(rng.choice(b - a, min(1000, b - a), replace=False))
        d = _direct(xs_[idx], hs_[idx], xs_, ms_, hs_, PTG_ACC)
        errs.append(_acc_err(acc[a + idx], d)[0])
    print(f"[C4] per-system accel rms median {np.median(errs):.3e} max {max(errs):.3e}", flush=True)
    assert max(errs) < 3e-3
    # three systems bit-equal to their own single-system walk (geometric criterion, no expansion)
    P = Forest(x, m, h, off, precision="float32")
    rp = P.evaluate(potential=True, accel=True)
    fa, fp = rp["accel"].cpu().numpy(), rp["potential"].cpu().numpy()
    for s in (0, 255, 511):
        xs_, ms_, hs_ = systems[s]
        u = engine.DeviceTree(xs_, ms_, hs_)
        o = engine.walk(u, xs_, hs_, u.order, PTG_POT | PTG_ACC, group_size=64, theta=0.7,
                        precision="float32").cpu().numpy()
        assert np.array_equal(fp[off[s]:off[s + 1]], o[:, 0]), s
        assert np.array_equal(fa[off[s]:off[s + 1]], o[:, 1:4]), s
    # graph replays equal the eager result, with eager work between them: the walk stage alone, and
    # the whole step (deferred build, groups, field scales, radii, walk) in one graph
    ref = torch.cat([r["potential"][:, None], r["accel"]], 1).clone()
    replay, out = F.capture(potential=True, accel=True, **kw)
    replay_step, out_step = F.capture_step(potential=True, accel=True, **kw)
    for k in range(3):
        out.zero_()
        replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), k
        out_step.zero_()
        replay_step()
        torch.cuda.synchronize()
        assert torch.equal(out_step, ref), k
        P.evaluate(potential=True, accel=True)


# ------------------------------------------------------------------------------------------------
# C5
# ------------------------------------------------------------------------------------------------
def test_c5_learned_vs_theta07_1e7():
    from pytreegrav_amd import ConstructTree
    from pytreegrav_amd.learned import OpeningMLP

    x, m = plummer_truncated(10_000_000, seed=0)
    h = np.full(len(m), 0.01)
    dev = engine.device()
    xd, md, hd = (torch.from_numpy(a).to(dev) for a in (x, m, h))
    print("[C5] plummer 1e7 learned vs theta 0.7", flush=True)
    rng = np.random.default_rng(12345)
    sel = np.sort(rng.choice(len(x), 10_000, replace=False))
    d = _direct(x[sel], h[sel], x, m, h, PTG_POT | PTG_ACC)
    mask = PTG_POT | PTG_ACC
    model = OpeningMLP.load()
    res = {}
    for name, kw in (("theta0.7", dict(criterion="theta", theta=0.7)),
                     ("learned", dict(criterion="learned", alpha=5e-4, model=model, expansion=4e-3, expansion_rel=3.0)),
                     ("learned_uncapped", dict(criterion="learned", alpha=5e-4, model=model, expansion=4e-3))):
        tree = ConstructTree(xd, md, hd, quadrupole=name != "theta0.7")
        ctr = torch.zeros(4, dtype=torch.int64, device=dev)
        out = engine.self_fields(tree.dev, xd, hd, mask, group_size=64, counters=ctr, **kw).cpu().numpy()
        ea, ema = _acc_err(out[sel, 1:4], d[:, 1:4])
        ep, epm = _pot_err(out[sel, 0], d[:, 0])
        res[name] = dict(ea=ea, ema=ema, ep=ep, epm=epm, inter=int(ctr[0]), visits=int(ctr[1]))
        print(f"[C5] {name}: {res[name]}", flush=True)
    L, P, U = res["learned"], res["theta0.7"], res["learned_uncapped"]
    # the benchmark configuration: acceleration RMS and max and potential RMS errors at most the
    # theta=0.7 walk's
    assert L["ea"] <= 1.8e-3 and L["ea"] <= P["ea"] and L["ema"] <= P["ema"] and L["ep"] <= P["ep"]
    # the learned walk evaluates far fewer per-target interactions and visits fewer elements
    assert L["inter"] < 0.3 * P["inter"] and L["visits"] < P["visits"]
    # without the relative cap the expansion folds more (fewer interactions) at about the same RMS
    # error, but its worst error is larger
    assert U["inter"] < L["inter"] and U["ea"] <= 1.05 * P["ea"] and U["ema"] > L["ema"]
