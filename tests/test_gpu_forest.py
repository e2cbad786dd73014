"""Forest (batched independent systems): every system of a batched build and walk must equal the
single-system GPU path run on that system alone, bit for bit (tree arrays, FP32 and FP64 fields),
and its fields must match FP64 direct summation to the walk's accuracy.  The HIP-graph replay of the
batched walk stage must reproduce the eager result on every replay."""

import numpy as np
import pytest
import torch

from gpu_cases import case
from pytreegrav_amd import Forest, engine
from pytreegrav_amd._lib import PTG_ACC, PTG_POT, PtgError
from pytreegrav_amd.sampling import plummer_truncated

pytestmark = pytest.mark.gpu


def _systems():
    """Uneven systems, including the reference tests' collision cases (re-keying, bucket chains)."""
    out = []
    rng = np.random.default_rng(21)
    for k, n in enumerate([2, 3, 700, 4096, 257, 12000, 64, 1500]):
        x, m = plummer_truncated(n, seed=100 + k)
        a = 10 ** rng.uniform(-1, 1)
        out.append((x * a + rng.normal(size=3) * 5, m, np.full(n, 0.01 * a)))
    for name in ("dups", "rekey", "bucket_chain", "all_identical", "dups_origin", "clustered"):
        out.append(case(name))
    return out


def _concat(systems):
    x = np.vstack([s[0] for s in systems])
    m = np.concatenate([s[1] for s in systems])
    h = np.concatenate([s[2] for s in systems])
    off = np.concatenate([[0], np.cumsum([len(s[0]) for s in systems])])
    return x, m, h, off


def _node_ranges(t):
    roots = t.roots.cpu().numpy().astype(np.int64)
    ends = np.append(roots[1:], t.n + t.nnodes)
    return roots, ends


def _sub(t, off, s, name):
    """System s's rows of a forest array, particles then nodes, with links renumbered locally."""
    roots, ends = _node_ranges(t)
    a, b = off[s], off[s + 1]
    arr = getattr(t, name).cpu().numpy()
    rows = np.concatenate([arr[a:b], arr[roots[s]:ends[s]]])
    if name in ("first_subnode", "next_branch"):
        r = rows.astype(np.int64)
        part = (r >= 0) & (r < t.n)
        node = r >= t.n
        r[part] -= a
        r[node] = r[node] - roots[s] + (b - a)
        rows = r
    return rows


ARRAYS = ("coordinates", "masses", "softenings", "sizes", "deltas", "first_subnode", "next_branch", "nchild")


@pytest.mark.parametrize("quadrupole", [False, True])
def test_forest_tree_equals_single_builds(quadrupole):
    systems = _systems()
    x, m, h, off = _concat(systems)
    print(f"[forest] build nsys={len(systems)} n={len(x)} quad={quadrupole}", flush=True)
    f = engine.DeviceTree(x, m, h, quadrupole=quadrupole, offsets=off)
    cubes = f.cubes.cpu().numpy()
    order = f.order.cpu().numpy()
    nn_total = 0
    for s, (xs, ms, hs) in enumerate(systems):
        u = engine.DeviceTree(xs, ms, hs, quadrupole=quadrupole)
        nn_total += u.nnodes
        assert np.array_equal(cubes[s, :3], u.center) and cubes[s, 3] == u.size, s
        assert np.array_equal(order[off[s]:off[s + 1]] - off[s], u.order.cpu().numpy()), s
        for name in ARRAYS + (("quadrupoles",) if quadrupole else ()):
            single = getattr(u, name).cpu().numpy()
            if name in ("first_subnode", "next_branch"):
                single = single.astype(np.int64)
            got = _sub(f, off, s, name)
            if name in ("quadrupoles", "sizes", "deltas"):  # node rows only: particle rows are not written
                got, single = got[off[s + 1] - off[s]:], single[u.n:]
            assert np.array_equal(got, single), (s, name)
    assert f.nnodes == nn_total


@pytest.mark.parametrize("precision,quadrupole", [("float32", False), ("float32", True), ("float64", False),
                                                  ("float64", True)])
def test_forest_fields_equal_single_walks(precision, quadrupole):
    systems = _systems()
    x, m, h, off = _concat(systems)
    gs = 64 if precision == "float32" else 8
    print(f"[forest] walk prec={precision} quad={quadrupole}", flush=True)
    F = Forest(x, m, h, off, quadrupole=quadrupole, precision=precision, group_size=gs)
    r = F.evaluate(potential=True, accel=True)
    fa, fp = r["accel"].cpu().numpy(), r["potential"].cpu().numpy()
    for s, (xs, ms, hs) in enumerate(systems):
        u = engine.DeviceTree(xs, ms, hs, quadrupole=quadrupole)
        o = engine.walk(u, xs, hs, u.order, PTG_POT | PTG_ACC, quadrupole=quadrupole, group_size=gs, theta=0.7,
                        precision=precision).cpu().numpy()
        assert np.array_equal(fp[off[s]:off[s + 1]], o[:, 0]), s
        assert np.array_equal(fa[off[s]:off[s + 1]], o[:, 1:4]), s


def test_forest_accuracy_against_direct_sum():
    systems = _systems()[:8]
    x, m, h, off = _concat(systems)
    F = Forest(x, m, h, off, quadrupole=True, precision="float32")
    a = F.evaluate(potential=False, accel=True, theta=0.5)["accel"].cpu().numpy()
    for s, (xs, ms, hs) in enumerate(systems):
        d = engine.direct(xs, hs, xs, ms, hs, PTG_ACC, self_mode=True).cpu().numpy()
        err = np.sqrt(np.mean(np.sum((a[off[s]:off[s + 1]] - d) ** 2, 1)) / np.mean(np.sum(d * d, 1)))
        assert err < 2e-3, (s, err)


def test_forest_graph_replays_match_eager():
    systems = _systems()[:8]
    x, m, h, off = _concat(systems)
    F = Forest(x, m, h, off, quadrupole=False, precision="float32")
    ref = F.evaluate(potential=True, accel=True)
    ref = torch.cat([ref["potential"][:, None], ref["accel"]], 1).clone()
    replay, out = F.capture(potential=True, accel=True)
    for k in range(4):
        out.zero_()
        torch.cuda.synchronize()
        replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), k
        # eager work between replays
        e = F.evaluate(potential=True, accel=True)
        assert torch.equal(torch.cat([e["potential"][:, None], e["accel"]], 1), ref)


def test_forest_single_system_equals_plain_tree():
    x, m, h = case("plummer_1e5_h")
    F = Forest(x, m, h, [0, len(x)], precision="float32")
    r = F.evaluate(potential=True, accel=True)
    u = engine.DeviceTree(x, m, h)
    o = engine.walk(u, x, h, u.order, PTG_POT | PTG_ACC, group_size=64, theta=0.7, precision="float32")
    assert torch.equal(r["potential"], o[:, 0]) and torch.equal(r["accel"], o[:, 1:4])


def test_forest_rejects_bad_offsets():
    x, m, h = case("uniform_1e4")
    with pytest.raises(ValueError):
        Forest(x, m, h, [0, 1, len(x)])  # a one-particle system
    with pytest.raises(ValueError):
        Forest(x, m, h, [0, 5000, len(x) - 1])  # does not cover all particles
    # the library checks the device offsets itself too
    with pytest.raises(PtgError):
        engine.DeviceTree(x, m, h, offsets=np.array([0, 5000, 5001, len(x)]))


def test_forest_learned_equals_single_learned_walks():
    """With the same per-system tolerance, the learned-criterion forest walk equals each system's
    single-tree learned walk bit for bit; with alpha the per-system tolerances come from one
    sample walk and the fields match FP64 direct summation."""
    from pytreegrav_amd import learned

    systems = _systems()[:8]
    x, m, h, off = _concat(systems)
    model = learned.OpeningMLP.load()
    eps = np.array([1e-3 * (k + 1) for k in range(len(systems))])
    F = Forest(x, m, h, off, precision="float32", criterion="learned")
    r = F.evaluate(potential=True, accel=True, eps=eps, model=model)
    fa, fp = r["accel"].cpu().numpy(), r["potential"].cpu().numpy()
    for s, (xs, ms, hs) in enumerate(systems):
        u = engine.DeviceTree(xs, ms, hs, quadrupole=True)
        o = engine.walk(u, xs, hs, u.order, PTG_POT | PTG_ACC, group_size=64, theta=2.0, precision="float32",
                        criterion="learned", eps=float(eps[s]), model=model).cpu().numpy()
        assert np.array_equal(fp[off[s]:off[s + 1]], o[:, 0]), s
        assert np.array_equal(fa[off[s]:off[s + 1]], o[:, 1:4]), s
    a = Forest(x, m, h, off, criterion="learned").evaluate(potential=False, accel=True, alpha=1.6e-3)["accel"]
    a = a.cpu().numpy()
    for s, (xs, ms, hs) in enumerate(systems[2:], start=2):
        d = engine.direct(xs, hs, xs, ms, hs, PTG_ACC, self_mode=True).cpu().numpy()
        err = np.sqrt(np.mean(np.sum((a[off[s]:off[s + 1]] - d) ** 2, 1)) / np.mean(np.sum(d * d, 1)))
        assert err < 4e-3, (s, err)


def test_forest_autotune_and_learned_capture():
    systems = _systems()[:8]
    x, m, h, off = _concat(systems)
    F = Forest(x, m, h, off, group_size="auto", criterion="learned")
    r = F.evaluate(potential=True, accel=True, alpha=1.6e-3)
    assert F.tuning["choice"] in engine.TUNE_SIZES and F.group_size == F.tuning["choice"]
    replay, out = F.capture(potential=True, accel=True, alpha=1.6e-3)
    ref = F.evaluate(potential=True, accel=True, alpha=1.6e-3)
    for _ in range(3):
        out.zero_()
        replay()
        torch.cuda.synchronize()
        assert torch.equal(out[:, 0], ref["potential"]) and torch.equal(out[:, 1:4], ref["accel"])


def _tree_rows(t, nn):
    """The element rows of a tree's arrays (particles, then nn nodes) on the host."""
    ne = t.n + nn
    rows = {}
    for name in ARRAYS + ("slot", "child_base"):
        rows[name] = getattr(t, name)[:ne].cpu().numpy()
    for name in ("sizes", "deltas"):  # particle rows are never written on the device (zero by definition)
        rows[name] = rows[name][t.n:]
    rows["order"] = t.order.cpu().numpy()
    if t.quadrupoles is not None:
        rows["quadrupoles"] = t.quadrupoles[t.n:ne].cpu().numpy()
    return rows


@pytest.mark.parametrize("quadrupole", [False, True])
def test_deferred_build_equals_two_call_build(quadrupole):
    """ptg_build_deferred (no host synchronisation, node arrays sized for a capacity) writes the same
    tree as ptg_build_sort + ptg_build_tree, bit for bit, including the reference's collision and
    coincident-point cases; over its capacity it reports the overflow and writes no nodes."""
    systems = _systems()
    x, m, h, off = _concat(systems)
    print(f"[deferred] quad={quadrupole} n={len(x)}", flush=True)
    e = engine.DeviceTree(x, m, h, quadrupole=quadrupole, offsets=off)
    d = engine.DeviceTree(x, m, h, quadrupole=quadrupole, offsets=off,
                          capacity=(e.nnodes + 1000, int(e.info.max_level) + 3))
    c = d.check()
    assert c.nnodes == e.nnodes and c.max_level == e.info.max_level and int(d.node_count.item()) == e.nnodes
    assert bool(c.has_coincident) == bool(e.info.has_coincident)
    re, rd = _tree_rows(e, e.nnodes), _tree_rows(d, e.nnodes)
    for k in re:
        assert np.array_equal(re[k], rd[k]), k
    assert torch.equal(e.cubes, d.cubes) and torch.equal(e.roots, d.roots)
    # the build's FP32 walk copies agree on every slot of the tree
    pe, pd = e.pack(engine.PTG_F32, False), d.pack(engine.PTG_F32, False)
    ne = e.n + e.nnodes
    for a, b in zip(pe["tensors"][:3], pd["tensors"][:3]):
        assert torch.equal(a[:ne], b[:ne])
    # a single tree goes through the same path as a one-system forest
    xs, ms, hs = systems[5]
    e1 = engine.DeviceTree(xs, ms, hs, quadrupole=quadrupole)
    d1 = engine.DeviceTree(xs, ms, hs, quadrupole=quadrupole, capacity=(e1.nnodes, int(e1.info.max_level)))
    assert d1.check().nnodes == e1.nnodes
    r1, r2 = _tree_rows(e1, e1.nnodes), _tree_rows(d1, e1.nnodes)
    for k in r1:
        assert np.array_equal(r1[k], r2[k]), k
    # over capacity: reported, nothing written past the count (the count reads 0)
    small = engine.DeviceTree(x, m, h, quadrupole=quadrupole, offsets=off, capacity=(e.nnodes // 2, 40))
    assert int(small.node_count.item()) == 0
    with pytest.raises(PtgError):
        small.check()
    low = engine.DeviceTree(x, m, h, quadrupole=quadrupole, offsets=off, capacity=(e.nnodes + 10, 3))
    with pytest.raises(PtgError):
        low.check()


@pytest.mark.parametrize("criterion", ["theta", "learned"])
def test_forest_capture_step_replays_the_whole_step(criterion):
    """capture_step records build + groups (+ field scales and radii) + walk in one graph: each replay
    equals an eager evaluation of the current positions bit for bit, also after the positions are
    changed in place; a capacity too small is reported by the replay."""
    systems = _systems()[:8]
    x, m, h, off = _concat(systems)
    kw = dict(alpha=1.6e-3, expansion=1e-3) if criterion == "learned" else {}
    F = Forest(x, m, h, off, precision="float32", criterion=criterion)
    replay, out = F.capture_step(potential=True, accel=True, **kw)
    rng = np.random.default_rng(3)
    for k in range(3):
        print(f"[capture_step] {criterion} replay {k}", flush=True)
        if k:  # new positions, written into the captured input buffer
            x = x + 1e-3 * rng.normal(size=x.shape) * np.repeat([s[2][0] * 10 for s in systems], np.diff(off))[:, None]
            F.pos.copy_(torch.from_numpy(x).to(F.pos.device))
        out.zero_()
        replay()
        torch.cuda.synchronize()
        E = Forest(x, m, h, off, precision="float32", criterion=criterion)
        r = E.evaluate(potential=True, accel=True, **kw)
        assert torch.equal(out[:, 0], r["potential"]) and torch.equal(out[:, 1:4], r["accel"]), k
    # positions that make the captured capacities too small: every system collapsed onto one point,
    # so all of its particles share a 63-bit key (more than the deferred build re-keys); the replay
    # reports it, and the walk it ran saw massless roots only
    F2 = Forest(x, m, h, off, precision="float32", criterion=criterion)
    r2, out2 = F2.capture_step(potential=True, accel=True, **kw)
    same = np.repeat(x[off[:-1]], np.diff(off), axis=0)
    F2.pos.copy_(torch.from_numpy(same).to(F2.pos.device))
    with pytest.raises(PtgError):
        r2()
    F2.pos.copy_(torch.from_numpy(x).to(F2.pos.device))  # and it recovers with fitting positions
    r2()
    torch.cuda.synchronize()
    assert torch.equal(out2[:, 1:4], r["accel"])


def test_single_system_captured_step_equals_self_fields():
    """A one-system forest's captured step (sync-free build, cell groups, field scale from the same
    64 sample groups, radii, walk) reproduces engine.self_fields bit for bit: the graph form of the
    self-gravity evaluation that the benchmark times beside the eager one."""
    from pytreegrav_amd.sampling import plummer_truncated

    n = 200_000
    x, m = plummer_truncated(n, seed=17)
    h = np.full(n, 0.01)
    dev = engine.device()
    xd, md, hd = (torch.from_numpy(a).to(dev) for a in (x, m, h))
    kw = dict(alpha=5e-4, expansion=4e-3, expansion_rel=3.0)
    t = engine.DeviceTree(xd, md, hd, quadrupole=True)
    ref = engine.self_fields(t, xd, hd, PTG_POT | PTG_ACC, theta=2.0, group_size=64, criterion="learned", **kw)
    F = Forest(xd, md, hd, [0, n], precision="float32", criterion="learned", field_samples=64)
    replay, out = F.capture_step(potential=True, accel=True, theta=2.0, **kw)
    for _ in range(2):
        out.zero_()
        replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref)


def test_resident_builder_sizes_grows_and_matches_eager_builds():
    """engine.ResidentBuilder: the first call builds with exact counts, later calls build sync-free
    into arrays sized from it and equal the eager build; a step whose tree outgrows the capacities
    is reported by verify() (TreeOverflow, capacities grown) and the rebuilt tree fits."""
    from pytreegrav_amd.sampling import plummer_truncated

    n = 50_000
    x, m = plummer_truncated(n, seed=23)
    h = np.full(n, 0.01)
    dev = engine.device()
    xd, md, hd = (torch.from_numpy(a).to(dev) for a in (x, m, h))
    b = engine.ResidentBuilder(quadrupole=True, slack=1.0, extra_levels=0)
    t0 = b.verify(b(xd, md, hd))
    cap0 = b.capacity
    assert not getattr(t0, "deferred", False) and cap0[0] >= t0.nnodes
    t1 = b.verify(b(xd, md, hd))
    assert t1.deferred and t1.counts().nnodes == t0.nnodes
    ne = t0.n + t0.nnodes
    assert torch.equal(t1.coordinates[:ne], t0.coordinates[:ne]) and torch.equal(t1.masses[:ne], t0.masses[:ne])
    # a more clustered set needs more nodes and levels than the capacities allow
    xc = x * np.where(np.arange(n) % 2 == 0, 1e-3, 1.0)[:, None]
    xcd = torch.from_numpy(xc).to(dev)
    tc = b(xcd, md, hd)
    with pytest.raises(engine.ResidentBuilder.TreeOverflow):
        b.verify(tc)
    assert b.capacity is None or b.capacity[0] > cap0[0] or b.capacity[1] > cap0[1]
    t2 = b.verify(b(xcd, md, hd))
    e = engine.DeviceTree(xcd, md, hd, quadrupole=True)
    nn = t2.counts().nnodes if getattr(t2, "deferred", False) else t2.nnodes
    assert nn == e.nnodes
