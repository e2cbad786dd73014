"""Column density on the GPU (treewalk.py:821-1276, frontend.ColumnDensity :1321-1457) against the
oracle restatement (itself pinned by closed-form chords in test_column_cpu.py): the ray-traced walk
to 1e-12 relative (both exact, sums in the same depth-first order), the binned estimator to 1e-12 at
the reference's grouping, and the frontend's dispatch, warnings and errors."""

import numpy as np
import pytest
import torch

import oracle
from pytreegrav_amd import ColumnDensity, ConstructTree, engine
from pytreegrav_amd.misc import random_rotation

pytestmark = pytest.mark.gpu


def cloud(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 3))
    m = np.full(n, 1.0 / n)
    h = np.full(n, 3.0 * n ** (-1 / 3))
    return x, m, h


def rel(a, b):
    s = np.abs(b).max()
    return np.abs(a - b).max() / (s if s > 0 else 1.0)


@pytest.mark.parametrize("nrays", [6, 13])
def test_rays_match_oracle(nrays):
    x, m, h = cloud(3000, 1)
    print(f"[column] rays n=3000 nrays={nrays}", flush=True)
    rng = np.random.default_rng(2)
    rays = np.vstack([np.eye(3), -np.eye(3)]) if nrays == 6 else rng.normal(size=(nrays, 3))
    rays = rays / np.linalg.norm(rays, axis=1, keepdims=True)
    t = ConstructTree(x, m, h)
    o = oracle.build_tree(x, m, h)
    targets = np.vstack([x[:500], rng.normal(size=(100, 3)) * 3])
    got = engine.column_rays(t.dev, targets, rays).cpu().numpy()
    want = oracle.column_rays(o, targets, rays)
    assert rel(got, want) < 1e-12


def test_per_target_rays_match_oracle():
    x, m, h = cloud(2000, 3)
    t = ConstructTree(x, m, h)
    o = oracle.build_tree(x, m, h)
    rays = np.vstack([np.eye(3), -np.eye(3)])
    rot = np.stack([rays @ random_rotation(k) for k in range(300)])
    got = engine.column_rays(t.dev, x[:300], rot, per_target=True).cpu().numpy()
    want = oracle.column_rays(o, x[:300], rot, per_target=True)
    assert rel(got, want) < 1e-12


@pytest.mark.parametrize("gs,theta", [(16, 0.5), (1, 0.5), (16, 0.8), (64, 0.3), (48, 0.5), (100, 0.5)])
def test_binned_matches_oracle(gs, theta):
    x, m, h = cloud(4000, 4)
    t = ConstructTree(x, m, h)
    o = oracle.build_tree(x, m, h)
    order = t.TreewalkIndices
    targets = x[order]  # Morton order, as the frontend groups them
    got = engine.column_binned(t.dev, targets, gs, theta).cpu().numpy()
    want = oracle.column_binned(o, targets, gs, theta)
    assert rel(got, want) < 1e-12


def test_frontend_matches_oracle_in_input_order():
    x, m, h = cloud(2500, 5)
    o = oracle.build_tree(x, m, h)
    # rays: exact, independent of the grouping
    rays = np.random.default_rng(6).normal(size=(5, 3))
    got = ColumnDensity(x, m, h, rays=rays.copy())
    want = oracle.column_rays(o, x, rays / np.linalg.norm(rays, axis=1, keepdims=True))
    assert got.shape == (2500, 5) and rel(got, want) < 1e-12
    # binned: groups of 16 consecutive Morton targets (>= 1000 targets)
    gotb = ColumnDensity(x, m, h)
    idx = o.TreewalkIndices
    wantb = np.empty((2500, 6))
    wantb[idx] = oracle.column_binned(o, x[idx], 16, 0.5)
    assert gotb.shape == (2500, 6) and rel(gotb, wantb) < 1e-12
    # randomize_rays: target of Morton rank k uses the grid rotated by random_rotation(k)
    g6 = np.vstack([np.eye(3), -np.eye(3)])
    gotr = ColumnDensity(x, m, h, rays=6, randomize_rays=True)
    rot = np.stack([g6 @ random_rotation(k) for k in range(2500)])
    wantr = np.empty((2500, 6))
    wantr[idx] = oracle.column_rays(o, x[idx], rot, per_target=True)
    assert rel(gotr, wantr) < 1e-12


def test_frontend_small_sets_walk_per_target_and_trees_can_be_reused():
    x, m, h = cloud(500, 7)
    o = oracle.build_tree(x, m, h)
    idx = o.TreewalkIndices
    got = ColumnDensity(x, m, h)  # fewer than 1000 targets: per-target binned walk
    want = np.empty((500, 6))
    want[idx] = oracle.column_binned(o, x[idx], 1, 0.5)
    assert rel(got, want) < 1e-12
    # a supplied tree holding other particles: the targets are sorted on their own
    cols, tree = ColumnDensity(x, m, h, rays=6, return_tree=True)
    targets = np.random.default_rng(8).normal(size=(50, 3))
    got2 = ColumnDensity(targets, None, np.ones(50), rays=6, tree=tree)
    want2 = oracle.column_rays(o, targets, np.vstack([np.eye(3), -np.eye(3)]))
    assert rel(got2, want2) < 1e-12
    # device tensors in, device tensors out
    xt = torch.from_numpy(x).cuda()
    got3 = ColumnDensity(xt, torch.from_numpy(m).cuda(), torch.from_numpy(h).cuda(), rays=6)
    assert isinstance(got3, torch.Tensor) and got3.is_cuda
    assert rel(got3.cpu().numpy(), cols) < 1e-15


def test_frontend_checks():
    x, m, h = cloud(200, 9)
    with pytest.warns(UserWarning, match="point-like"):
        hh = h.copy()
        hh[0] = 0.0
        ColumnDensity(x, m, hh, rays=6)
    with pytest.raises(Exception, match="must be 2D"):
        ColumnDensity(x, m, h, rays=np.ones(3))
    with pytest.raises(Exception, match="3D vectors"):
        ColumnDensity(x, m, h, rays=np.ones((4, 2)))
    with pytest.raises(Exception, match="not supported"):
        ColumnDensity(x, m, h, rays=2.5)
    with pytest.raises(ValueError, match="device must be"):
        ColumnDensity(x, m, h, rays=6, device="tpu")
    with pytest.raises(ValueError, match="ray-traced path"):
        ColumnDensity(x, m, h, device="cuda")
    try:
        import healpy  # noqa: F401
    except ImportError:
        with pytest.raises(ImportError, match="healpy"):
            ColumnDensity(x, m, h, healpix=1)


@pytest.mark.parametrize("gs", [16, 48, 100])
def test_binned_counters_count_every_pair_when_nothing_is_accepted(gs):
    """theta -> 0 accepts no node, so every group visits every element of the tree once and evaluates
    every particle for every one of its targets: pairs = n * nt, visits = groups * (n + nnodes)."""
    x, m, h = cloud(3000, 6)
    t = ConstructTree(x, m, h)
    targets = x[t.TreewalkIndices][:1000]
    ctr = torch.zeros(2, dtype=torch.int64, device="cuda")
    got = engine.column_binned(t.dev, targets, gs, 1e-9, counters=ctr).cpu().numpy()
    assert ctr[0].item() == 3000 * 1000
    assert ctr[1].item() == -(-1000 // gs) * t.dev.ne
    o = oracle.build_tree(x, m, h)
    assert rel(got, oracle.column_binned(o, targets, gs, 1e-9)) < 1e-12


def test_ray_counters_accumulate_and_checks():
    x, m, h = cloud(2000, 7)
    t = ConstructTree(x, m, h)
    rays = np.vstack([np.eye(3), -np.eye(3)])
    ctr = torch.zeros(1, dtype=torch.int64, device="cuda")
    engine.column_rays(t.dev, x[:500], rays, counters=ctr)
    one = ctr[0].item()
    assert one >= 500 * 6  # every walk visits at least the root
    engine.column_rays(t.dev, x[:500], rays, counters=ctr)
    assert ctr[0].item() == 2 * one
    with pytest.raises(ValueError):
        engine.column_rays(t.dev, x[:500], rays, counters=torch.zeros(1, dtype=torch.int32, device="cuda"))
    with pytest.raises(ValueError):
        engine.column_binned(t.dev, x[:500], 16, 0.5, counters=torch.zeros(1, dtype=torch.int64, device="cuda"))


@pytest.mark.parametrize("gs", [16, 64])
def test_binned_heavy_groups_split_across_wavefronts_match_oracle(gs):
    """At N=1e5 some groups of consecutive Morton targets straddle the whole cloud and visit the whole
    tree (far past the 32768-visit budget): they are finished by the pass that deals their subtrees out
    to 64 wavefronts and summed share by share; every group must still equal the oracle."""
    x, m, h = cloud(100_000, 8)
    t = ConstructTree(x, m, h)
    targets = x[t.TreewalkIndices]
    got = engine.column_binned(t.dev, targets, gs, 0.5).cpu().numpy()
    o = oracle.build_tree(x, m, h)
    want = oracle.column_binned(o, targets, gs, 0.5, nthreads=16)
    assert rel(got, want) < 1e-12
    again = engine.column_binned(t.dev, targets, gs, 0.5).cpu().numpy()
    assert np.array_equal(got, again)  # deterministic: shares are summed in a fixed order


def test_wavefront_ray_walk_equals_stackless_walk():
    """The wavefront walk (64 targets along one ray, the union of their paths in depth-first order)
    applies each lane's own tests in its own order: the per-thread stackless walk's visit counts
    exactly and its columns to rounding, for shared and per-target rays."""
    x, m, h = cloud(20_000, 9)
    t = ConstructTree(x, m, h)
    L = engine._lib.load()
    rays = np.vstack([np.eye(3), -np.eye(3), [[0.6, 0.0, 0.8]]])
    rot = np.stack([rays @ random_rotation(k) for k in range(777)])
    for tg, ry, pt in ((x[:5000], rays, 0), (x[:777], rot, 1)):
        nt, nr = len(tg), ry.shape[-2]
        tp = engine.to_dev(tg, (-1, 3))
        rr = engine.to_dev(ry)
        outs, visits = [], []
        pack = engine._lib.ctypes.byref(t.dev.pack(engine.PTG_F64, False)["wt"])
        # wavefront walk in target order, in per-ray orders across the ray (shared rays), stackless walk
        order = engine._ray_orders(tp, rr) if not pt else None
        for wt, od in ((pack, None), (pack, order), (None, None)):
            if wt is not None and od is None and len(outs) == 1:
                continue
            out = torch.empty((nt, nr), dtype=torch.float64, device="cuda")
            err = torch.zeros(1, dtype=torch.int32, device="cuda")
            ctr = torch.zeros(1, dtype=torch.int64, device="cuda")
            engine.check(L.ptg_column_rays(engine._lib.ctypes.byref(t.dev.arrays), t.dev.n, t.dev.nnodes, wt,
                                           engine.ptr(tp), nt, engine.ptr(rr), nr, pt, engine.ptr(od),
                                           engine.ptr(out), engine.ptr(err), engine.ptr(ctr), engine.stream()),
                         "ptg_column_rays")
            assert int(err.item()) == 0
            outs.append(out.cpu().numpy())
            visits.append(int(ctr.item()))
        assert len(set(visits)) == 1
        for o in outs[1:]:
            # same terms in the same order; the kernels may contract multiply-adds differently, which
            # grazing chords (sqrt(1 - q^2) with q near 1) amplify to about 1e-13 of the largest column
            assert rel(outs[0], o) < 1e-12
        if order is not None:
            assert np.array_equal(outs[0], outs[1])  # the packet composition does not change a column
            assert np.array_equal(np.sort(order.cpu().numpy(), 1), np.tile(np.arange(nt), (nr, 1)))


def test_device_ray_rotations_equal_host_rotations():
    """The per-target grids of randomize_rays, drawn on the device, equal misc.random_rotation's
    (numpy's seeded Mersenne Twister) to the last few bits (device cos/sin may differ by an ulp)."""
    rays = np.random.default_rng(9).normal(size=(7, 3))
    rays /= np.linalg.norm(rays, axis=1, keepdims=True)
    for first, nt in ((0, 3000), (2**32 - 64, 64), (123456789, 100)):
        got = engine.random_rotations(rays, nt, first_seed=first).cpu().numpy()
        want = np.stack([rays @ random_rotation(first + k) for k in range(nt)])
        assert got.shape == (nt, 7, 3) and np.abs(got - want).max() < 1e-14, first
    with pytest.raises(RuntimeError):
        engine.random_rotations(rays, 65, first_seed=2**32 - 64)
