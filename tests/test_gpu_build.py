"""GPU tree build against the oracle restatement of octree.py: bit-exact Morton order, index-free
topology and FP64 moments (monopole and quadrupole), on the reference's own test inputs, plus
invariants at larger N."""

import numpy as np
import pytest

import oracle
from gpu_cases import BUILD_CASES, case
from treeprint import assert_same_tree

pytestmark = pytest.mark.gpu


def _gpu_tree(x, m, h, quadrupole):
    import warnings

    from pytreegrav_amd import ConstructTree

    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return ConstructTree(x, m, h, quadrupole=quadrupole)


@pytest.mark.parametrize("name", BUILD_CASES)
@pytest.mark.parametrize("quadrupole", [False, True])
def test_build_bit_exact_against_oracle(name, quadrupole):
    x, m, h = case(name)
    print(f"[build] case={name} n={len(x)} quad={quadrupole}", flush=True)
    g = _gpu_tree(x, m, h, quadrupole)
    o = oracle.build_tree(x, m, h, quadrupole=quadrupole)
    assert g.NumParticles == o.NumParticles
    assert np.array_equal(g.TreewalkIndices, o.TreewalkIndices)
    assert g.HasCoincidentPoints == o.HasCoincidentPoints
    assert g.NumNodes == o.node_hi
    assert_same_tree(g, o, with_quads=quadrupole)


def test_build_plummer_1e6_bit_exact():
    x, m, h = case("plummer_1e6")
    print("[build] case=plummer_1e6", flush=True)
    g = _gpu_tree(x, m, h, True)
    o = oracle.build_tree(x, m, h, quadrupole=True)
    assert np.array_equal(g.TreewalkIndices, o.TreewalkIndices)
    assert_same_tree(g, o, with_quads=True)


def test_build_is_deterministic():
    x, m, h = case("plummer_1e5_h")
    a = _gpu_tree(x, m, h, True)
    b = _gpu_tree(x, m, h, True)
    for k in ("Coordinates", "Masses", "Softenings", "Deltas", "Quadrupoles", "FirstSubnode", "NextBranch"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


@pytest.mark.parametrize("n", [3_000_000])
def test_large_build_invariants(n):
    """No oracle at this size: mass conservation, the Morton order is a permutation sorted by key,
    every element is the child of exactly one node, and the link walk covers slots 0..n-1 in order."""
    from pytreegrav_amd.sampling import plummer_truncated

    x, m = plummer_truncated(n, seed=11)
    h = np.full(n, 0.01)
    print(f"[build] invariants n={n}", flush=True)
    t = _gpu_tree(x, m, h, False)
    assert np.array_equal(np.sort(t.TreewalkIndices), np.arange(n))
    keys = oracle.morton_keys(x, *oracle.bounding_cube(x))
    assert np.all(np.diff(keys[t.TreewalkIndices]) >= 0)
    assert np.isclose(t.Masses[n], m.sum(), rtol=1e-12)
    nc = t.dev.nchild.cpu().numpy()
    assert nc[:n].sum() == 0 and nc[n:].sum() == t.NumNodes - 1
    # links: following FirstSubnode/NextBranch visits particle slots in increasing order
    First, Next = t.FirstSubnode, t.NextBranch
    nxt_particle = Next[:n]
    # a particle's NextBranch is either the next slot or a node whose leftmost particle is that slot
    ok = (nxt_particle[:-1] == np.arange(1, n)) | (nxt_particle[:-1] >= n)
    assert ok.all() and nxt_particle[-1] == -1
    nodes = np.nonzero(nxt_particle[:-1] >= n)[0]
    # descend FirstSubnode from those nodes to a particle: must be the next slot
    cur = nxt_particle[nodes]
    for _ in range(100):
        isn = cur >= n
        if not isn.any():
            break
        cur[isn] = First[cur[isn]]
    assert np.array_equal(cur, nodes + 1)


@pytest.mark.parametrize("quadrupole", [False, True])
def test_build_fills_the_walk_copy_like_ptg_pack_walk(quadrupole):
    """The FP32 walk copy the build writes (from its moment pass) equals ptg_pack_walk's, bit for bit,
    for one tree, a forest and a capacity-sized sync-free build."""
    from pytreegrav_amd import engine
    from pytreegrav_amd._lib import PTG_F32

    x, m, h = case("plummer_3e4")
    off = np.array([0, 7000, 19000, len(m)], dtype=np.int64)
    for offsets in (None, off):
        t = engine.DeviceTree(x, m, h, quadrupole=quadrupole, offsets=offsets)
        built = [a.cpu().numpy() for a in t._packs[(PTG_F32, False)]["tensors"][:3]]
        fresh = [a.cpu().numpy() for a in t.pack(PTG_F32, False, cache=False, norms=True)["tensors"][:3]]
        for b, f in zip(built, fresh):
            assert np.array_equal(b, f)
        if quadrupole:  # aux[3] of a node slot: its quadrupole's Frobenius norm
            slot = t.slot.cpu().numpy()[t.n:]
            q = t.quadrupoles.cpu().numpy()[t.n:]
            want = np.sqrt((q * q).sum(1))
            got = built[1][slot, 3].astype(np.float64)
            assert np.allclose(got, want, rtol=1e-6, atol=1e-30)
    # the sync-free build (capacity-sized arrays) writes the same copy over its slots
    e = engine.DeviceTree(x, m, h, quadrupole=quadrupole)
    d = engine.DeviceTree(x, m, h, quadrupole=quadrupole, capacity=(e.nnodes + 500, int(e.info.max_level) + 3))
    ne = d.n + d.check().nnodes
    built = [a[:ne].cpu().numpy() for a in d._packs[(PTG_F32, False)]["tensors"][:3]]
    fresh = [a[:ne].cpu().numpy() for a in d.pack(PTG_F32, False, cache=False, norms=True)["tensors"][:3]]
    for b, f in zip(built, fresh):
        assert np.array_equal(b, f)


@pytest.mark.parametrize("quadrupole", [False, True])
def test_particle_rows_left_unwritten_are_never_read(quadrupole):
    """The build does not write the particle rows of sizes, deltas and quadrupoles (all zero by
    definition).  With those rows poisoned with NaN, the FP32 and FP64 walk copies, the binned column
    walk (which reads the tree arrays directly) and the host views are unchanged."""
    import torch

    from pytreegrav_amd import ConstructTree, engine
    from pytreegrav_amd._lib import PTG_F32, PTG_F64

    x, m, h = case("plummer_3e4")
    g = ConstructTree(x, m, h, quadrupole=quadrupole)
    t, n = g.dev, g.dev.n
    order = g.TreewalkIndices
    targets = x[order][:3000]
    before = [[a.cpu().numpy() for a in t.pack(p, q, cache=False, norms=True)["tensors"][:4] if a is not None]
              for p, q in ((PTG_F32, False), (PTG_F64, False), (PTG_F32, quadrupole))]
    col = engine.column_binned(t, targets, 16, 0.5).cpu().numpy()
    t.sizes[:n] = float("nan")
    t.deltas[:n] = float("nan")
    if quadrupole:
        t.quadrupoles.view(-1, 9)[:n] = float("nan")
    torch.cuda.synchronize()
    after = [[a.cpu().numpy() for a in t.pack(p, q, cache=False, norms=True)["tensors"][:4] if a is not None]
             for p, q in ((PTG_F32, False), (PTG_F64, False), (PTG_F32, quadrupole))]
    for b, a in zip(before, after):
        assert len(b) == len(a)
        for u, v in zip(b, a):
            assert np.array_equal(u, v)
    assert np.array_equal(engine.column_binned(t, targets, 16, 0.5).cpu().numpy(), col)
    # host views: particle rows zero, node rows as on the device
    for view, dev in ((g.Sizes, t.sizes), (g.Deltas, t.deltas)):
        assert np.all(view[:n] == 0.0) and np.array_equal(view[n:], dev.cpu().numpy()[n:])
    if quadrupole:
        assert np.all(g.Quadrupoles[:n] == 0.0)
