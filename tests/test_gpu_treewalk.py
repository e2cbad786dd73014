"""The reference's per-target walks and brute-force sums as exported by its frontend
(treewalk.py:22-341, bruteforce.py:7-386), on the GPU.

Subtree walks follow the reference's own test of the ``no`` argument (tests/test_subtree_walks.py:
127-157): every field is a plain sum over source elements and the acceptance test depends only on
the node under test, so the walks of the root's children sum to the whole walk (1e-12), and an
explicit root is the default walk.  Below the root each node is checked both ways: a node the
target opens equals the sum of its children's walks, a node it accepts equals its monopole from
the tree's own moments; a particle start is that one pair.  A walk that ignored ``no`` or ran past
the subtree's end fails the sums by about the number of children.  The many-target forms equal
the oracle's per-target walk (group size 1: the same accepted elements, so equal counters) and the
brute-force forms the oracle's direct sums."""

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def cloud(N=2000, seed=42, quadrupole=False):
    from pytreegrav_amd import ConstructTree

    rng = np.random.default_rng(seed)
    x = np.ascontiguousarray(rng.normal(size=(N, 3)))
    m = rng.uniform(0.5, 1.5, N) / N
    h = np.full(N, 2.0 * N ** (-1 / 3))
    return x, m, h, ConstructTree(x, m, h, quadrupole=quadrupole)


def children(tree, no):
    """Direct children of element `no` in depth-first order (FirstSubnode, then NextBranch until
    the parent's own NextBranch)."""
    kids, stop = [], tree.NextBranch[no]
    c = tree.FirstSubnode[no] if no >= tree.NumParticles else -1
    while c > -1 and c != stop:
        kids.append(int(c))
        c = tree.NextBranch[c]
    return kids


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


def test_root_has_several_children():
    _, _, _, tree = cloud()
    assert 2 <= len(children(tree, tree.NumParticles)) <= 8


@pytest.mark.parametrize("quad", [False, True])
@pytest.mark.parametrize("kind", ["potential", "accel"])
def test_subtrees_sum_to_whole(kind, quad):
    from pytreegrav_amd import AccelWalk, AccelWalk_quad, PotentialWalk, PotentialWalk_quad

    walk = {("potential", False): PotentialWalk, ("potential", True): PotentialWalk_quad,
            ("accel", False): AccelWalk, ("accel", True): AccelWalk_quad}[(kind, quad)]
    x, _, h, tree = cloud(quadrupole=quad)
    kids = children(tree, tree.NumParticles)
    print(f"[subtree] kind={kind} quad={quad} root children={len(kids)}", flush=True)
    for i in (0, 1234):
        whole = walk(x[i], tree, h[i], -1, 0.7)
        parts = sum(walk(x[i], tree, h[i], k, 0.7) for k in kids)
        assert np.any(np.abs(whole) > 0)
        assert rel(parts, whole) < RTOL
        # a walk that ignored `no` would return the whole once per child
        assert rel(walk(x[i], tree, h[i], kids[0], 0.7), whole) > 1e-3


def test_explicit_root_matches_default():
    from pytreegrav_amd import AccelWalk, PotentialWalk

    x, _, h, tree = cloud()
    root = tree.NumParticles
    assert PotentialWalk(x[3], tree, h[3], root, 0.7) == PotentialWalk(x[3], tree, h[3], -1, 0.7)
    assert np.array_equal(AccelWalk(x[3], tree, h[3], root, 0.7), AccelWalk(x[3], tree, h[3], -1, 0.7))


def test_every_node_opened_or_accepted_as_the_reference_decides():
    """Nodes in the first levels below the root, for a target inside the cloud and one outside."""
    from pytreegrav_amd import PotentialWalk, acceptance_criterion

    x, _, h, tree = cloud()
    n = tree.NumParticles
    C, M, S, D, H = tree.Coordinates, tree.Masses, tree.Sizes, tree.Deltas, tree.Softenings
    frontier, nodes = [n], []
    while frontier and len(nodes) < 40:
        nxt = []
        for no in frontier:
            for k in children(tree, no):
                if k >= n:
                    nodes.append(k)
                    nxt.append(k)
        frontier = nxt
    targets = [(x[7], h[7]), (np.array([2.5, -1.0, 0.5]), 0.0)]
    n_open = n_acc = 0
    for pos, ht in targets:
        for no in nodes[:40]:
            r = float(np.sqrt(np.sum((C[no] - pos) ** 2)))
            got = PotentialWalk(pos, tree, ht, no, 0.7)
            if acceptance_criterion(r, max(H[no], ht), S[no], D[no], 0.7):
                n_acc += 1
                assert abs(got - (-M[no] / r)) <= 1e-13 * abs(M[no] / r)
            else:
                n_open += 1
                parts = sum(PotentialWalk(pos, tree, ht, k, 0.7) for k in children(tree, no))
                assert abs(parts - got) <= RTOL * abs(got)
    print(f"[subtree] opened {n_open} accepted {n_acc}", flush=True)
    assert n_open > 0 and n_acc > 0


def test_particle_start_is_one_pair():
    from pytreegrav_amd import AccelWalk, PotentialWalk

    x, _, h, tree = cloud()
    n = tree.NumParticles
    C, M, H = tree.Coordinates[:n], tree.Masses[:n], tree.Softenings[:n]  # particle rows (tree order)
    pos, ht = x[11], 0.3  # a softening large enough that near pairs take the spline branch
    d = np.sqrt(np.sum((C - pos) ** 2, axis=1))
    near = [int(p) for p in np.argsort(d)[:6]]
    for p in near + [int(np.argmax(d))]:
        r, hh = float(d[p]), max(float(H[p]), ht)
        if r == 0:  # the target's own particle: the self term is skipped
            assert PotentialWalk(pos, tree, ht, p, 0.7) == 0.0
            continue
        phi = M[p] * oracle.potential_kernel(r, hh) if r < hh else -M[p] / r
        g = M[p] * (oracle.force_kernel(r, hh) if r < hh else 1.0 / r ** 3) * (C[p] - pos)
        assert abs(PotentialWalk(pos, tree, ht, p, 0.7) - phi) <= 1e-13 * abs(phi)
        assert rel(AccelWalk(pos, tree, ht, p, 0.7), g) <= 1e-13


@pytest.mark.parametrize("quad,theta", [(False, 0.7), (True, 0.5), (False, 1.0)])
def test_target_tree_matches_oracle(quad, theta):
    import torch

    from pytreegrav_amd import AccelTarget_tree, PotentialTarget_tree, engine

    x, m, h, tree = cloud(N=5000, seed=3, quadrupole=quad)
    rng = np.random.default_rng(5)
    xt = np.concatenate([rng.uniform(-3, 3, (2000, 3)), x[:50]])  # 50 targets sit on particles
    ht = rng.uniform(0.0, 0.2, len(xt))
    print(f"[target_tree] quad={quad} theta={theta} nt={len(xt)}", flush=True)
    phi = PotentialTarget_tree(xt, ht, tree, G=1.7, theta=theta, quadrupole=quad)
    acc = AccelTarget_tree(xt, ht, tree, G=1.7, theta=theta, quadrupole=quad)
    ot = oracle.build_tree(x, m, h, quadrupole=quad)
    ref, rctr = oracle.fields_grouped(ot, xt, ht, potential=True, accel=True, quadrupole=quad, group_size=1,
                                      G=1.7, theta=theta, return_counters=True)
    assert rel(phi, ref["potential"]) < 1e-10
    assert rel(acc, ref["accel"]) < 1e-10
    ctr = torch.zeros(4, dtype=torch.int64, device=engine.device())
    engine.walk_subtree(tree.dev, xt, ht, -1, engine.fields_mask(potential=True, accel=True), quadrupole=quad,
                        theta=theta, counters=ctr)
    assert int(ctr[0]) == rctr["interactions"] and int(ctr[1]) == rctr["visits"]


def test_target_tree_device_inputs_and_defaults():
    import torch

    from pytreegrav_amd import AccelTarget_tree, AccelTarget_tree_parallel, engine

    x, _, h, tree = cloud()
    xt = np.random.default_rng(1).normal(size=(300, 3))
    host = AccelTarget_tree(xt, None, tree)
    dev = AccelTarget_tree(torch.tensor(xt, device=engine.device()), None, tree)
    assert isinstance(dev, torch.Tensor) and dev.is_cuda
    assert np.array_equal(dev.cpu().numpy(), host)
    assert np.array_equal(AccelTarget_tree_parallel(xt, np.zeros(300), tree), host)


def test_subtree_walk_arguments():
    import torch

    from pytreegrav_amd import PotentialWalk, PotentialWalk_quad, _lib, engine

    x, _, h, tree = cloud()
    with pytest.raises(ValueError, match="not an element"):
        PotentialWalk(x[0], tree, h[0], tree.NumNodes, 0.7)
    with pytest.raises(ValueError, match="quadrupole"):
        PotentialWalk_quad(x[0], tree, h[0], -1, 0.7)
    # a start slot outside the walk tree (forced from the host through the C ABI) walks nothing and
    # is reported by the check
    L = _lib.load()
    p = tree.dev.pack(_lib.PTG_F64, False)
    bad = torch.tensor([tree.NumNodes + 5], dtype=torch.int32, device=engine.device())
    pos = torch.tensor(x[:2], device=engine.device())
    soft = torch.tensor(h[:2], device=engine.device())
    perm = torch.arange(2, device=engine.device())
    out = torch.full((2, 1), 7.0, dtype=torch.float64, device=engine.device())
    nb = L.ptg_walk_workspace_bytes(2, 1)
    ws = torch.empty(nb, dtype=torch.uint8, device=engine.device())
    st = L.ptg_walk_subtree(_lib.ctypes.byref(p["wt"]), _lib.ptr(bad), _lib.ptr(pos), _lib.ptr(soft), _lib.ptr(perm), 2,
                            1, 0, 0.7, 1.0, _lib.ptr(out), None, _lib.ptr(ws), nb, engine.stream())
    assert st == 0
    assert L.ptg_walk_check(_lib.ptr(ws), engine.stream()) == -1  # PTG_ERR_ARG
    assert torch.count_nonzero(out).item() == 0


@pytest.mark.parametrize("name", ["potential", "accel", "tidal"])
def test_bruteforce_forms_match_oracle(name):
    import pytreegrav_amd as p

    rng = np.random.default_rng(9)
    x = rng.normal(size=(1500, 3))
    x[10] = x[20]  # a coincident pair: r = 0 contributes nothing
    m = rng.uniform(0.5, 1.5, 1500)
    h = rng.uniform(0.01, 0.3, 1500)
    xt = np.concatenate([rng.normal(size=(400, 3)), x[:5]])
    ht = rng.uniform(0.0, 0.2, len(xt))
    self_f = {"potential": p.Potential_bruteforce, "accel": p.Accel_bruteforce, "tidal": p.TidalTensor_bruteforce}
    tgt_f = {"potential": p.PotentialTarget_bruteforce, "accel": p.AccelTarget_bruteforce,
             "tidal": p.TidalTensorTarget_bruteforce}
    ref_s = oracle.fields_self_bruteforce(x, m, h, G=0.5, **{name: True})[name]
    ref_t = oracle.fields_target_bruteforce(xt, ht, x, m, h, G=0.5, **{name: True})[name]
    got_s = self_f[name](x, m, h, G=0.5)
    got_t = tgt_f[name](xt, ht, x, m, h, G=0.5)
    assert got_s.shape == ref_s.shape and got_t.shape == ref_t.shape
    assert rel(got_s, ref_s) < 1e-12 and rel(got_t, ref_t) < 1e-12
    par = {"potential": p.Potential_bruteforce_parallel, "accel": p.Accel_bruteforce_parallel,
           "tidal": p.TidalTensor_bruteforce_parallel}[name]
    assert np.array_equal(par(x, m, h, G=0.5), got_s)
    if name != "tidal":
        sym = {"potential": p.Potential_bruteforce_symmetric, "accel": p.Accel_bruteforce_symmetric}[name]
        assert np.array_equal(sym(x, m, h, G=0.5), got_s)


def test_bruteforce_invariants_of_the_reference_tests():
    """tests/test_bruteforce_symmetric.py:56-76, 169-172: momentum conservation (the per-target GPU
    sums cancel to rounding, not exactly as the reference's symmetrised sum does, hence 1e-12),
    coincident unsoftened particles give finite zeros, and G scales every field linearly."""
    import pytreegrav_amd as p

    rng = np.random.default_rng(7)
    x = rng.normal(size=(500, 3))
    m = rng.uniform(0.5, 1.5, 500)
    h = rng.uniform(0.05, 0.3, 500)
    a = p.Accel_bruteforce_symmetric(x, m, h)
    net = np.abs((m[:, None] * a).sum(axis=0)).max()
    assert net < 1e-12 * np.abs(m[:, None] * a).sum()
    z = np.zeros((4, 3))
    phi0, a0 = p.Potential_bruteforce(z, np.ones(4), np.zeros(4)), p.Accel_bruteforce(z, np.ones(4), np.zeros(4))
    assert np.all(np.isfinite(phi0)) and np.all(np.isfinite(a0))
    assert not np.any(phi0) and not np.any(a0)
    for f in (p.Potential_bruteforce, p.Accel_bruteforce, p.TidalTensor_bruteforce):
        assert rel(f(x[:200], m[:200], h[:200], 3.0), 3.0 * f(x[:200], m[:200], h[:200])) < 1e-14


def test_column_density_drivers_match_oracle():
    """ColumnDensity_tree / _grouped / Binned_grouped (treewalk.py:1031-1294) against the oracle's
    ray-traced and binned walks (1e-12); the grouped ray trace equals the per-target one bit for bit,
    as in the reference; randomize_rays rotates target i's grid by random_rotation(i)."""
    import pytreegrav_amd as p
    from pytreegrav_amd.misc import random_rotation

    rng = np.random.default_rng(4)
    N = 20000
    x = rng.normal(size=(N, 3))
    m = np.full(N, 1.0 / N)
    r = np.full(N, 3 * N ** (-1 / 3))
    tree = p.ConstructTree(x, m, r)
    o = oracle.build_tree(x, m, r)
    xt = x[:300]
    rays = np.vstack([np.eye(3), -np.eye(3)])
    got = p.ColumnDensity_tree(xt, tree, rays)
    assert rel(got, oracle.column_rays(o, xt, rays)) < 1e-12
    assert np.array_equal(p.ColumnDensity_grouped(xt, rays, tree, group_size=16), got)
    rot = np.stack([rays @ random_rotation(i) for i in range(len(xt))])
    assert rel(p.ColumnDensity_tree(xt, tree, rays, randomize_rays=True),
               oracle.column_rays(o, xt, rot, per_target=True)) < 1e-12
    assert rel(p.ColumnDensity_tree(xt, tree, None, theta=0.7), oracle.column_binned(o, xt, 1, 0.7)) < 1e-12
    xs = x[:1000][p._morton_order(x[:1000]).cpu().numpy()]  # Morton-ordered targets: compact groups
    assert rel(p.ColumnDensityBinned_grouped(xs, tree, 0.5, 16), oracle.column_binned(o, xs, 16, 0.5)) < 1e-12
    with pytest.raises(ValueError):
        p.ColumnDensity_grouped(xt, rays, tree, group_size=0)
