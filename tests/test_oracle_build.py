"""Oracle keys, sort, radix split and moments (octree.py) against independent checks.

Keys: the bit interleave against a per-bit loop; quantisation against the dyadic cell each point
lies in.  Sort: numpy.argsort(kind="stable").  Topology: brute-force enumeration of every dyadic
cell holding >= 2 points (the reference's node definition, octree.py:817-1010), plus the
reference's own link-walk, child-sum, re-key and duplicate tests (tests/test_radix_build.py,
tests/test_nonunique_positions.py).  Moments: per-node COM/mass/hmax recomputed directly from
the particles each node contains, root quadrupole analytic (tests/test_quadrupole.py:16-44).
"""

import numpy as np
import pytest

import oracle


def _interleave_loop(qx, qy, qz):
    k = 0
    for b in range(21):
        k |= ((qx >> b) & 1) << (3 * b)
        k |= ((qy >> b) & 1) << (3 * b + 1)
        k |= ((qz >> b) & 1) << (3 * b + 2)
    return k


def test_spread_bits_is_the_bit_interleave():
    rng = np.random.default_rng(0)
    for q in list(rng.integers(0, 1 << 21, 200)) + [0, 1, (1 << 21) - 1]:
        q = int(q)
        assert oracle.spread_bits(q) == _interleave_loop(q, 0, 0)


def test_keys_locate_each_point_in_its_dyadic_cell():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(3000, 3))
    c, s = oracle.bounding_cube(x)
    assert s == (x.max(0) - x.min(0)).max()
    assert np.array_equal(c, 0.5 * (x.max(0) + x.min(0)))
    keys = oracle.morton_keys(x, c, s)
    assert np.all(keys >= 0)
    cmin = c - 0.5 * s
    for i in range(0, 3000, 7):
        k = int(keys[i])
        q = [0, 0, 0]
        for b in range(21):
            for d in range(3):
                q[d] |= ((k >> (3 * b + d)) & 1) << b
        for d in range(3):
            lo = cmin[d] + q[d] * s / 2 ** 21
            hi = cmin[d] + (q[d] + 1) * s / 2 ** 21
            tol = 4 * np.finfo(float).eps * s
            assert lo - tol <= x[i, d] <= hi + tol


@pytest.mark.parametrize("n", [1, 2, 1000, 50000])
def test_radix_argsort_is_stable_argsort(n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 1 << 63, n, dtype=np.int64)
    keys[: n // 3] = keys[0]  # ties
    keys[n // 2: n // 2 + n // 5] = rng.integers(0, 64, n // 5)
    assert np.array_equal(oracle.radix_argsort(keys), np.argsort(keys, kind="stable"))


# ---------------------------------------------------------------------------------------------
# topology
# ---------------------------------------------------------------------------------------------


def walk_nodes(t):
    """DFS over FirstSubnode/NextBranch: for each node, (size, particle-slot range), children."""
    n = t.NumParticles
    out = {}

    def kids(no):
        res, c, stop = [], t.FirstSubnode[no], t.NextBranch[no]
        while c != -1 and c != stop:
            res.append(int(c))
            c = t.NextBranch[c]
            assert len(res) <= 8
        return res

    stack = [n]
    budget = 8 * t.NumNodes + 64
    while stack:
        budget -= 1
        assert budget > 0, "link chain does not terminate"
        no = stack.pop()
        if no < n:
            continue
        k = kids(no)
        out[no] = k
        stack.extend(reversed(k))
    return out


def subtree_particles(t, no, kids):
    if no < t.NumParticles:
        return [no]
    r = []
    for c in kids[no]:
        r += subtree_particles(t, c, kids)
    return r


def brute_force_cells(x):
    """Every dyadic cell (level, prefix) of the root cube holding >= 2 points: the node set."""
    c, s = oracle.bounding_cube(x)
    keys = oracle.morton_keys(x, c, s)
    cells = set()
    for L in range(0, 22):
        shift = 63 - 3 * L
        pref = keys >> shift if L > 0 else np.zeros_like(keys)
        uniq, inv, cnt = np.unique(pref, return_inverse=True, return_counts=True)
        for u in np.nonzero(cnt >= 2)[0]:
            cells.add((L, frozenset(np.nonzero(inv == u)[0].tolist())))
    return cells


@pytest.mark.parametrize("n,seed", [(2, 0), (3, 1), (50, 2), (400, 3), (3000, 4)])
def test_node_set_equals_dyadic_cells_with_two_or_more_points(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.random((n, 3)) ** 3  # uneven, so chains of single-child nodes occur
    t = oracle.build_tree(x, np.ones(n), np.zeros(n))
    kids = walk_nodes(t)
    root_size = t.Sizes[n]
    got = set()
    for no in kids:
        L = int(round(np.log2(root_size / t.Sizes[no])))
        assert t.Sizes[no] == root_size * 0.5 ** L
        got.add((L, frozenset(t.TreewalkIndices[subtree_particles(t, no, kids)].tolist())))
    assert got == brute_force_cells(x)
    assert len(kids) == t.NumNodes - n


def test_lone_particle():
    t = oracle.build_tree(np.array([[0.2, 0.3, 0.4]]), np.array([2.0]), np.array([0.1]))
    assert t.NumParticles == 1 and t.NumNodes == 2
    assert t.Masses[1] == 2.0 and t.FirstSubnode[1] == 0 and t.NextBranch[0] == -1


@pytest.mark.parametrize("n", [1, 2, 997, 5000])
def test_links_reach_every_particle_once(n):
    rng = np.random.default_rng(n)
    x = rng.random((n, 3))
    t = oracle.build_tree(x, np.full(n, 1.0 / n), np.zeros(n))
    seen, no = [], n
    for _ in range(4 * t.NumNodes + 16):
        if no <= -1:
            break
        if no < n:
            seen.append(no)
            no = t.NextBranch[no]
        else:
            no = t.FirstSubnode[no]
    assert np.array_equal(np.array(seen), np.arange(n))  # DFS visits Morton slots in order
    assert np.array_equal(np.sort(t.TreewalkIndices), np.arange(n))


def test_node_masses_equal_sum_of_children_and_moments_are_direct_sums():
    n = 4000
    rng = np.random.default_rng(5)
    x = rng.random((n, 3))
    m = rng.random(n) / n
    h = rng.random(n) * 0.01
    t = oracle.build_tree(x, m, h)
    kids = walk_nodes(t)
    xs, ms, hs = x[t.TreewalkIndices], m[t.TreewalkIndices], h[t.TreewalkIndices]
    for no, k in kids.items():
        assert np.isclose(sum(t.Masses[c] for c in k), t.Masses[no], rtol=1e-12)
        p = subtree_particles(t, no, kids)
        assert np.isclose(ms[p].sum(), t.Masses[no], rtol=1e-12)
        com = (ms[p, None] * xs[p]).sum(0) / ms[p].sum()
        assert np.allclose(t.Coordinates[no], com, rtol=1e-12, atol=1e-14)
        assert t.Softenings[no] == hs[p].max()


def test_deltas_measure_com_offset_from_the_cell_centre():
    n = 3000
    rng = np.random.default_rng(8)
    x = rng.random((n, 3))
    t = oracle.build_tree(x, np.full(n, 1.0 / n), np.zeros(n))
    t0 = oracle.build_tree(x, np.full(n, 1.0 / n), np.zeros(n), compute_moments=False)
    kids = walk_nodes(t)
    for no in kids:
        geo = t0.Coordinates[no]  # geometric centre before the moments pass
        assert np.isclose(t.Deltas[no], np.linalg.norm(t.Coordinates[no] - geo), rtol=1e-12, atol=1e-15)
        # the cell centre is the centre of the dyadic cell holding the node's particles
        p = subtree_particles(t, no, kids)
        assert np.all(np.abs(t.Coordinates[p] - geo) <= 0.5 * t.Sizes[no] * (1 + 1e-9))


@pytest.mark.parametrize("n", [8, 50, 500, 5000])
def test_root_quadrupole_matches_analytic(n):
    rng = np.random.default_rng(0)
    x, m = rng.random((n, 3)), rng.random(n)
    t = oracle.build_tree(x, m, np.zeros(n), quadrupole=True)
    com = t.Coordinates[n]
    r = x - com
    Q_ref = 3.0 * np.einsum("i,ik,il->kl", m, r, r) - np.eye(3) * np.sum(m * np.sum(r ** 2, axis=1))
    assert np.max(np.abs(t.Quadrupoles[n] - Q_ref)) / np.max(np.abs(Q_ref)) < 1e-10
    assert np.all(t.Quadrupoles[:n] == 0)


def test_rekeying_separates_points_sharing_a_key():
    """tests/test_radix_build.py:82-98 inputs."""
    np.random.seed(7)
    corners = np.array([[0.0, 0, 0], [1, 1, 1]])
    cluster = 0.5 + 1e-8 * np.random.rand(300, 3)
    x = np.vstack([corners, cluster])
    n = len(x)
    t = oracle.build_tree(x, np.ones(n) / n, np.full(n, 1e-3))
    assert t.Sizes[n] == 1.0
    assert np.isclose(t.Masses[n], 1.0)
    assert np.array_equal(np.sort(t.TreewalkIndices), np.arange(n))
    assert not t.HasCoincidentPoints
    kids = walk_nodes(t)
    assert max(-np.log2(t.Sizes[no]) for no in kids) > 21  # re-keyed levels exist
    # fully separated: every node holds >= 2 particles and distinct points end in distinct leaves
    for no in kids:
        assert len(subtree_particles(t, no, kids)) >= 2


def test_coincident_points_go_to_buckets():
    """tests/test_nonunique_positions.py:167-213 inputs: duplicates at the origin."""
    x = np.array([[0.0, 0, 0], [0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    t = oracle.build_tree(x, np.repeat(0.2, 5), np.repeat(0.1, 5))
    assert t.HasCoincidentPoints
    assert np.isclose(t.Masses[5], 1.0)
    assert np.array_equal(np.sort(t.TreewalkIndices), np.arange(5))


def test_large_bucket_is_chained_by_eights():
    """BuildBucket (octree.py:1128-1175): >8 coincident points chain 7 leaves + 1 node per level."""
    x = np.vstack([np.zeros((20, 3)), np.eye(3)])
    n = len(x)
    t = oracle.build_tree(x, np.ones(n), np.zeros(n))
    kids = walk_nodes(t)
    chain = [no for no, k in kids.items() if all(c < n for c in k[:7]) and len(k) == 8 and k[7] >= n]
    assert len(chain) >= 2
    assert np.array_equal(np.sort(t.TreewalkIndices), np.arange(n))
    assert np.isclose(t.Masses[n], n)


def test_duplicates_conserve_mass():
    np.random.seed(11)
    x = np.random.rand(1000, 3)
    x = np.vstack([x, x[:50]])
    n = len(x)
    t = oracle.build_tree(x, np.ones(n) / n, np.repeat(0.02, n))
    assert np.isclose(t.Masses[n], 1.0)
    assert np.array_equal(np.sort(t.TreewalkIndices), np.arange(n))
    assert t.HasCoincidentPoints
