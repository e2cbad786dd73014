"""Target groups of the walk (ptg_target_groups), pinned exactly.

The cell-aligned grouping has no reference counterpart, so it is restated here from its definition
with numpy: a new capped cell starts at sorted position t when t starts a system, or when the
smallest cell holding t-1 and t (the cell of the levels their Morton keys share, counted within
t's system) holds more than 4*group_size targets; a capped cell of len targets is cut into
ceil(len / group_size) consecutive chunks of ceil(len / nchunks) targets, the last one shorter.
The vectorised restatement (per-level searchsorted) is checked against a brute-force count of every
cell on small inputs; the GPU output must equal it position for position.  The fixed grouping is
the reference's runs of group_size (grouped_treewalk.py:179-182), restarted at every system."""

import numpy as np
import pytest

from gpu_cases import case


def common_levels(a, b):
    """Whole 3-bit levels (0..21) two 63-bit keys share, elementwise."""
    L = np.zeros(np.shape(a), np.int64)
    for lev in range(1, 22):
        s = np.uint64(63 - 3 * lev)
        L += (a >> s) == (b >> s)  # shared levels are nested, so the count is the deepest one
    return L


def cell_cuts(keys, soff, cap):
    """cut[t]: a capped cell starts at sorted position t (keys sorted within each system)."""
    keys = np.asarray(keys, np.uint64)
    n = len(keys)
    cut = np.zeros(n, bool)
    for s in range(len(soff) - 1):
        a, b = int(soff[s]), int(soff[s + 1])
        if b <= a:
            continue
        k = keys[a:b]
        assert np.all(k[1:] >= k[:-1]), "keys must be sorted within each system"
        cut[a] = True
        if b - a < 2:
            continue
        L = common_levels(k[:-1], k[1:])
        size = np.full(b - a - 1, b - a, np.int64)  # level 0: the whole system
        for lev in range(1, 22):
            sel = np.nonzero(L == lev)[0]
            if len(sel):
                pk = k >> np.uint64(63 - 3 * lev)
                p = pk[sel + 1]
                size[sel] = np.searchsorted(pk, p, "right") - np.searchsorted(pk, p, "left")
        cut[a + 1:b] = size > cap
    return cut


def cell_cuts_brute(keys, soff, cap):
    """The same cuts by counting each pair's cell member by member (small inputs)."""
    keys = [int(k) for k in keys]
    n = len(keys)
    cut = np.zeros(n, bool)
    for s in range(len(soff) - 1):
        a, b = int(soff[s]), int(soff[s + 1])
        for t in range(a, b):
            if t == a:
                cut[t] = True
                continue
            L = 0
            while L < 21 and (keys[t - 1] >> (63 - 3 * (L + 1))) == (keys[t] >> (63 - 3 * (L + 1))):
                L += 1
            sh = 63 - 3 * L
            members = sum(1 for j in range(a, b) if L == 0 or (keys[j] >> sh) == (keys[t] >> sh))
            cut[t] = members > cap
    return cut


def groups_from_cuts(cut, gs):
    """Group starts (with n appended) of near-equal chunks of at most gs in every capped cell."""
    n = len(cut)
    starts = np.append(np.nonzero(cut)[0], n)
    out = []
    for ps, pe in zip(starts[:-1], starts[1:]):
        ln = int(pe - ps)
        nch = -(-ln // gs)
        out.extend(range(int(ps), int(pe), -(-ln // nch)))
    out.append(n)
    return np.array(out, np.int64)


def cell_groups(keys, soff, gs):
    return groups_from_cuts(cell_cuts(keys, soff, 4 * gs), gs)


def fixed_groups(soff, gs):
    out = [np.arange(int(a), int(b), gs) for a, b in zip(soff[:-1], soff[1:])]
    return np.append(np.concatenate(out), soff[-1]).astype(np.int64)


def _clustered_keys(rng, n, depth):
    """Sorted keys drawn from a few nested prefixes, so that cells of every size occur."""
    keys = np.zeros(n, np.uint64)
    for lev in range(depth):
        digit = rng.integers(0, 8 if lev < 3 else 2, n).astype(np.uint64)
        keys |= digit << np.uint64(60 - 3 * lev)
    keys |= rng.integers(0, 1 << 20, n).astype(np.uint64) * (rng.random(n) < 0.5)
    return np.sort(keys)


@pytest.mark.parametrize("seed", range(6))
def test_vectorised_cuts_equal_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 400))
    keys = _clustered_keys(rng, n, int(rng.integers(1, 8)))
    if seed % 2:
        keys[: n // 3] = keys[0]  # a run of identical keys
        keys = np.sort(keys)
    cuts = np.sort(rng.choice(np.arange(1, n), size=min(3, n - 1), replace=False))
    for soff in (np.array([0, n]), np.concatenate([[0], cuts, [n]])):
        # per-system sorting, as the forest's keys are
        k = keys.copy()
        for a, b in zip(soff[:-1], soff[1:]):
            k[a:b] = np.sort(rng.permutation(keys)[: b - a])
        for cap in (4, 12, 40, 256):
            assert np.array_equal(cell_cuts(k, soff, cap), cell_cuts_brute(k, soff, cap))


def test_capped_cells_hold_at_most_cap_and_groups_at_most_gs():
    rng = np.random.default_rng(3)
    keys = _clustered_keys(rng, 5000, 6)
    for gs in (1, 5, 16, 64):
        cut = cell_cuts(keys, np.array([0, 5000]), 4 * gs)
        starts = np.append(np.nonzero(cut)[0], 5000)
        assert np.diff(starts).max() <= 4 * gs
        g = groups_from_cuts(cut, gs)
        assert g[0] == 0 and g[-1] == 5000 and np.all(np.diff(g) >= 1) and np.diff(g).max() <= gs
        # every capped-cell start is a group start
        assert np.all(np.isin(starts, g))


def test_hand_case():
    # levels: keys 0,0 share all 21 levels; the third differs at level 1 from both
    k = np.array([0, 0, 1 << 60, (1 << 60) + 1], np.uint64)
    # cap 4 (gs 1): no cell holds more than 4, so one capped cell of 4 -> four groups of 1
    assert np.array_equal(cell_cuts(k, np.array([0, 4]), 4), [True, False, False, False])
    assert np.array_equal(cell_groups(k, np.array([0, 4]), 1), [0, 1, 2, 3, 4])
    # cap 2 (the whole set is the root cell of 4 > 2): cut at 2, where the root is the smallest
    # common cell; the pair (2, 3) shares 20 levels (a cell of 2)
    assert np.array_equal(cell_cuts(k, np.array([0, 4]), 2), [True, False, True, False])
    assert np.array_equal(fixed_groups(np.array([0, 3, 7]), 2), [0, 2, 3, 5, 7])


# ----------------------------------------------------------------------------------------------
# GPU: the HIP grouping equals the restatement position for position

def _gpu_groups(x, perm, gs, mode, keys=None, tree=None):
    from pytreegrav_amd import engine

    off, cnt = engine.target_groups(x, perm, gs, tree=tree, mode=mode, keys=keys)
    G = int(cnt.item())
    return off.cpu().numpy()[: G + 1].astype(np.int64)


CASES = ["n1", "n2", "n3", "all_identical", "dups", "rekey", "clustered", "uniform_1e4", "dup40_6e4",
         "lattice", "plummer_1e5_h"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("gs", [1, 7, 64, 256])
def test_gpu_cell_groups_equal_restatement(name, gs):
    from pytreegrav_amd import ConstructTree

    x, m, h = case(name)
    print(f"[groups] {name} gs={gs} n={len(x)}", flush=True)
    t = ConstructTree(x, m, h)
    keys = t.dev.keys.cpu().numpy()
    n = len(x)
    ref = cell_groups(keys, np.array([0, n]), gs)
    assert np.array_equal(_gpu_groups(x, t.dev.order, gs, "cells", keys=t.dev.keys), ref)
    # the key pass of the grouping (no keys given) gives the same keys, hence the same groups
    assert np.array_equal(_gpu_groups(x, t.dev.order, gs, "cells"), ref)
    assert np.array_equal(_gpu_groups(x, t.dev.order, gs, "fixed"), fixed_groups(np.array([0, n]), gs))


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1023, 1024, 1025, 2047, 2049, 5000])
def test_gpu_cell_groups_ragged_tiles(n):
    """Sizes around the 1024-target tile, with cells larger than a tile and runs across tiles."""
    from pytreegrav_amd import ConstructTree

    rng = np.random.default_rng(n)
    x = np.vstack([rng.normal(size=(n // 2, 3)) * 1e-3, rng.random((n - n // 2, 3))])
    t = ConstructTree(x, np.full(n, 1.0 / n), np.full(n, 1e-3))
    keys = t.dev.keys.cpu().numpy()
    for gs in (3, 64, 200):
        ref = cell_groups(keys, np.array([0, n]), gs)
        assert np.array_equal(_gpu_groups(x, t.dev.order, gs, "cells", keys=t.dev.keys), ref), gs


@pytest.mark.gpu
def test_gpu_cell_groups_plummer_1e6():
    from pytreegrav_amd import ConstructTree

    x, m, h = case("plummer_1e6")
    t = ConstructTree(x, m, h)
    keys = t.dev.keys.cpu().numpy()
    ref = cell_groups(keys, np.array([0, len(x)]), 64)
    got = _gpu_groups(x, t.dev.order, 64, "cells", keys=t.dev.keys)
    assert len(got) == len(ref) and np.array_equal(got, ref)


@pytest.mark.gpu
@pytest.mark.parametrize("gs", [5, 64])
def test_gpu_forest_groups_equal_restatement(gs):
    """Forest targets: cells are counted within each system and every system start is a cut."""
    from pytreegrav_amd import Forest
    from pytreegrav_amd.sampling import plummer_truncated

    rng = np.random.default_rng(8)
    parts = []
    for k, n in enumerate([2, 3, 300, 1024, 1500, 5000, 70]):
        xs, _ = plummer_truncated(n, seed=40 + k)
        parts.append(xs * 10 ** rng.uniform(-1, 1) + rng.normal(size=3))
    parts.append(case("dups")[0])
    parts.append(case("all_identical")[0])
    x = np.vstack(parts)
    off = np.concatenate([[0], np.cumsum([len(p) for p in parts])])
    n = len(x)
    F = Forest(x, np.full(n, 1.0 / n), np.full(n, 1e-3), off, precision="float32")
    keys = F.tree.keys.cpu().numpy()
    ref = cell_groups(keys, off, gs)
    assert np.array_equal(_gpu_groups(x, F.tree.order, gs, "cells", keys=F.tree.keys, tree=F.tree), ref)
    assert np.array_equal(_gpu_groups(x, F.tree.order, gs, "cells", tree=F.tree), ref)
    assert np.array_equal(_gpu_groups(x, F.tree.order, gs, "fixed", tree=F.tree), fixed_groups(off, gs))


def field_sample_rows(offsets, ngroups, soff, samples, gs):
    """Sorted positions of the field-scale sample targets: per system, groups first + k * count //
    samples of its groups, each padded to gs targets by repeating its last target."""
    starts = np.asarray(offsets[:ngroups], np.int64)
    rows = []
    for s in range(len(soff) - 1):
        g0 = int(np.searchsorted(starts, soff[s], "left"))
        g1 = int(np.searchsorted(starts, soff[s + 1], "left")) if s + 1 < len(soff) - 1 else ngroups
        for k in range(samples):
            g = g0 + (k * (g1 - g0)) // samples
            a, b = int(offsets[g]), int(offsets[g + 1])
            rows.extend(a + min(j, b - a - 1) for j in range(gs))
    return np.array(rows, np.int64)


@pytest.mark.gpu
@pytest.mark.parametrize("samples,gs", [(64, 64), (7, 16)])
def test_gpu_field_sample_equals_restatement(samples, gs):
    """ptg_field_sample gathers the sample targets the restatement picks, bit for bit, for one tree
    and for a forest (per-system groups)."""
    import torch

    from pytreegrav_amd import ConstructTree, Forest, _lib, engine
    from pytreegrav_amd._lib import check, ptr

    x, m, h = case("plummer_1e5_h")
    t = ConstructTree(x, m, h).dev
    n = len(x)
    off, cnt = engine.target_groups(x, t.order, gs, keys=t.keys)
    G = int(cnt.item())
    pos, hh, _, g = engine.field_scale_sample(t, samples, gs, (off, cnt))
    idx = field_sample_rows(off.cpu().numpy(), G, np.array([0, n]), samples, gs)
    coords = t.coordinates[:n].cpu().numpy()
    assert g == gs and np.array_equal(pos.cpu().numpy(), coords[idx])
    assert np.array_equal(hh.cpu().numpy(), t.softenings[:n].cpu().numpy()[idx])

    rng = np.random.default_rng(4)
    parts = [x[:3000] * 0.5, x[3000:3500] + 4.0, x[3500:9000] * 2.0 - 7.0, case("dups")[0]]
    xf = np.vstack(parts)
    soff = np.concatenate([[0], np.cumsum([len(p) for p in parts])])
    nf = len(xf)
    F = Forest(xf, np.full(nf, 1.0 / nf), rng.random(nf) * 0.01, soff, precision="float32")
    goff, gcnt = engine.target_groups(xf, F.tree.order, gs, tree=F.tree, keys=F.tree.keys)
    Gf = int(gcnt.item())
    rows = len(parts) * samples * gs
    pf = torch.empty((rows, 3), dtype=torch.float64, device=engine.device())
    hf = torch.empty(rows, dtype=torch.float64, device=engine.device())
    check(_lib.load().ptg_field_sample(ptr(goff), ptr(gcnt), F.tree.forest_ref, nf, samples, gs,
                                       ptr(F.tree.coordinates), ptr(F.tree.softenings), ptr(pf), ptr(hf),
                                       engine.stream()), "ptg_field_sample")
    idx = field_sample_rows(goff.cpu().numpy(), Gf, soff, samples, gs)
    assert np.array_equal(pf.cpu().numpy(), F.tree.coordinates[:nf].cpu().numpy()[idx])
    assert np.array_equal(hf.cpu().numpy(), F.tree.softenings[:nf].cpu().numpy()[idx])
