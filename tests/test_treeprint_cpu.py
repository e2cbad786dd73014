"""The fingerprint helper itself: equal for equal trees, sensitive to a changed node field,
a changed visit order and a dropped subtree (mistakes planted in oracle copies, on the CPU)."""

import copy

import numpy as np
import pytest

import oracle
from treeprint import assert_same_tree, fingerprint


def _tree():
    rng = np.random.default_rng(0)
    x = rng.random((500, 3))
    return oracle.build_tree(x, rng.random(500), rng.random(500) * 0.01, quadrupole=True)


class _Copy:
    def __init__(self, t):
        for k in ("NumParticles", "NumNodes", "FirstSubnode", "NextBranch", "Masses", "Coordinates", "Deltas",
                  "Softenings", "Sizes", "Quadrupoles", "TreewalkIndices"):
            setattr(self, k, copy.deepcopy(getattr(t, k)))


def test_identical_trees_match():
    t = _tree()
    assert_same_tree(t, _Copy(t), with_quads=True)


def test_detects_changed_moment_by_one_ulp():
    t = _tree()
    c = _Copy(t)
    c.Masses[t.NumParticles + 3] = np.nextafter(c.Masses[t.NumParticles + 3], 1.0)
    with pytest.raises(AssertionError):
        assert_same_tree(t, c)


def test_detects_changed_visit_order():
    t = _tree()
    c = _Copy(t)
    c.TreewalkIndices[[0, 1]] = c.TreewalkIndices[[1, 0]]
    with pytest.raises(AssertionError):
        assert_same_tree(t, c)


def test_detects_dropped_child():
    t = _tree()
    c = _Copy(t)
    n = t.NumParticles
    root_first = c.FirstSubnode[n]
    c.FirstSubnode[n] = c.NextBranch[root_first]
    with pytest.raises(AssertionError):
        assert_same_tree(t, c)
    o, nodes = fingerprint(t)
    assert len(o) == n
