"""The reference's per-target tree walks (treewalk.py:22-341) on the GPU.

``PotentialWalk``/``AccelWalk`` and their ``_quad`` forms walk one position through the tree, over
the whole tree or, with ``no``, over the subtree of one element only (the reference stops at
``subtree_end``, treewalk.py:28-34); ``PotentialTarget_tree``/``AccelTarget_tree`` walk many
targets, one walk per target.  All of them run the FP64 walk of libptg_hip.so with the reference's
branches and one target per group (ptg_walk_subtree), so each result is the reference's per-target
sum: the same accepted elements, the same self skip at r = 0, the same softening branches.
"""

import numpy as np
import torch

from . import engine
from .frontend import _atleast_1d, _atleast_2d, _f64, _finish, _is_dev, _zeros_like_len, checkTreeQuadrupoles


def acceptance_criterion(r, h, size, delta, theta):
    """The opening test of one node (treewalk.py:22-25): accept its multipole if r exceeds both
    size/theta + delta and h + 0.6 size + delta."""
    return r > max(size / theta + delta, h + size * 0.6 + delta)


def subtree_end(tree, no):
    """The element a walk started at `no` stops at (treewalk.py:28-34): NextBranch[no], -1 at the root."""
    return tree.NextBranch[no]


def _walk_one(pos, tree, softening, no, theta, mask, quadrupole):
    checkTreeQuadrupoles(tree, quadrupole)
    p = _f64(pos)
    p = p.reshape(1, 3) if isinstance(p, torch.Tensor) else np.asarray(p, dtype=np.float64).reshape(1, 3)
    h = np.full(1, float(softening))
    out = engine.walk_subtree(tree.dev, p, h, no, mask, quadrupole=quadrupole, theta=theta)
    return out[0].cpu().numpy()


def PotentialWalk(pos, tree, softening=0, no=-1, theta=0.7):
    """Potential at one position (3,) from the tree, or from the subtree of element `no` (a particle
    index or a node index NumParticles + k; -1 = the whole tree), G = 1 (treewalk.py:78-118)."""
    return float(_walk_one(pos, tree, softening, no, theta, engine.fields_mask(potential=True), False)[0])


def PotentialWalk_quad(pos, tree, softening=0, no=-1, theta=0.7):
    """PotentialWalk with the nodes' quadrupole moments (treewalk.py:121-165); needs a tree built
    with quadrupole=True."""
    return float(_walk_one(pos, tree, softening, no, theta, engine.fields_mask(potential=True), True)[0])


def AccelWalk(pos, tree, softening=0, no=-1, theta=0.7):
    """Acceleration (3,) at one position from the tree or the subtree of `no`, G = 1
    (treewalk.py:168-217)."""
    return _walk_one(pos, tree, softening, no, theta, engine.fields_mask(accel=True), False)


def AccelWalk_quad(pos, tree, softening=0, no=-1, theta=0.7):
    """AccelWalk with the nodes' quadrupole moments (treewalk.py:220-277)."""
    return _walk_one(pos, tree, softening, no, theta, engine.fields_mask(accel=True), True)


def _target_tree(pos_target, softening_target, tree, G, theta, quadrupole, mask):
    checkTreeQuadrupoles(tree, quadrupole)
    on_dev = _is_dev(pos_target)
    pos_target = _atleast_2d(_f64(pos_target))
    if softening_target is None:
        softening_target = _zeros_like_len(pos_target)
    softening_target = _atleast_1d(_f64(softening_target))
    out = engine.walk_subtree(tree.dev, pos_target, softening_target, -1, mask, quadrupole=quadrupole, theta=theta,
                              G=G)
    return _finish(out, on_dev)


def PotentialTarget_tree(pos_target, softening_target, tree, G=1.0, theta=0.7, quadrupole=False):
    """Potential (N,) at the target positions (N, 3), one walk per target (treewalk.py:280-304)."""
    return _target_tree(pos_target, softening_target, tree, G, theta, quadrupole,
                        engine.fields_mask(potential=True))[:, 0]


def AccelTarget_tree(pos_target, softening_target, tree, G=1.0, theta=0.7, quadrupole=False):
    """Acceleration (N, 3) at the target positions, one walk per target (treewalk.py:311-337);
    softening_target None means zero."""
    return _target_tree(pos_target, softening_target, tree, G, theta, quadrupole, engine.fields_mask(accel=True))


# the reference compiles a thread-parallel copy of each; every GPU walk is parallel
PotentialTarget_tree_parallel = PotentialTarget_tree
AccelTarget_tree_parallel = AccelTarget_tree
