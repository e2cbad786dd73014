"""The reference's per-target tree walks (treewalk.py:22-341) and column-density drivers
(:1031-1298) on the GPU.

``PotentialWalk``/``AccelWalk`` and their ``_quad`` forms walk one position through the tree, over
the whole tree or, with ``no``, over the subtree of one element only (the reference stops at
``subtree_end``, treewalk.py:28-34); ``PotentialTarget_tree``/``AccelTarget_tree`` walk many
targets, one walk per target.  All of them run the FP64 walk of libptg_hip.so with the reference's
branches and one target per group (ptg_walk_subtree), so each result is the reference's per-target
sum: the same accepted elements, the same self skip at r = 0, the same softening branches.  The
column-density drivers call the ray-traced and angle-binned walks of colden.hip.
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


# ---- column density drivers (treewalk.py:1031-1298); the per-ray physics is in colden.hip ----

def _targets(pos_target):
    on_dev = _is_dev(pos_target)
    return engine.to_dev(_atleast_2d(_f64(pos_target)), (-1, 3)), on_dev


def _ray_array(rays):
    rays = rays.cpu().numpy() if isinstance(rays, torch.Tensor) else np.asarray(rays, dtype=np.float64)
    if rays.ndim != 2 or rays.shape[1] != 3:
        raise ValueError("rays must be an (N_rays, 3) array of directions")
    return rays


def ColumnDensity_tree(pos_target, tree, rays=None, randomize_rays=False, theta=0.7):
    """Column density from each target to infinity through the tree's particles (uniform spheres
    of radius = softening), (N, N_rays), or (N, 6) sky-bin averages when rays is None
    (treewalk.py:1031-1080).  randomize_rays rotates target i's ray grid by random_rotation(i), i its
    index in pos_target.  Every ray is traced exactly (no opening-angle truncation, as the
    reference's ray walks); theta applies to the binned estimate, one target per traversal."""
    pos, on_dev = _targets(pos_target)
    if rays is None:
        out = engine.column_binned(tree.dev, pos, 1, theta)
    else:
        rays = _ray_array(rays)
        if randomize_rays:
            out = engine.column_rays(tree.dev, pos, engine.random_rotations(rays, pos.shape[0]), per_target=True)
        else:
            out = engine.column_rays(tree.dev, pos, rays)
    return _finish(out, on_dev)


def ColumnDensity_grouped(pos_target, rays, tree, group_size=16):
    """Ray-traced column density, (N, N_rays) (treewalk.py:1087-1191).  The reference's grouped
    traversal is bit-identical to the per-target one at every group size, and so is this: the GPU
    walk forms its own wavefront groups (targets adjacent across each ray), so group_size is only
    checked."""
    if int(group_size) < 1:
        raise ValueError("group_size must be at least 1")
    pos, on_dev = _targets(pos_target)
    return _finish(engine.column_rays(tree.dev, pos, _ray_array(rays)), on_dev)


def ColumnDensityBinned_grouped(pos_target, tree, theta=0.5, group_size=16):
    """Angle-binned column density, (N, 6), one traversal per group of group_size consecutive
    targets (Morton-ordered targets make compact groups; group_size=1 is the per-target walk)
    (treewalk.py:1197-1294)."""
    if int(group_size) < 1:
        raise ValueError("group_size must be at least 1")
    pos, on_dev = _targets(pos_target)
    return _finish(engine.column_binned(tree.dev, pos, int(group_size), theta), on_dev)


ColumnDensity_tree_parallel = ColumnDensity_tree
ColumnDensity_grouped_parallel = ColumnDensity_grouped
ColumnDensityBinned_grouped_parallel = ColumnDensityBinned_grouped
