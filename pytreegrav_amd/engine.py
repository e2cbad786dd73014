"""Device-side engine: build, pack, walk and direct sum on the GPU through libptg_hip.

torch is used only to hold device memory and to name the current HIP stream; every computation
is a libptg_hip kernel.  Inputs may be numpy arrays (copied to the device) or torch tensors already
on the GPU (used in place).
"""

import numpy as np
import torch

from . import _lib
from ._lib import PTG_ACC, PTG_F32, PTG_F64, PTG_POT, PTG_TIDAL, BuildInfo, TreeArrays, WalkTree, check, ptr


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("pytreegrav_amd needs a ROCm GPU: torch.cuda.is_available() is False")
    return torch.device("cuda", torch.cuda.current_device())


def stream():
    return _lib._vp(torch.cuda.current_stream().cuda_stream)


def to_dev(a, shape=None):
    """float64 contiguous device tensor from a numpy array or torch tensor."""
    if isinstance(a, torch.Tensor):
        t = a.to(device=device(), dtype=torch.float64)
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64))).to(device())
    t = t.contiguous()
    if shape is not None:
        t = t.reshape(shape)
    return t


def workspace(nbytes):
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device())


def fields_mask(potential=False, accel=False, tidal=False):
    return (PTG_POT if potential else 0) | (PTG_ACC if accel else 0) | (PTG_TIDAL if tidal else 0)


def field_width(mask):
    return (1 if mask & PTG_POT else 0) + (3 if mask & PTG_ACC else 0) + (9 if mask & PTG_TIDAL else 0)


def unpack_fields(out, mask):
    """Split an (N, W) buffer into the reference's dict of views (grouped_treewalk.py:387-397)."""
    res = {}
    o = 0
    if mask & PTG_POT:
        res["potential"] = out[:, 0]
        o = 1
    if mask & PTG_ACC:
        res["accel"] = out[:, o:o + 3]
        o += 3
    if mask & PTG_TIDAL:
        res["tidal"] = out[:, o:o + 9].reshape(-1, 3, 3)
    return res


# ---------------------------------------------------------------------------------------------
# tree build
# ---------------------------------------------------------------------------------------------
class NonFiniteInput(ValueError):
    """A position, mass or softening passed to the build was NaN or infinite."""


class DeviceTree:
    """Device arrays of one octree in the reference's element layout (octree.py:8-28).

    offsets (optional, nsys + 1 increasing integers from 0 to n): build a forest of nsys independent
    systems, system s being particles [offsets[s], offsets[s+1]) (each needs >= 2 particles).  Every
    system gets its own cube, keys and tree, exactly as if it were built alone; particles stay
    system-major in Morton order and the root of system s sits at walk slot s."""

    def __init__(self, pos, m, h, quadrupole=False, compute_moments=True, offsets=None, prepack=True, capacity=None):
        L = _lib.load()
        dev = device()
        pos = to_dev(pos, (-1, 3))
        n = pos.shape[0]
        m = to_dev(m, (n,))
        h = to_dev(h, (n,))
        s = stream()
        self.forest = None
        self.nsys = 1
        self.deferred = capacity is not None
        if self.deferred:
            return self._build_deferred(L, dev, pos, m, h, n, quadrupole, compute_moments, offsets, prepack, capacity)
        if offsets is not None:
            if isinstance(offsets, torch.Tensor):
                off = offsets.to(device=dev, dtype=torch.int64).contiguous()
            else:
                off = torch.from_numpy(np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))).to(dev)
            nsys = off.numel() - 1
            if nsys < 1:
                raise ValueError("offsets must hold nsys + 1 >= 2 entries")
            self.nsys = nsys
            self.sys_offsets = off
            self.roots = torch.empty(nsys, dtype=torch.int32, device=dev)
            self.cubes = torch.empty((nsys, 4), dtype=torch.float64, device=dev)
            self.forest = _lib.Forest(nsys, ptr(off).value, ptr(self.roots).value, ptr(self.cubes).value)
        fref = _lib.ctypes.byref(self.forest) if self.forest is not None else None
        ws = workspace(L.ptg_build_workspace_bytes(n, self.nsys))
        info = BuildInfo()
        st = L.ptg_build_sort(ptr(pos), ptr(m), ptr(h), n, fref, ptr(ws), ws.numel(), _lib.ctypes.byref(info), s)
        if st != 0 and info.nonfinite:
            raise NonFiniteInput("non-finite position, mass or softening")
        check(st, "ptg_build_sort")
        nn = int(info.nnodes)
        ne = n + nn
        f64 = dict(dtype=torch.float64, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        self.n, self.nnodes, self.ne = n, nn, ne
        self.quadrupole = bool(quadrupole)
        self.info = info
        self.coordinates = torch.empty((ne, 3), **f64)
        self.masses = torch.empty(ne, **f64)
        self.softenings = torch.empty(ne, **f64)
        self.sizes = torch.empty(ne, **f64)
        self.deltas = torch.empty(ne, **f64)
        self.quadrupoles = torch.empty((ne, 9), **f64) if quadrupole else None
        self.first_subnode = torch.empty(ne, **i32)
        self.next_branch = torch.empty(ne, **i32)
        self.order = torch.empty(n, dtype=torch.int64, device=dev)
        self.nchild = torch.empty(ne, **i32)
        self.child_base = torch.empty(ne, **i32)
        self.slot = torch.empty(ne, **i32)
        self.keys = torch.empty(n, dtype=torch.int64, device=dev)  # 63-bit Morton keys of the particle slots
        # per-node child statistics for the learned criterion's features (trees with quadrupoles)
        stats = bool(quadrupole) and bool(compute_moments)
        self.node_stats = torch.empty((nn, 5), **f64) if stats else None
        self.arrays = TreeArrays(
            ptr(self.coordinates).value, ptr(self.masses).value, ptr(self.softenings).value, ptr(self.sizes).value,
            ptr(self.deltas).value, ptr(self.quadrupoles).value if quadrupole else None,
            ptr(self.first_subnode).value, ptr(self.next_branch).value, ptr(self.order).value,
            ptr(self.nchild).value, ptr(self.child_base).value, ptr(self.slot).value, ptr(self.keys).value,
            ptr(self.node_stats).value if stats else None)
        nws = workspace(L.ptg_node_workspace_bytes(n, nn, int(info.max_level)))
        # recorded by the build once the particle slots exist (order, keys): target groups can start
        # on a side stream while the nodes and moments are still being built
        self.particles_ready = torch.cuda.Event()
        self.particles_ready.record()  # creates the event; the build records it again
        self._packs = {}
        # prepack: the build also writes the FP32 monopole walk copy, in slot order (cheaper than a
        # separate ptg_pack_walk, whose stores are scattered)
        pk = self._new_pack(PTG_F32, False) if prepack else None
        check(L.ptg_build_tree(ptr(pos), ptr(m), ptr(h), _lib.ctypes.byref(info), fref, int(quadrupole),
                               int(compute_moments), ptr(ws), ws.numel(), ptr(nws), nws.numel(),
                               _lib.ctypes.byref(self.arrays), _lib.ctypes.byref(pk["wt"]) if pk else None,
                               _lib._vp(self.particles_ready.cuda_event), s),
              "ptg_build_tree")
        if pk:
            self._packs[(PTG_F32, False)] = pk
        # ws / nws go back to torch's stream-ordered caching allocator here; later allocations on
        # this stream are ordered after the build kernels that use them

    def _build_deferred(self, L, dev, pos, m, h, n, quadrupole, compute_moments, offsets, prepack, capacity):
        """ptg_build_deferred: no host synchronisation (capturable).  capacity = (node capacity, level
        capacity); the node arrays and the walk copy are sized for n + node capacity, the build writes
        the node count to self.node_count (device) and its counts to self.counts_dev.  A single tree is
        built as a one-system forest, so its cube stays on the device.  self.nnodes is the capacity;
        check() reads the counts and raises if the tree did not fit."""
        nn_cap, level_cap = int(capacity[0]), int(capacity[1])
        if offsets is None:
            offsets = [0, n]
        off = offsets.to(device=dev, dtype=torch.int64).contiguous() if isinstance(offsets, torch.Tensor) else \
            torch.from_numpy(np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))).to(dev)
        nsys = off.numel() - 1
        self.nsys = nsys
        self.sys_offsets = off
        self.roots = torch.empty(nsys, dtype=torch.int32, device=dev)
        self.cubes = torch.empty((nsys, 4), dtype=torch.float64, device=dev)
        self.forest = _lib.Forest(nsys, ptr(off).value, ptr(self.roots).value, ptr(self.cubes).value)
        ne = n + nn_cap
        f64 = dict(dtype=torch.float64, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        self.n, self.nnodes, self.ne = n, nn_cap, ne
        self.level_cap = level_cap
        self.quadrupole = bool(quadrupole)
        self.coordinates = torch.empty((ne, 3), **f64)
        self.masses = torch.empty(ne, **f64)
        self.softenings = torch.empty(ne, **f64)
        self.sizes = torch.empty(ne, **f64)
        self.deltas = torch.empty(ne, **f64)
        self.quadrupoles = torch.empty((ne, 9), **f64) if quadrupole else None
        self.first_subnode = torch.empty(ne, **i32)
        self.next_branch = torch.empty(ne, **i32)
        self.order = torch.empty(n, dtype=torch.int64, device=dev)
        self.nchild = torch.empty(ne, **i32)
        self.child_base = torch.empty(ne, **i32)
        self.slot = torch.empty(ne, **i32)
        self.keys = torch.empty(n, dtype=torch.int64, device=dev)
        stats = bool(quadrupole) and bool(compute_moments)
        self.node_stats = torch.empty((nn_cap, 5), **f64) if stats else None
        self.node_count = torch.zeros(1, **i32)
        self.counts_dev = torch.zeros(_lib.ctypes.sizeof(_lib.BuildCounts), dtype=torch.uint8, device=dev)
        self.arrays = TreeArrays(
            ptr(self.coordinates).value, ptr(self.masses).value, ptr(self.softenings).value, ptr(self.sizes).value,
            ptr(self.deltas).value, ptr(self.quadrupoles).value if quadrupole else None,
            ptr(self.first_subnode).value, ptr(self.next_branch).value, ptr(self.order).value,
            ptr(self.nchild).value, ptr(self.child_base).value, ptr(self.slot).value, ptr(self.keys).value,
            ptr(self.node_stats).value if stats else None, ptr(self.node_count).value)
        info = BuildInfo()
        info.n, info.nnodes, info.max_level = n, nn_cap, level_cap
        self.info = info
        ws = workspace(L.ptg_build_workspace_bytes(n, nsys))
        nws = workspace(L.ptg_node_workspace_bytes(n, nn_cap, level_cap))
        # the event that lets the target groups start on a side stream once the particle slots are
        # placed; none inside a graph capture (the captured step runs on one stream)
        self.particles_ready = None
        if not torch.cuda.is_current_stream_capturing():
            self.particles_ready = torch.cuda.Event()
            self.particles_ready.record()  # creates the event; the build records it again
        self._packs = {}
        pk = self._new_pack(PTG_F32, False) if prepack else None
        check(L.ptg_build_deferred(ptr(pos), ptr(m), ptr(h), n, _lib.ctypes.byref(self.forest), int(quadrupole),
                                   int(compute_moments), nn_cap, level_cap, ptr(ws), ws.numel(), ptr(nws), nws.numel(),
                                   _lib.ctypes.byref(self.arrays), _lib.ctypes.byref(pk["wt"]) if pk else None,
                                   ptr(self.counts_dev),
                                   _lib._vp(self.particles_ready.cuda_event) if self.particles_ready else None,
                                   stream()), "ptg_build_deferred")
        if pk:
            self._packs[(PTG_F32, False)] = pk

    def counts(self):
        """The counts of a deferred build (synchronises): a _lib.BuildCounts."""
        host = self.counts_dev.cpu().numpy().tobytes()
        return _lib.BuildCounts.from_buffer_copy(host)

    def check(self):
        """Raise if a deferred build did not fit its capacities or saw non-finite input."""
        c = self.counts()
        if c.nonfinite:
            raise NonFiniteInput("non-finite position, mass or softening")
        if c.overflow:
            why = [w for b, w in ((_lib.PTG_OVERFLOW_NODES, f"{c.nnodes} nodes > capacity {self.nnodes}"),
                                  (_lib.PTG_OVERFLOW_LEVELS, f"level {c.max_level} > capacity {self.level_cap}"),
                                  (_lib.PTG_OVERFLOW_COLLISIONS,
                                   f"more than {_lib.PTG_DEFERRED_COLLISION_CAP} particles share 63-bit keys"))
                   if c.overflow & b]
            raise _lib.PtgError("deferred tree build over capacity: " + "; ".join(why))
        return c

    @property
    def forest_ref(self):
        """ctypes reference to the forest descriptor (None for a single tree)."""
        return _lib.ctypes.byref(self.forest) if self.forest is not None else None

    @property
    def center(self):
        if getattr(self, "deferred", False):
            return np.array(self.counts().center[:])
        return np.array(self.info.center[:])

    @property
    def size(self):
        if getattr(self, "deferred", False):
            return float(self.counts().size)
        return float(self.info.size)

    @property
    def has_coincident(self):
        return bool(self.info.has_coincident)

    def pack_learned(self, model, eps, quadrupole=False, bf16=False, scale=None):
        """FP32 walk copy whose node slots carry the learned opening radii at tolerance eps (times
        the device scalar `scale` if given) (ptg_open_radii); cached per (model, tolerance, moment
        order, bf16)."""
        from .learned import open_radii

        # the scale tensor is kept in the cache entry, so its address cannot be handed to another
        # tensor while the entry exists
        skey = None if scale is None else scale.data_ptr()
        key = ("learned", model.key(), float(eps), skey, bool(quadrupole), bool(bf16))
        if key not in self._packs:
            # take over the plain FP32 pack if there is one (its aux[3] is overwritten, so it leaves
            # the plain cache and is rebuilt there when needed again)
            p = self._packs.pop((PTG_F32, bool(quadrupole)), None) or self.pack(PTG_F32, quadrupole, cache=False,
                                                                                 norms=False)
            open_radii(self, p, model, eps, bf16=bf16, scale=scale)
            for k in [k for k in self._packs if k[0] == "learned"]:  # one learned pack per tree
                del self._packs[k]
            p["scale_ref"] = scale
            self._packs[key] = p
        return self._packs[key]

    def _new_pack(self, prec, quadrupole):
        """Device buffers and descriptor of a walk copy (filled by ptg_pack_walk or by the build)."""
        dev = device()
        dt = torch.float32 if prec == PTG_F32 else torch.float64
        ne = self.ne
        xyzm = torch.empty((ne, 4), dtype=dt, device=dev)
        aux = torch.empty((ne, 4), dtype=dt, device=dev)
        kids = torch.empty((ne, 2), dtype=torch.int32, device=dev)
        quad = torch.empty((ne, 8), dtype=dt, device=dev) if quadrupole else None
        # a forest (every deferred tree is one) ignores the origin: FP32 slots are relative to each
        # system's cube on the device
        origin = self.center if prec == PTG_F32 and self.forest is None else np.zeros(3)
        wt = WalkTree(ptr(xyzm).value, ptr(aux).value, ptr(kids).value, ptr(quad).value if quad is not None else None,
                      ne, prec, int(bool(quadrupole)), (_lib._dbl * 3)(*origin))
        return {"wt": wt, "tensors": (xyzm, aux, kids, quad), "norms": True}

    def pack(self, prec, quadrupole, cache=True, norms=None):
        """Sibling-contiguous walk copy (cached per precision and moment order).  aux[3] holds each
        node's quadrupole norm, which only the error-bounded criterion reads: norms=True demands it,
        norms=False skips it (aux[3] = -1, and the quadrupoles are not read) for a pack whose aux[3]
        will carry learned opening radii, None takes either."""
        key = (prec, bool(quadrupole))
        if cache and key in self._packs and (norms is not True or self._packs[key]["norms"]):
            return self._packs[key]
        if getattr(self, "deferred", False):
            self.check()  # a deferred tree over its capacity has no nodes to copy (synchronises)
        norms = norms is not False or bool(quadrupole)
        p = self._new_pack(prec, quadrupole)
        wt = p["wt"]
        arrays = self.arrays
        if not norms and arrays.quadrupoles:
            arrays = TreeArrays.from_buffer_copy(arrays)
            arrays.quadrupoles = None
        check(_lib.load().ptg_pack_walk(_lib.ctypes.byref(arrays), self.n, self.nnodes, self.forest_ref,
                                        _lib.ctypes.byref(wt), stream()), "ptg_pack_walk")
        p["norms"] = norms or not self.arrays.quadrupoles
        if cache:
            self._packs[key] = p
        return p


class ResidentBuilder:
    """Tree builds of a particle set that moves from step to step (same count and systems), without
    host synchronisation inside the build: the first call builds with exact counts (two host
    reads), later calls build with ptg_build_deferred into arrays sized by the last known node count
    and depth times `slack` (plus a few levels).  After the step that used a tree, verify(tree) reads
    its counts (one host read): if the tree did not fit, it raises TreeOverflow after growing the
    capacities, and the caller redoes the step; a tree that fit updates the capacities it is sized
    from."""

    class TreeOverflow(RuntimeError):
        pass

    def __init__(self, quadrupole=False, offsets=None, slack=1.1, extra_levels=4):
        self.quadrupole, self.offsets = bool(quadrupole), offsets
        self.slack, self.extra_levels = float(slack), int(extra_levels)
        self.capacity = None

    def _size(self, nnodes, max_level):
        self.capacity = (int(nnodes * self.slack) + 1024, min(255, int(max_level) + self.extra_levels))

    def __call__(self, pos, m, h):
        if self.capacity is None:
            t = DeviceTree(pos, m, h, quadrupole=self.quadrupole, offsets=self.offsets)
            self._size(t.nnodes, t.info.max_level)
            return t
        return DeviceTree(pos, m, h, quadrupole=self.quadrupole, offsets=self.offsets, capacity=self.capacity)

    def verify(self, tree):
        if not getattr(tree, "deferred", False):
            return tree
        c = tree.counts()
        if c.nonfinite:
            raise NonFiniteInput("non-finite position, mass or softening")
        if c.overflow:
            # the counts report the tree's true node count and depth even when they did not fit
            self._size(max(c.nnodes, self.capacity[0]), max(c.max_level, self.capacity[1]))
            if c.overflow & _lib.PTG_OVERFLOW_COLLISIONS:
                self.capacity = None  # beyond the deferred build's re-keying: build with exact counts
            raise ResidentBuilder.TreeOverflow("tree over its capacities; capacities grown, redo the step")
        if c.nnodes * self.slack > self.capacity[0] or c.max_level + self.extra_levels > self.capacity[1]:
            self._size(c.nnodes, c.max_level)
        return tree


def morton_order(pos):
    """Device int64 permutation sorting points into Morton order (grouped_treewalk.py:269-275)."""
    L = _lib.load()
    pos = to_dev(pos, (-1, 3))
    n = pos.shape[0]
    order = torch.empty(n, dtype=torch.int64, device=device())
    if n:
        ws = workspace(L.ptg_build_workspace_bytes(n, 1))
        check(L.ptg_morton_order(ptr(pos), n, ptr(order), ptr(ws), ws.numel(), stream()), "ptg_morton_order")
    return order


_walk_ws = {}


def target_groups(pos, perm, group_size, tree=None, mode="cells", keys=None):
    """Groups of the Morton-sorted targets pos[perm[t]]: (offsets, count) device int32 tensors for
    ptg_walk (ptg_target_groups).  mode "cells": cell-aligned groups; "fixed": runs of group_size.
    tree: the forest whose own particles are the targets (groups then respect its systems).
    keys: the targets' sorted Morton keys if known (tree.keys for a tree's own particles)."""
    L = _lib.load()
    pos = to_dev(pos, (-1, 3))
    n = pos.shape[0]
    offsets = torch.empty(n + 1, dtype=torch.int32, device=device())
    count = torch.empty(1, dtype=torch.int32, device=device())
    fref = tree.forest_ref if tree is not None else None
    nsys = tree.nsys if tree is not None else 1
    if mode not in ("cells", "fixed"):
        raise ValueError(f"mode must be 'cells' or 'fixed', got {mode!r}")
    m = _lib.PTG_GROUPS_CELLS if mode == "cells" else _lib.PTG_GROUPS_RUNS
    ws = workspace(L.ptg_groups_workspace_bytes(n, nsys))
    check(L.ptg_target_groups(ptr(pos), ptr(perm), ptr(keys), n, fref, int(group_size), m, ptr(offsets), ptr(count),
                              ptr(ws), ws.numel(), stream()), "ptg_target_groups")
    return offsets, count


def walk(tree, pos, soft, perm, mask, quadrupole=False, group_size=8, theta=0.7, G=1.0, precision="float64",
         counters=None, grouping=None, groups=None, criterion="theta", eps=None, model=None, bf16=False,
         eps_scale=None, sorted_targets=False, expansion=None, variant=0, expansion_rel=0.0,
         _events_end=None, _check=True, _ws_slot=0):
    """Grouped walk of `tree` at targets pos[perm[t]]; returns an (nt, W) float64 device tensor in the
    original target order (FieldsTarget_grouped, grouped_treewalk.py:350-397).

    grouping: "fixed" = runs of group_size sorted targets (the reference's groups), "cells" =
    cell-aligned groups (ptg_target_groups); default "fixed" for float64, "cells" for float32.
    groups: precomputed (offsets, count) from target_groups, overriding grouping.
    criterion: "theta" (the reference's geometric test) or "error" (FP32 only: also bound the
    dropped quadrupole term 3.5|Q|/r^4 by eps, an absolute acceleration tolerance for G=1; theta is
    then a loose cap; the tree must carry quadrupoles), or "learned" (FP32 only: accept a node only
    beyond the opening radius the MLP `model` predicts for tolerance eps; theta is the guard; bf16
    runs the MLP with bfloat16 weights; eps_scale: optional device float64 scalar multiplying eps).
    sorted_targets: pos and soft are already in the sorted order (pos[t] is target perm[t]), as the
    tree's own coordinates and softenings are for self gravity; perm then only places the output.
    expansion: tolerance (absolute acceleration, G=1, times eps_scale when given) of the group
    expansion: accepted elements far enough from a group's centre are folded into a third-order
    Taylor expansion about it instead of being evaluated per target (FP32, monopole interactions,
    potential and/or acceleration).  expansion_rel: also fold an
    element only when its truncation bound is below expansion_rel x its own field M / D^2 at the group
    centre (a cap on R / D; 0 = none).  variant: the FP32 walk
    kernel's occupancy / interaction-list variant (an index into WALK_VARIANTS; the autotuner picks
    it with group_size="auto"); results depend on it at most at the FP32 rounding level."""
    L = _lib.load()
    pos = to_dev(pos, (-1, 3))
    nt = pos.shape[0]
    soft = to_dev(soft, (nt,))
    W = field_width(mask)
    out = torch.empty((nt, W), dtype=torch.float64, device=device())
    if nt == 0:
        return out
    prec = PTG_F32 if precision in ("float32", "f32", np.float32, torch.float32) else PTG_F64
    if group_size == "auto":
        if groups is not None:
            raise ValueError("group_size='auto' cannot be combined with precomputed groups (they were formed for "
                             "one group size; a walk never splits a supplied group)")
        rec = autotune_group_size(tree, pos, soft, perm, mask, quadrupole=quadrupole, theta=theta,
                                  precision=precision, grouping=grouping, criterion=criterion, eps=eps,
                                  model=model, bf16=bf16, expansion=expansion, eps_scale=eps_scale,
                                  expansion_rel=expansion_rel)
        group_size, variant = rec["choice"], rec["variant"]
    if not 0 <= int(variant) < len(WALK_VARIANTS):
        raise ValueError(f"variant must index WALK_VARIANTS (0..{len(WALK_VARIANTS) - 1}), got {variant!r}")
    if grouping is None:
        grouping = "cells" if prec == PTG_F32 else "fixed"
    if grouping not in ("fixed", "cells"):
        raise ValueError(f"grouping must be 'fixed' or 'cells', got {grouping!r}")
    forest = tree.nsys > 1 or tree.forest is not None
    if tree.nsys > 1 and (nt != tree.n or perm.data_ptr() != tree.order.data_ptr()):
        raise ValueError("a forest is walked at its own particles (perm = tree.order)")
    if groups is None and (grouping == "cells" or tree.nsys > 1):
        own = nt == tree.n and perm.data_ptr() == tree.order.data_ptr()
        groups = target_groups(pos, perm, group_size, tree=tree if forest else None, mode=grouping,
                               keys=tree.keys if own else None)
    goff, gcnt = groups if groups is not None else (None, None)
    crit = _lib.Criterion(_lib.PTG_CRIT_THETA, 0, 0.0)
    crit.variant = int(variant)
    if criterion == "error":
        if prec != PTG_F32:
            raise ValueError("criterion='error' is available for the float32 walk")
        if not tree.quadrupole:
            raise ValueError("criterion='error' needs a tree built with quadrupole=True (it bounds the quadrupole term)")
        if eps is None or not eps > 0:
            raise ValueError("criterion='error' needs eps > 0")
        crit = _lib.Criterion(_lib.PTG_CRIT_QERR, 0, float(eps))
        crit.variant = int(variant)
    elif criterion == "learned":
        if prec != PTG_F32:
            raise ValueError("criterion='learned' is available for the float32 walk")
        if not tree.quadrupole:
            raise ValueError("criterion='learned' needs a tree built with quadrupole=True (a feature is |Q|)")
        if model is None:
            raise ValueError("criterion='learned' needs model= (pytreegrav_amd.learned.OpeningMLP)")
        if eps is None or not eps > 0:
            raise ValueError("criterion='learned' needs eps > 0")
        crit = _lib.Criterion(_lib.PTG_CRIT_RADIUS, 0, float(eps))
        crit.variant = int(variant)
    elif criterion != "theta":
        raise ValueError(f"criterion must be 'theta', 'error' or 'learned', got {criterion!r}")
    if expansion is not None:
        if prec != PTG_F32 or quadrupole or (int(mask) & PTG_TIDAL):
            raise ValueError("the group expansion runs on the float32 walk with monopole interactions, "
                             "for the potential and/or the acceleration")
        if not expansion > 0:
            raise ValueError("expansion must be a positive tolerance")
        if eps_scale is not None and eps_scale.numel() < (tree.nsys if forest else 1):
            raise ValueError("eps_scale needs one value per system of the forest")
        if not expansion_rel >= 0:
            raise ValueError("expansion_rel must be >= 0")
        crit.group_expansion = 1
        crit.expansion_eps = float(expansion)
        crit.expansion_rel = float(expansion_rel)
        crit.eps_scale = eps_scale.data_ptr() if eps_scale is not None else None
    if criterion == "learned":
        p = tree.pack_learned(model, eps, quadrupole, bf16, scale=eps_scale)
    else:
        p = tree.pack(prec, quadrupole, norms=True if criterion == "error" else None)
    wsb = L.ptg_walk_workspace_bytes(nt, int(group_size))
    ws = _walk_workspace(wsb, _ws_slot)
    check(L.ptg_walk(_lib.ctypes.byref(p["wt"]), tree.forest_ref if forest else None, ptr(pos), ptr(soft), ptr(perm),
                     int(bool(sorted_targets)), nt, int(mask), int(bool(quadrupole)),
                     int(group_size), ptr(goff), ptr(gcnt), float(theta), _lib.ctypes.byref(crit), float(G),
                     int(prec == PTG_F64), ptr(out),
                     ptr(counters), ptr(ws), ws.numel(), stream()), "ptg_walk")
    if _events_end is not None:
        _events_end.record()
    if _check:  # synchronises; a caller that skips it calls check_walk(_ws_slot) later
        check(L.ptg_walk_check(ptr(ws), stream()), "ptg_walk")
    return out


def _walk_workspace(nbytes, slot=0):
    """Walk workspace per (device, stream, slot), reused across walks on that stream; its group
    counters, spill stacks and overflow flag belong to the last walk run on it, so walks on
    different streams never share one."""
    key = (torch.cuda.current_device(), torch.cuda.current_stream().cuda_stream, slot)
    ws = _walk_ws.get(key)
    if ws is None or ws.numel() < nbytes:
        ws = workspace(nbytes)
        _walk_ws[key] = ws
    return ws


def check_walk(slot=0):
    """Overflow check of the last walk run on workspace `slot` (synchronises the stream)."""
    ws = _walk_ws.get((torch.cuda.current_device(), torch.cuda.current_stream().cuda_stream, slot))
    if ws is not None:
        check(_lib.load().ptg_walk_check(ptr(ws), stream()), "ptg_walk")


def walk_subtree(tree, pos, soft, no, mask, quadrupole=False, theta=0.7, G=1.0, counters=None):
    """Per-target FP64 walks of the subtree of element `no` (a particle index or a node index
    NumParticles + k; -1 = the root) at every target pos[t] (ptg_walk_subtree; treewalk.py:78-277
    with the stop at subtree_end, :28-34); (nt, W) float64 device tensor in target order, scaled by G."""
    if tree.nsys > 1 or tree.forest is not None:
        raise ValueError("subtree walks run on a single tree, not a forest")
    no = int(no)
    if no < 0:
        no = tree.n  # the root, as the reference defaults
    if no >= tree.ne:
        raise ValueError(f"no={no} is not an element of the tree (it has {tree.ne})")
    if quadrupole and not tree.quadrupole:
        raise ValueError("a quadrupole walk needs a tree built with quadrupole=True")
    L = _lib.load()
    pos = to_dev(pos, (-1, 3))
    nt = pos.shape[0]
    soft = to_dev(soft, (nt,))
    out = torch.empty((nt, field_width(mask)), dtype=torch.float64, device=device())
    if nt == 0:
        return out
    perm = morton_order(pos) if nt > 1 else torch.zeros(1, dtype=torch.int64, device=device())
    p = tree.pack(PTG_F64, quadrupole)
    start = tree.slot[no:no + 1]  # the element's walk slot, read by the kernel (no host round trip)
    ws = _walk_workspace(L.ptg_walk_workspace_bytes(nt, 1))
    check(L.ptg_walk_subtree(_lib.ctypes.byref(p["wt"]), ptr(start), ptr(pos), ptr(soft), ptr(perm), nt, int(mask),
                             int(bool(quadrupole)), float(theta), float(G), ptr(out), ptr(counters), ptr(ws),
                             ws.numel(), stream()), "ptg_walk_subtree")
    check(L.ptg_walk_check(ptr(ws), stream()), "ptg_walk_subtree")
    return out


_side_streams = {}


def _side_stream():
    d = torch.cuda.current_device()
    if d not in _side_streams:
        _side_streams[d] = torch.cuda.Stream()  # default priority
    return _side_streams[d]


def self_fields(tree, pos, soft, mask, theta=None, G=1.0, quadrupole=False, group_size=64, criterion="theta",
                alpha=None, model=None, bf16=False, counters=None, walk_events=None, expansion=None,
                stats=None, expansion_rel=0.0, variant=0):
    """Self-gravity FP32 fields of the particles a tree was built from, in their input order.

    The cell-aligned target groups are formed on a side stream while the main stream packs the
    tree, runs the field-scale sample walk and (criterion "learned") the radius pass; the walk
    waits for both.  criterion "error"/"learned" take the per-node tolerance alpha x (RMS field)
    and need a tree built with quadrupoles; theta defaults to 0.7 ("theta") or 2.0 (the guard of
    the other two).  expansion: group-expansion tolerance in units of the RMS field (see walk);
    None walks without it.  walk_events: optional (start, end) torch.cuda.Event pair recorded around
    the walk launch alone.  stats: optional dict that receives "groups", the device int32 count of
    the target groups walked.  variant: the FP32 walk kernel variant (WALK_VARIANTS)."""
    L = _lib.load()
    pos = to_dev(pos, (-1, 3))
    n = pos.shape[0]
    soft = to_dev(soft, (n,))
    if n != tree.n:
        raise ValueError("self_fields evaluates the tree's own particles")
    if criterion not in ("theta", "error", "learned"):
        raise ValueError(f"criterion must be 'theta', 'error' or 'learned', got {criterion!r}")
    if theta is None:
        theta = 0.7 if criterion == "theta" else 2.0
    if criterion != "theta":
        if alpha is None or not alpha > 0:
            raise ValueError(f"criterion={criterion!r} needs alpha > 0")
        if criterion == "learned" and model is None:
            from .learned import OpeningMLP

            model = OpeningMLP.load()
    perm = tree.order
    dev = device()
    main = torch.cuda.current_stream()
    side = _side_stream()
    # the side stream starts once the build has written the particle slots (the group pass reads
    # only the tree's order and keys); its buffers come from the side stream's own pool (never
    # memory the build may still be using) and are marked as used by the main stream, which reads
    # the groups later
    side.wait_stream(main) if getattr(tree, "particles_ready", None) is None else side.wait_event(tree.particles_ready)
    with torch.cuda.stream(side):
        offsets = torch.empty(n + 1, dtype=torch.int32, device=dev)
        count = torch.empty(1, dtype=torch.int32, device=dev)
        gws = workspace(L.ptg_groups_workspace_bytes(n, tree.nsys))
        check(L.ptg_target_groups(ptr(pos), ptr(perm), ptr(tree.keys), n, tree.forest_ref, int(group_size),
                                  _lib.PTG_GROUPS_CELLS, ptr(offsets), ptr(count), ptr(gws), gws.numel(), stream()),
              "ptg_target_groups")
        # the field-scale sample targets read only the particle slots, final once the particles are
        # placed, so they are gathered here, off the main stream's path
        sample = field_scale_sample(tree, cells=(offsets, count)) if criterion != "theta" or expansion is not None \
            else None
    offsets.record_stream(main)
    count.record_stream(main)
    if stats is not None:
        stats["groups"] = count
    for s_t in sample[:3] if sample is not None else ():
        s_t.record_stream(main)
    eps, scale, exp_eps = None, None, None
    tree.pack(PTG_F32, quadrupole, norms=criterion != "learned")
    if criterion != "theta" or expansion is not None:
        main.wait_stream(side)  # the sample targets come from the side stream
        if criterion == "error":
            rms = field_scale(tree, sample=sample)
            eps = alpha * rms
            exp_eps = None if expansion is None else expansion * rms
        else:  # the tolerances stay on the device: no host round trip
            scale = field_scale(tree, sample=sample, on_device=True)
            exp_eps = expansion
            if criterion == "learned":
                eps = alpha
                tree.pack_learned(model, eps, quadrupole, bf16, scale=scale)
    main.wait_stream(side)
    if walk_events is not None:
        walk_events[0].record()
    # targets read from the tree's Morton-ordered copies (the same doubles as pos[perm]) instead of
    # gathering pos through the permutation
    out = walk(tree, tree.coordinates[:n], tree.softenings[:n], perm, mask, quadrupole=quadrupole,
               group_size=group_size, theta=theta, G=G, precision="float32", counters=counters,
               groups=(offsets, count), criterion=criterion, eps=eps, model=model, bf16=bf16, eps_scale=scale,
               sorted_targets=True, expansion=exp_eps, expansion_rel=expansion_rel, variant=variant,
               _events_end=walk_events[1] if walk_events is not None else None)
    if scale is not None:
        check_walk(1)  # the field-scale sample walk was not checked when it ran
    return out


def _check_counters(counters, k):
    if counters is not None and (not isinstance(counters, torch.Tensor) or counters.dtype != torch.int64
                                 or not counters.is_cuda or not counters.is_contiguous() or counters.numel() < k):
        raise ValueError(f"counters must be a contiguous int64 device tensor of at least {k} entries")


def _ray_orders(tp, rays):
    """(nr, nt) int64: for each ray u, the targets sorted by the 2D Morton order of their coordinates
    in a plane perpendicular to u (ptg_morton_order with the third coordinate zero)."""
    orders = []
    for u in rays.cpu().numpy():
        u = u / np.linalg.norm(u)
        a = np.array([1.0, 0.0, 0.0]) if abs(u[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
        e1 = np.cross(u, a)
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(u, e1)
        basis = torch.tensor(np.stack([e1, e2, np.zeros(3)], 1), dtype=torch.float64, device=tp.device)
        orders.append(morton_order(tp @ basis))
    return torch.stack(orders).contiguous()


def random_rotations(rays, nt, first_seed=0):
    """(nt, nr, 3) float64 device tensor: the unit rays (nr, 3) rotated, for target k, by the
    rotation of misc.random_rotation(first_seed + k) (ptg_random_rotations)."""
    L = _lib.load()
    r = to_dev(rays, (-1, 3))
    out = torch.empty((nt, r.shape[0], 3), dtype=torch.float64, device=device())
    check(L.ptg_random_rotations(int(first_seed), int(nt), ptr(r), r.shape[0], ptr(out), stream()),
          "ptg_random_rotations")
    return out


def column_rays(tree, targets, rays, per_target=False, counters=None):
    """Ray-traced column density (ptg_column_rays): (nt, nr) float64 device tensor for targets (nt, 3)
    and unit rays (nr, 3), or per-target rays (nt, nr, 3) when per_target.  counters: optional device
    int64 tensor; [0] accumulates the elements visited."""
    L = _lib.load()
    tp = to_dev(targets, (-1, 3))
    nt = tp.shape[0]
    r = to_dev(rays)
    nr = r.shape[-2] if per_target else r.shape[0]
    if (per_target and tuple(r.shape) != (nt, nr, 3)) or (not per_target and (r.dim() != 2 or r.shape[1] != 3)):
        raise ValueError("rays must be (nr, 3), or (nt, nr, 3) with per_target")
    _check_counters(counters, 1)
    out = torch.empty((nt, nr), dtype=torch.float64, device=device())
    err = torch.zeros(1, dtype=torch.int32, device=device())
    snap = None if counters is None else counters.clone()
    # shared rays: per ray, the targets in 2D Morton order of their position across the ray, so the
    # 64 targets of a wavefront walk nearly the same path
    order = _ray_orders(tp, r) if not per_target and nt > 64 else None
    # the wavefront walk over the FP64 pack; a tree deeper than its stack reruns the stackless walk
    for wt in (_lib.ctypes.byref(tree.pack(PTG_F64, False)["wt"]), None):
        check(L.ptg_column_rays(_lib.ctypes.byref(tree.arrays), tree.n, tree.nnodes, wt, ptr(tp), nt, ptr(r), nr,
                                int(bool(per_target)), ptr(order) if wt is not None else None, ptr(out), ptr(err),
                                ptr(counters), stream()), "ptg_column_rays")
        e = int(err.item())
        if e & 1:
            raise RuntimeError("ptg_column_rays: corrupt tree links")
        if not e:
            break
        if snap is not None:
            counters.copy_(snap)
    return out


def column_binned(tree, targets, group_size=16, theta=0.5, counters=None):
    """Angular-binned column density (ptg_column_binned): (nt, 6) float64 device tensor.  counters:
    optional device int64 tensor of 2; accumulates [0] element-target pairs, [1] elements visited."""
    L = _lib.load()
    tp = to_dev(targets, (-1, 3))
    nt = tp.shape[0]
    out = torch.empty((nt, 6), dtype=torch.float64, device=device())
    err = torch.zeros(1, dtype=torch.int32, device=device())
    _check_counters(counters, 2)
    wt, ws = None, None
    if group_size <= 64:  # one wavefront per group over the FP64 walk pack
        wt = _lib.ctypes.byref(tree.pack(PTG_F64, False)["wt"])
        ws = workspace(L.ptg_column_workspace_bytes(int(group_size)))
    check(L.ptg_column_binned(_lib.ctypes.byref(tree.arrays), tree.n, tree.nnodes, wt, ptr(tp), nt, int(group_size),
                              float(theta), ptr(out), ptr(err), ptr(ws), 0 if ws is None else ws.numel(), ptr(counters),
                              stream()),
          "ptg_column_binned")
    e = int(err.item())
    if e:
        raise RuntimeError("ptg_column_binned: " + ("corrupt tree links" if e & 1 else "traversal stack overflow"))
    return out


def direct(pos_t, soft_t, pos_s, m_s, h_s, mask, G=1.0, self_mode=False, precision="float64"):
    """LDS-tiled direct sum (bruteforce.py); (nt, W) float64 device tensor."""
    L = _lib.load()
    pos_t = to_dev(pos_t, (-1, 3))
    nt = pos_t.shape[0]
    soft_t = to_dev(soft_t, (nt,))
    pos_s = to_dev(pos_s, (-1, 3))
    ns = pos_s.shape[0]
    m_s = to_dev(m_s, (ns,))
    h_s = to_dev(h_s, (ns,))
    out = torch.empty((nt, field_width(mask)), dtype=torch.float64, device=device())
    prec = PTG_F32 if precision in ("float32", "f32", np.float32, torch.float32) else PTG_F64
    check(L.ptg_direct(ptr(pos_t), ptr(soft_t), nt, ptr(pos_s), ptr(m_s), ptr(h_s), ns, prec, int(mask),
                       int(bool(self_mode)), float(G), ptr(out), stream()), "ptg_direct")
    return out


def field_scale(tree, groups=64, group_size=64, theta=1.0, cells=None, on_device=False, sample=None):
    """RMS |acceleration| (G=1) of the tree's own particles, estimated by a theta=1 FP32 walk of
    `groups` sample groups of `group_size` Morton-consecutive particles spread evenly over the
    Morton order; turns a relative accuracy target into the absolute tolerance of the error-bounded
    and learned criteria.  cells: the (offsets, count) cell-aligned groups of the tree's own
    particles, if already formed: each sample is then one of those groups (its last target repeated
    to fill group_size), so no sample straddles two distant cells, whose bounding box would make
    its walk open most of the tree.  sample: what field_scale_sample returned, gathered beforehand
    (e.g. on another stream); groups, group_size and cells are then not used.  on_device: return a device float64 scalar instead of a float
    (no host synchronisation)."""
    pos, h, perm, gs = field_scale_sample(tree, groups, group_size, cells) if sample is None else sample
    a = walk(tree, pos, h, perm, PTG_ACC, group_size=gs,
             theta=theta, precision="float32", grouping="fixed", _check=not on_device, _ws_slot=1)
    rms = torch.empty(1, dtype=torch.float64, device=device())
    check(_lib.load().ptg_rms_rows(ptr(a), a.shape[0], 3, 0, 3, 1, ptr(rms), stream()), "ptg_rms_rows")
    return rms if on_device else float(rms.item())


def field_scale_sample(tree, groups=64, group_size=64, cells=None):
    """The field-scale sample targets (see field_scale): (positions, softenings, output rows, group
    size), read from the particle slots [0, n) of the tree only, on the current stream."""
    n = tree.n
    k = min(groups, max(1, n // group_size))
    gs = min(group_size, n)
    j = torch.arange(gs, device=device(), dtype=torch.int64)[None, :]
    if cells is not None:
        # groups k * count / groups of the cell-aligned groups, each padded to gs targets
        off, cnt = cells
        pos = torch.empty((k * gs, 3), dtype=torch.float64, device=device())
        h = torch.empty(k * gs, dtype=torch.float64, device=device())
        check(_lib.load().ptg_field_sample(ptr(off), ptr(cnt), None, n, k, gs, ptr(tree.coordinates),
                                           ptr(tree.softenings), ptr(pos), ptr(h), stream()), "ptg_field_sample")
        return pos, h, torch.arange(k * gs, device=device()), gs
    else:
        starts = (torch.arange(k, device=device(), dtype=torch.int64) * (n // k))
        idx = (starts[:, None] + j).reshape(-1)
    pos = tree.coordinates.index_select(0, idx)  # particle slots [0, n) are in Morton order
    h = tree.softenings.index_select(0, idx)
    return pos, h, torch.arange(idx.shape[0], device=device()), gs


# ---------------------------------------------------------------------------------------------
# autotuner: target-group size x walk kernel variant
# ---------------------------------------------------------------------------------------------
TUNE_SIZES = (32, 64, 128, 256)
# FP32 walk kernel variants (ptg_criterion.variant): waves per SIMD the one-target-per-lane kernel is
# compiled for, and the interaction-list length at which it evaluates the lists (LDS per wave)
WALK_VARIANTS = ("5 waves/SIMD, lists of 64", "4 waves/SIMD, lists of 64", "6 waves/SIMD, lists of 32")
_tuned = {}


def _bucket(n):
    return int(round(np.log2(max(n, 1)) * 2))  # half-octave buckets of the problem size


def distribution_key(tree, pos):
    """Coarse statistics of the sources and targets that the best group size and kernel variant
    depend on: tree depth, nodes per particle (in eighths), and how far the targets spread beyond
    the source cube (the log2 of their extent over its side, rounded)."""
    depth = int(tree.info.max_level)
    npp = int(round(8.0 * tree.nnodes / max(tree.n, 1)))
    ext = (pos.max(0).values - pos.min(0).values).max().item()
    spread = int(round(np.log2(max(ext, 1e-300) / max(tree.size, 1e-300)))) if tree.size > 0 else 0
    return depth, npp, spread


def tuning_log():
    """Every autotuning decision taken in this process: {key: {"choice", "variant", "ms", "interactions"}},
    "ms" keyed by "group_size/variant"."""
    return dict(_tuned)


def autotune_group_size(tree, pos, soft, perm, mask, quadrupole=False, theta=0.7, precision="float32", grouping=None,
                        criterion="theta", eps=None, model=None, bf16=False, sizes=TUNE_SIZES, reps=2, force=False,
                        expansion=None, eps_scale=None, expansion_rel=0.0):
    """Pick the walk's target-group size and kernel variant by timing the walk (groups included).

    Candidates: every group size in `sizes`; for the FP32 one-target-per-lane kernels (groups of
    at most 64) also every WALK_VARIANTS entry (occupancy and interaction-list LDS).  The choice
    depends on the target count, the tree and the opening criterion, not on the values the walk
    returns: every candidate evaluates the same sums to the same truncation order (a bigger group
    opens nodes for its whole bounding box, so it trades more interactions per target for fewer,
    longer traversals; the variants give the same fields, bit for bit or to FP32 rounding).  Decisions are cached per (tree
    size, target count, distribution statistics, fields, moment order, theta, precision, grouping,
    criterion, tolerance, expansion and its relative cap) and listed by tuning_log()."""
    pos = to_dev(pos, (-1, 3))
    nt = pos.shape[0]
    key = (_bucket(tree.n), _bucket(nt), distribution_key(tree, pos), int(mask), bool(quadrupole), float(theta),
           str(precision), grouping, criterion, None if eps is None else float(eps),
           None if model is None else model.key(), bool(bf16), None if expansion is None else float(expansion),
           float(expansion_rel))
    if key in _tuned and not force:
        return _tuned[key]
    fp32 = precision in ("float32", "f32", np.float32, torch.float32)
    cands = []
    for gs in sizes:
        for v in (range(len(WALK_VARIANTS)) if fp32 and int(gs) <= 64 else (0,)):
            cands.append((int(gs), v))
    times, inter = {}, {}
    for gs, v in cands:
        ctr = torch.zeros(4, dtype=torch.int64, device=device())
        kw = dict(quadrupole=quadrupole, group_size=gs, theta=theta, precision=precision, grouping=grouping,
                  criterion=criterion, eps=eps, model=model, bf16=bf16, expansion=expansion, eps_scale=eps_scale,
                  expansion_rel=expansion_rel, variant=v)
        walk(tree, pos, soft, perm, mask, counters=ctr, **kw)  # warm-up (and interaction count)
        torch.cuda.synchronize()
        inter[f"{gs}/{v}"] = int(ctr[0].item())
        best = float("inf")
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            walk(tree, pos, soft, perm, mask, **kw)
            e1.record()
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1))
        times[f"{gs}/{v}"] = best
    best_key = min(times, key=times.get)
    gs, v = (int(t) for t in best_key.split("/"))
    rec = {"choice": gs, "variant": v, "variant_name": WALK_VARIANTS[v], "ms": times, "interactions": inter,
           "distribution": {"max_level": key[2][0], "nodes_per_particle_x8": key[2][1],
                            "target_spread_log2": key[2][2]}}
    _tuned[key] = rec
    return rec
