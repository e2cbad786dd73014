"""Batched gravity of many independent systems ("forest"): one build, one pack, one grouping and one
walk launch cover every system.

The reference solves one particle set per call (frontend.Accel / Potential, src/pytreegrav/
frontend.py:391-520, building one Octree each).  Here systems of uneven size are concatenated;
every system still gets its own bounding cube, Morton keys, octree, target groups and FP32 origin,
so each system's result equals what the single-system path gives for it alone, bit for bit.  The
walk stage (groups and walk; the pack is made once before) can be captured into a HIP graph and
replayed.
"""

import numpy as np
import torch

from . import _lib
from ._lib import PTG_F32, PTG_F64, check, ptr
from .engine import DeviceTree, device, fields_mask, field_width, stream, to_dev, unpack_fields, workspace


def _prec(precision):
    return PTG_F32 if precision in ("float32", "f32", np.float32, torch.float32) else PTG_F64


class Forest:
    """nsys independent systems: particles [offsets[s], offsets[s+1]) of (pos, m, h) form system s.

    precision "float32" (default): FP32 walk with cell-aligned groups of up to group_size targets;
    "float64": the reference's FP64 branch set with runs of group_size targets per system."""

    def __init__(self, pos, m, h, offsets, quadrupole=False, precision="float32", group_size=None):
        n = len(pos)
        off = np.asarray(offsets.cpu() if isinstance(offsets, torch.Tensor) else offsets, dtype=np.int64)
        if off.ndim != 1 or off.size < 2 or off[0] != 0 or off[-1] != n or np.any(np.diff(off) < 2):
            raise ValueError("offsets must rise from 0 to n with at least two particles per system")
        self.pos = to_dev(pos, (-1, 3))
        self.m = to_dev(m, (n,))
        self.h = to_dev(h, (n,))
        self.offsets = off
        self.n, self.nsys = n, off.size - 1
        self.quadrupole = bool(quadrupole)
        self.prec = _prec(precision)
        self.group_size = int(group_size) if group_size else (64 if self.prec == PTG_F32 else 8)
        self.tree = DeviceTree(self.pos, self.m, self.h, quadrupole=quadrupole, offsets=off)
        self._graphs = {}

    def _stage(self, mask, theta, G, out, ws, goff, gcnt, counters=None):
        """Enqueue pack, groups and walk (no allocation, no host synchronisation)."""
        L = _lib.load()
        t = self.tree
        p = t.pack(self.prec, self.quadrupole)
        mode = _lib.PTG_GROUPS_CELLS if self.prec == PTG_F32 else _lib.PTG_GROUPS_RUNS
        gws = self._gws
        check(L.ptg_target_groups(ptr(self.pos), ptr(t.order), ptr(t.keys), self.n, t.forest_ref, self.group_size,
                                  mode, ptr(goff), ptr(gcnt), ptr(gws), gws.numel(), stream()), "ptg_target_groups")
        crit = _lib.Criterion(_lib.PTG_CRIT_THETA, 0, 0.0)
        # targets read from the tree's Morton-ordered coordinates and softenings (sorted order)
        check(L.ptg_walk(_lib.ctypes.byref(p["wt"]), t.forest_ref, ptr(t.coordinates), ptr(t.softenings),
                         ptr(t.order), 1, self.n,
                         int(mask), int(self.quadrupole), self.group_size, ptr(goff), ptr(gcnt), float(theta),
                         _lib.ctypes.byref(crit), float(G), int(self.prec == PTG_F64), ptr(out), ptr(counters),
                         ptr(ws), ws.numel(), stream()), "ptg_walk")

    def _buffers(self, mask):
        L = _lib.load()
        dev = device()
        self._gws = workspace(L.ptg_groups_workspace_bytes(self.n, self.nsys))
        ws = workspace(L.ptg_walk_workspace_bytes(self.n, self.group_size))
        out = torch.empty((self.n, field_width(mask)), dtype=torch.float64, device=dev)
        goff = torch.empty(self.n + 1, dtype=torch.int32, device=dev)
        gcnt = torch.empty(1, dtype=torch.int32, device=dev)
        return out, ws, goff, gcnt

    def evaluate(self, potential=True, accel=True, tidal=False, theta=0.7, G=1.0, counters=None):
        """Fields of every particle from its own system, in input order: dict of device views
        ("potential" (n,), "accel" (n,3), "tidal" (n,3,3))."""
        mask = fields_mask(potential, accel, tidal)
        out, ws, goff, gcnt = self._buffers(mask)
        self.tree.pack(self.prec, self.quadrupole)
        self._stage(mask, theta, G, out, ws, goff, gcnt, counters)
        check(_lib.load().ptg_walk_check(ptr(ws), stream()), "ptg_walk")
        return unpack_fields(out, mask)

    def capture(self, potential=True, accel=True, tidal=False, theta=0.7, G=1.0):
        """Capture the walk stage (groups + walk over all systems) into a HIP graph.  Returns
        (replay, out): replay() re-runs the stage and checks the traversal overflow flag; out is
        the (n, W) result buffer every replay rewrites."""
        mask = fields_mask(potential, accel, tidal)
        out, ws, goff, gcnt = self._buffers(mask)
        self.tree.pack(self.prec, self.quadrupole)  # the pack is built once, outside the graph
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):  # warm-up on the capture stream
            self._stage(mask, theta, G, out, ws, goff, gcnt)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            self._stage(mask, theta, G, out, ws, goff, gcnt)
        keep = (g, out, ws, goff, gcnt, self._gws)
        self._graphs[(mask, theta, G)] = keep

        def replay():
            g.replay()
            check(_lib.load().ptg_walk_check(ptr(ws), stream()), "ptg_walk")
            return out

        return replay, out

    def split(self, arr):
        """Per-system views of an (n, ...) array in input order."""
        return [arr[self.offsets[s]:self.offsets[s + 1]] for s in range(self.nsys)]


def BatchedAccelPotential(pos, m, h, offsets, theta=0.7, G=1.0, quadrupole=False, precision="float32",
                          group_size=None, return_device=False):
    """Accel and potential of many independent systems in one batched build and walk: returns
    (accel (n,3), potential (n,)) in input order, each particle feeling only its own system."""
    f = Forest(pos, m, h, offsets, quadrupole=quadrupole, precision=precision, group_size=group_size)
    r = f.evaluate(potential=True, accel=True, theta=theta, G=G)
    if return_device:
        return r["accel"], r["potential"]
    return r["accel"].cpu().numpy(), r["potential"].cpu().numpy()
