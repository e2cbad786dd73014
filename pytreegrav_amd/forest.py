"""Batched gravity of many independent systems ("forest"): one build, one pack, one grouping and one
walk launch cover every system.

The reference solves one particle set per call (frontend.Accel / Potential, src/pytreegrav/
frontend.py:391-520, building one Octree each).  Here systems of uneven size are concatenated;
every system still gets its own bounding cube, Morton keys, octree, target groups and FP32 origin,
so each system's result equals what the single-system path gives for it alone, bit for bit.  The
walk stage (groups and walk; the pack is made once before) can be captured into a HIP graph and
replayed.  The learned opening criterion is available per system: each system's tolerance is
alpha x its own RMS field, estimated by one sample walk over all systems.
"""

import time

import numpy as np
import torch

from . import _lib
from ._lib import PTG_ACC, PTG_F32, PTG_F64, PTG_TIDAL, check, ptr
from .engine import (TUNE_SIZES, DeviceTree, device, fields_mask, field_width, stream, to_dev, unpack_fields,
                     workspace)


def _prec(precision):
    return PTG_F32 if precision in ("float32", "f32", np.float32, torch.float32) else PTG_F64


class Forest:
    """nsys independent systems: particles [offsets[s], offsets[s+1]) of (pos, m, h) form system s.

    precision "float32" (default): FP32 walk with cell-aligned groups of up to group_size targets;
    "float64": the reference's FP64 branch set with runs of group_size targets per system.
    group_size "auto": chosen by timing the walk stage at 32/64/128/256 (autotune()).
    criterion "learned" (FP32): the tree is built with quadrupoles for the opening-radius features.
    field_samples: cell groups per system walked to estimate its RMS field (the scale of the
    learned criterion's and the expansion's tolerances)."""

    def __init__(self, pos, m, h, offsets, quadrupole=False, precision="float32", group_size=None,
                 criterion="theta", field_samples=8):
        n = len(pos)
        off = np.asarray(offsets.cpu() if isinstance(offsets, torch.Tensor) else offsets, dtype=np.int64)
        if off.ndim != 1 or off.size < 2 or off[0] != 0 or off[-1] != n or np.any(np.diff(off) < 2):
            raise ValueError("offsets must rise from 0 to n with at least two particles per system")
        if criterion not in ("theta", "learned"):
            raise ValueError(f"criterion must be 'theta' or 'learned', got {criterion!r}")
        self.pos = to_dev(pos, (-1, 3))
        self.m = to_dev(m, (n,))
        self.h = to_dev(h, (n,))
        self.offsets = off
        self.n, self.nsys = n, off.size - 1
        self.quadrupole = bool(quadrupole)
        self.prec = _prec(precision)
        if criterion == "learned" and self.prec != PTG_F32:
            raise ValueError("criterion='learned' runs on the float32 walk")
        self.criterion = criterion
        self.auto = group_size == "auto"
        self.group_size = 64 if self.auto or not group_size else int(group_size)
        if self.prec == PTG_F64 and not group_size:
            self.group_size = 8
        self.tree = DeviceTree(self.pos, self.m, self.h, quadrupole=quadrupole or criterion == "learned", offsets=off)
        self._graphs = {}
        self.tuning = None
        self.field_samples = int(field_samples)  # sample groups per system of the RMS-field estimate
        self._sample_consts = {}

    # ---------------------------------------------------------------------------------------------
    def _groups(self, gs, goff, gcnt, gws, t=None):
        t = t or self.tree
        mode = _lib.PTG_GROUPS_CELLS if self.prec == PTG_F32 else _lib.PTG_GROUPS_RUNS
        check(_lib.load().ptg_target_groups(ptr(self.pos), ptr(t.order), ptr(t.keys), self.n, t.forest_ref, gs, mode,
                                            ptr(goff), ptr(gcnt), ptr(gws), gws.numel(), stream()),
              "ptg_target_groups")

    def _walk(self, p, mask, theta, G, out, ws, gs, goff, gcnt, crit, counters=None, t=None):
        t = t or self.tree
        # targets read from the tree's Morton-ordered coordinates and softenings (sorted order)
        check(_lib.load().ptg_walk(_lib.ctypes.byref(p["wt"]), t.forest_ref, ptr(t.coordinates), ptr(t.softenings),
                                   ptr(t.order), 1, self.n, int(mask), int(self.quadrupole), gs, ptr(goff),
                                   ptr(gcnt), float(theta), _lib.ctypes.byref(crit), float(G),
                                   int(self.prec == PTG_F64), ptr(out), ptr(counters), ptr(ws), ws.numel(),
                                   stream()), "ptg_walk")

    def _buffers(self, mask, gs):
        L = _lib.load()
        dev = device()
        gws = workspace(L.ptg_groups_workspace_bytes(self.n, self.nsys))
        ws = workspace(L.ptg_walk_workspace_bytes(self.n, gs))
        out = torch.empty((self.n, field_width(mask)), dtype=torch.float64, device=dev)
        goff = torch.empty(self.n + 1, dtype=torch.int32, device=dev)
        gcnt = torch.empty(1, dtype=torch.int32, device=dev)
        return out, ws, goff, gcnt, gws

    def system_field_scales(self, goff, gcnt, samples=None, gs=64, t=None, check_walk=True):
        """RMS |acceleration| (G=1) of each system, a device (nsys,) float64 array: a theta=1 FP32
        walk of `samples` of each system's cell-aligned groups (each padded to gs targets by
        repeating its last target), then one RMS per system (ptg_rms_rows).  t: the tree (default the
        forest's own); check_walk=False skips the overflow check, which synchronises (a captured step
        checks after the replay)."""
        t = t or self.tree
        samples = self.field_samples if samples is None else int(samples)
        dev = device()
        # everything stays on the device (the count of groups too), so a captured step can run this;
        # the sample forest's offsets, groups and order are constants, made once per (samples, gs)
        key = (samples, gs)
        if key not in self._sample_consts:
            # the samples form a forest of their own: system s owns sample targets [s, s+1) * samples * gs
            nt = self.nsys * samples * gs
            self._sample_consts[key] = (
                torch.arange(self.nsys + 1, device=dev, dtype=torch.int64) * (samples * gs),
                (torch.arange(self.nsys * samples + 1, device=dev, dtype=torch.int64) * gs).to(torch.int32),
                torch.full((1,), self.nsys * samples, dtype=torch.int32, device=dev),
                torch.arange(nt, device=dev, dtype=torch.int64))
        soff_s, goff_s, gcnt_s, perm = self._sample_consts[key]
        nt = perm.numel()
        pos = torch.empty((nt, 3), dtype=torch.float64, device=dev)
        h = torch.empty(nt, dtype=torch.float64, device=dev)
        check(_lib.load().ptg_field_sample(ptr(goff), ptr(gcnt), t.forest_ref, self.n, samples, gs,
                                           ptr(t.coordinates), ptr(t.softenings), ptr(pos), ptr(h), stream()),
              "ptg_field_sample")
        fs = _lib.Forest(self.nsys, ptr(soff_s).value, ptr(t.roots).value, ptr(t.cubes).value)
        out = torch.empty((nt, 3), dtype=torch.float64, device=dev)
        L = _lib.load()
        ws = workspace(L.ptg_walk_workspace_bytes(nt, gs))
        p = t.pack(PTG_F32, False, norms=False if self.criterion == "learned" else None)
        crit = _lib.Criterion(_lib.PTG_CRIT_THETA, 0, 0.0)
        check(L.ptg_walk(_lib.ctypes.byref(p["wt"]), _lib.ctypes.byref(fs), ptr(pos), ptr(h), ptr(perm), 1, nt,
                         PTG_ACC, 0, gs, ptr(goff_s), ptr(gcnt_s), 1.0, _lib.ctypes.byref(crit), 1.0, 0, ptr(out),
                         None, ptr(ws), ws.numel(), stream()), "ptg_walk")
        if check_walk:
            check(L.ptg_walk_check(ptr(ws), stream()), "ptg_walk")
        else:
            self._sample_ws = ws  # checked after the captured step's replay
        rms = torch.empty(self.nsys, dtype=torch.float64, device=dev)
        check(L.ptg_rms_rows(ptr(out), nt, 3, 0, 3, self.nsys, ptr(rms), stream()), "ptg_rms_rows")
        return rms

    def _prepare(self, gs, theta, alpha, eps, model, bf16, goff, gcnt, gws, mask=PTG_ACC, expansion=None,
                 expansion_rel=0.0, t=None, check_walk=True):
        """Groups, and for the learned criterion the opening radii; returns (pack, criterion, theta).
        expansion: group-expansion tolerance in units of each system's RMS field (None: off);
        expansion_rel: its cap relative to each folded element's own field (0: none)."""
        t = t or self.tree
        self._groups(gs, goff, gcnt, gws, t=t)
        if expansion is not None:
            if self.prec != PTG_F32 or self.quadrupole or (mask & PTG_TIDAL):
                raise ValueError("expansion runs on the float32 walk with monopole interactions, for the potential "
                                 "and/or the acceleration")
            if not expansion > 0:
                raise ValueError("expansion must be a positive tolerance (relative to each system's RMS field)")
            if not expansion_rel >= 0:
                raise ValueError("expansion_rel must be >= 0 (0: no cap)")
        if self.criterion != "learned":
            crit = _lib.Criterion(_lib.PTG_CRIT_THETA, 0, 0.0)
            self._expand(crit, expansion, None, goff, gcnt, expansion_rel, t=t, check_walk=check_walk)
            return t.pack(self.prec, self.quadrupole), crit, (0.7 if theta is None else theta)
        if model is None:
            from .learned import OpeningMLP

            model = OpeningMLP.load()
        if eps is not None:  # explicit per-system tolerances
            scale = torch.as_tensor(np.broadcast_to(np.asarray(eps, dtype=np.float64), (self.nsys,)).copy()).to(device())
            base = 1.0
        else:
            if alpha is None or not alpha > 0:
                raise ValueError("criterion='learned' needs alpha > 0 (or per-system eps)")
            scale, base = self.system_field_scales(goff, gcnt, t=t, check_walk=check_walk), float(alpha)
        p = t.pack_learned(model, base, self.quadrupole, bf16, scale=scale)
        self._scale = scale  # keeps the scale alive while the pack is cached under its address
        crit = _lib.Criterion(_lib.PTG_CRIT_RADIUS, 0, 0.0)
        self._expand(crit, expansion, scale if eps is None else None, goff, gcnt, expansion_rel, t=t,
                     check_walk=check_walk)
        return p, crit, (2.0 if theta is None else theta)

    def _expand(self, crit, expansion, rms, goff, gcnt, expansion_rel=0.0, t=None, check_walk=True):
        """Switch the group expansion on in crit at expansion x each system's RMS field (device)."""
        if expansion is None:
            return
        if rms is None:
            rms = self.system_field_scales(goff, gcnt, t=t, check_walk=check_walk)
        self._exp_scale = rms  # the walk (and any captured graph) reads it through crit
        crit.group_expansion = 1
        crit.expansion_eps = float(expansion)
        crit.expansion_rel = float(expansion_rel)
        crit.eps_scale = rms.data_ptr()

    def evaluate(self, potential=True, accel=True, tidal=False, theta=None, G=1.0, counters=None, alpha=None,
                 eps=None, model=None, bf16=False, expansion=None, expansion_rel=0.0):
        """Fields of every particle from its own system, in input order: dict of device views
        ("potential" (n,), "accel" (n,3), "tidal" (n,3,3)).  theta defaults to 0.7 (geometric
        criterion) or 2.0 (the guard of the learned criterion, whose per-system tolerance is
        alpha x the system's RMS field, or the given per-system eps).  expansion: group-expansion
        tolerance relative to each system's RMS field (float32 monopole walks; None: off);
        expansion_rel: its cap relative to each folded element's own field (0: none)."""
        if self.auto and self.tuning is None:
            self.autotune(potential, accel, tidal, theta, G, alpha=alpha, eps=eps, model=model, bf16=bf16,
                          expansion=expansion, expansion_rel=expansion_rel)
        gs = self.group_size
        mask = fields_mask(potential, accel, tidal)
        out, ws, goff, gcnt, gws = self._buffers(mask, gs)
        p, crit, th = self._prepare(gs, theta, alpha, eps, model, bf16, goff, gcnt, gws, mask, expansion,
                                    expansion_rel)
        self._walk(p, mask, th, G, out, ws, gs, goff, gcnt, crit, counters)
        check(_lib.load().ptg_walk_check(ptr(ws), stream()), "ptg_walk")
        return unpack_fields(out, mask)

    def autotune(self, potential=True, accel=True, tidal=False, theta=None, G=1.0, sizes=TUNE_SIZES, **kw):
        """Pick the group size by timing the walk stage (groups + walk, HIP events, best of two) at
        each candidate; records {"choice", "ms"} in self.tuning."""
        mask = fields_mask(potential, accel, tidal)
        times = {}
        for gs in sizes:
            out, ws, goff, gcnt, gws = self._buffers(mask, gs)
            p, crit, th = self._prepare(gs, theta, kw.get("alpha"), kw.get("eps"), kw.get("model"),
                                        kw.get("bf16", False), goff, gcnt, gws, mask, kw.get("expansion"),
                                        kw.get("expansion_rel", 0.0))
            best = float("inf")
            for _ in range(2):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                self._groups(gs, goff, gcnt, gws)
                self._walk(p, mask, th, G, out, ws, gs, goff, gcnt, crit)
                e1.record()
                torch.cuda.synchronize()
                best = min(best, e0.elapsed_time(e1))
            check(_lib.load().ptg_walk_check(ptr(ws), stream()), "ptg_walk")
            times[int(gs)] = best
        self.group_size = min(times, key=times.get)
        self.tuning = {"choice": self.group_size, "ms": times, "when": time.time()}
        return self.tuning

    def capture(self, potential=True, accel=True, tidal=False, theta=None, G=1.0, alpha=None, eps=None, model=None,
                bf16=False, expansion=None, expansion_rel=0.0):
        """Capture the walk stage (groups + walk over all systems) into a HIP graph.  The pack (and,
        for the learned criterion, the opening radii) are made once before the capture.  Returns
        (replay, out): replay() re-runs the stage and checks the traversal overflow flag; out is
        the (n, W) result buffer every replay rewrites."""
        gs = self.group_size
        mask = fields_mask(potential, accel, tidal)
        out, ws, goff, gcnt, gws = self._buffers(mask, gs)
        p, crit, th = self._prepare(gs, theta, alpha, eps, model, bf16, goff, gcnt, gws, mask, expansion,
                                    expansion_rel)
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):  # warm-up on the capture stream
            self._groups(gs, goff, gcnt, gws)
            self._walk(p, mask, th, G, out, ws, gs, goff, gcnt, crit)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            self._groups(gs, goff, gcnt, gws)
            self._walk(p, mask, th, G, out, ws, gs, goff, gcnt, crit)
        self._graphs[(mask, th, G)] = (g, out, ws, goff, gcnt, gws, p, crit, getattr(self, "_exp_scale", None))

        def replay():
            g.replay()
            check(_lib.load().ptg_walk_check(ptr(ws), stream()), "ptg_walk")
            return out

        return replay, out

    def capture_step(self, potential=True, accel=True, theta=None, G=1.0, alpha=None, model=None, bf16=False,
                     expansion=None, capacity=None, expansion_rel=0.0):
        """Capture one whole step -- tree build (ptg_build_deferred), target groups, for the learned
        criterion the per-system field scales and opening radii, and the walk -- into one HIP graph.
        Each replay recomputes every system's fields from the current contents of the input buffers
        self.pos, self.m, self.h (write new positions into them in place; same particle count and
        systems), with no host synchronisation inside the step.  capacity = (node capacity, level
        capacity) of the build, by default 1.25x the node count and 4 levels more than the tree built
        by the constructor.  Returns (replay, out): replay() runs the graph, then checks that the
        tree fit its capacities and that no traversal stack overflowed (raising otherwise; capture
        again with a larger capacity), and returns out, the (n, W) buffer every replay rewrites."""
        if self.prec != PTG_F32:
            raise ValueError("capture_step runs the float32 walk")
        mask = fields_mask(potential, accel, False)
        if self.auto and self.tuning is None:
            self.autotune(potential, accel, False, theta, G, alpha=alpha, model=model, bf16=bf16,
                          expansion=expansion, expansion_rel=expansion_rel)
        gs = self.group_size
        if self.criterion == "learned":
            if alpha is None or not alpha > 0:
                raise ValueError("capture_step with criterion='learned' needs alpha > 0")
            if model is None:
                from .learned import OpeningMLP

                model = OpeningMLP.load()  # loaded before the capture (file and host-to-device copy)
        if capacity is None:
            capacity = (int(self.tree.nnodes * 1.25) + 64, min(255, int(self.tree.info.max_level) + 4))
        out, ws, goff, gcnt, gws = self._buffers(mask, gs)
        quad_build = self.quadrupole or self.criterion == "learned"
        held = {}

        def step():
            t = DeviceTree(self.pos, self.m, self.h, quadrupole=quad_build, offsets=self.tree.sys_offsets,
                           capacity=capacity)
            p, crit, th = self._prepare(gs, theta, alpha, None, model, bf16, goff, gcnt, gws, mask, expansion,
                                        expansion_rel, t=t, check_walk=False)
            self._walk(p, mask, th, G, out, ws, gs, goff, gcnt, crit, t=t)
            # everything the graph reads through a captured pointer lives as long as the graph: the
            # tree, packs and tolerances made inside the step, and the model's parameters (loaded
            # outside it), whose memory would otherwise go back to the allocator for eager work
            held.update(tree=t, pack=p, crit=crit, model=model, scale=getattr(self, "_scale", None),
                        exp_scale=getattr(self, "_exp_scale", None), sample_ws=getattr(self, "_sample_ws", None))

        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):  # warm-up on the capture stream: the same step, eagerly
            step()
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        held["tree"].check()
        check(_lib.load().ptg_walk_check(ptr(ws), stream()), "ptg_walk")
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            step()
        state = dict(held)  # the captured step's tree, packs and scales live as long as the graph
        self._graphs[("step", mask, theta, G)] = (g, out, ws, goff, gcnt, gws, state)
        L = _lib.load()

        def replay():
            g.replay()
            state["tree"].check()  # synchronises: the tree fit its capacities, finite inputs
            check(L.ptg_walk_check(ptr(ws), stream()), "ptg_walk")
            if state["sample_ws"] is not None:
                check(L.ptg_walk_check(ptr(state["sample_ws"]), stream()), "ptg_walk")
            return out

        return replay, out

    def split(self, arr):
        """Per-system views of an (n, ...) array in input order."""
        return [arr[self.offsets[s]:self.offsets[s + 1]] for s in range(self.nsys)]


def BatchedAccelPotential(pos, m, h, offsets, theta=None, G=1.0, quadrupole=False, precision="float32",
                          group_size=None, return_device=False, criterion="theta", alpha=None, expansion=None):
    """Accel and potential of many independent systems in one batched build and walk: returns
    (accel (n,3), potential (n,)) in input order, each particle feeling only its own system."""
    f = Forest(pos, m, h, offsets, quadrupole=quadrupole, precision=precision, group_size=group_size,
               criterion=criterion)
    r = f.evaluate(potential=True, accel=True, theta=theta, G=G, alpha=alpha, expansion=expansion)
    if return_device:
        return r["accel"], r["potential"]
    return r["accel"].cpu().numpy(), r["potential"].cpu().numpy()
