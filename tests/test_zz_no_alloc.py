"""ptg_walk_noalloc (the walk with its workspace in the library's static scratch, for callers that
cannot allocate) against ptg_walk with a caller workspace: the same kernels on the same inputs, so
the fields must agree bit for bit, on two calls in a row (the second with a smaller input, which
reuses the scratch the first call left behind), in FP32 (cell groups, the group expansion) and FP64
(the reference's fixed groups), with the overflow flag cleared.  Calls through the C ABI."""

import ctypes

import numpy as np
import pytest
import torch

from pytreegrav_amd import _lib, engine
from pytreegrav_amd._lib import PTG_ACC, PTG_F32, PTG_F64, PTG_POT, Criterion, check, ptr
from pytreegrav_amd.sampling import plummer_truncated

pytestmark = pytest.mark.gpu

_c = ctypes


def _noalloc():
    return _lib.load().ptg_walk_noalloc


def _both(tree, xt, ht, perm, mask, gs, theta, prec, groups, crit):
    """(ptg_walk result, ptg_walk_noalloc result, overflow flag) for the same arguments."""
    L = _lib.load()
    dev = engine.device()
    nt = xt.shape[0]
    W = engine.field_width(mask)
    p = tree.pack(prec, False)
    goff, gcnt = groups if groups is not None else (None, None)
    exact = int(prec == PTG_F64)
    ref = torch.empty((nt, W), dtype=torch.float64, device=dev)
    ws = engine.workspace(L.ptg_walk_workspace_bytes(nt, gs))
    check(L.ptg_walk(_c.byref(p["wt"]), None, ptr(xt), ptr(ht), ptr(perm), 0, nt, mask, 0, gs, ptr(goff), ptr(gcnt),
                     theta, _c.byref(crit), 1.0, exact, ptr(ref), None, ptr(ws), ws.numel(), engine.stream()),
          "ptg_walk")
    check(L.ptg_walk_check(ptr(ws), engine.stream()), "ptg_walk")
    out = torch.full((nt, W), float("nan"), dtype=torch.float64, device=dev)
    err = torch.full((1,), -1, dtype=torch.int32, device=dev)
    check(_noalloc()(_c.byref(p["wt"]), None, ptr(xt), ptr(ht), ptr(perm), 0, nt, mask, 0, gs, ptr(goff), ptr(gcnt),
                     theta, _c.byref(crit), 1.0, exact, ptr(out), None, ptr(err), engine.stream()),
          "ptg_walk_noalloc")
    torch.cuda.synchronize()
    return ref, out, int(err.item())


def _targets(n, seed):
    rng = np.random.default_rng(seed)
    xt = rng.normal(size=(n, 3)) * 3.0
    return xt, np.full(n, 0.01)


@pytest.mark.parametrize("prec", [PTG_F32, PTG_F64])
def test_noalloc_walk_equals_workspace_walk(prec):
    from pytreegrav_amd import ConstructTree

    x, m = plummer_truncated(200_000, seed=91)
    h = np.full(len(m), 0.01)
    tree = ConstructTree(x, m, h).dev
    dev = engine.device()
    mask = PTG_POT | PTG_ACC
    gs = 64 if prec == PTG_F32 else 8
    for call, n in enumerate((60_000, 7_000)):  # the second call is smaller and reuses the scratch
        print(f"[noalloc] prec={prec} call={call} targets={n}", flush=True)
        xt, ht = _targets(n, 100 + call)
        xd = torch.from_numpy(xt).to(dev)
        hd = torch.from_numpy(ht).to(dev)
        perm = engine.morton_order(xd)
        groups = engine.target_groups(xd, perm, gs, mode="cells") if prec == PTG_F32 else None
        crit = Criterion(_lib.PTG_CRIT_THETA, 0, 0.0)
        ref, out, err = _both(tree, xd, hd, perm, mask, gs, 0.7, prec, groups, crit)
        assert err == 0
        assert torch.equal(out, ref), call
        if prec == PTG_F32:  # the group expansion path of the same kernel
            crit = Criterion(_lib.PTG_CRIT_THETA, 1, 0.0, 1e-4, None)
            ref, out, err = _both(tree, xd, hd, perm, mask, gs, 0.7, prec, groups, crit)
            assert err == 0
            assert torch.equal(out, ref), call


def test_noalloc_walk_rejects_bad_criteria_and_needs_err():
    from pytreegrav_amd import ConstructTree

    x, m = plummer_truncated(20_000, seed=92)
    h = np.full(len(m), 0.01)
    tree = ConstructTree(x, m, h).dev
    dev = engine.device()
    xd = torch.from_numpy(x).to(dev)
    hd = torch.from_numpy(h).to(dev)
    p = tree.pack(PTG_F32, False)
    groups = engine.target_groups(xd, tree.order, 64, mode="cells")
    out = torch.empty((len(x), 4), dtype=torch.float64, device=dev)
    err = torch.zeros(1, dtype=torch.int32, device=dev)
    f = _noalloc()
    crit = Criterion(_lib.PTG_CRIT_THETA, 2, 0.0, 1e-4, None)
    st = f(_c.byref(p["wt"]), None, ptr(xd), ptr(hd), ptr(tree.order), 0, len(x), 3, 0, 64, ptr(groups[0]),
           ptr(groups[1]), 0.7, _c.byref(crit), 1.0, 0, ptr(out), None, ptr(err), engine.stream())
    assert st != 0 and b"group_expansion" in _lib.load().ptg_last_error()
    crit = Criterion(_lib.PTG_CRIT_THETA, 0, 0.0)
    st = f(_c.byref(p["wt"]), None, ptr(xd), ptr(hd), ptr(tree.order), 0, len(x), 3, 0, 64, ptr(groups[0]),
           ptr(groups[1]), 0.7, _c.byref(crit), 1.0, 0, ptr(out), None, None, engine.stream())
    assert st != 0 and b"err" in _lib.load().ptg_last_error()
