"""Learned opening criterion: a 10-32-32-1 MLP that predicts each node's opening radius.

The reference opens a node unless r > max(size/theta + delta, h + 0.6 size + delta)
(treewalk.py:22-25).  Here the geometric test stays as the guard (at a loose theta), and a node is
accepted only when the group's distance r_min also exceeds a per-node radius
size * 10^(MLP(features) + margin), with the MLP trained on labels that direct summation generates:
for sampled nodes, the smallest distance beyond which the node's monopole stays within the
tolerance eps in every one of 32 directions.  Label generation, training (forward, backward, Adam)
and inference are libptg_hip kernels; torch only holds the buffers.

Usage:
    model = OpeningMLP.load()                      # the trained weights shipped with the package
    out = engine.walk(tree, ..., criterion="learned", model=model, eps=alpha * field_scale)
Training (on the GPU):  python -m pytreegrav_amd.learned --out pytreegrav_amd/opening_mlp.json
"""

import argparse
import itertools
import json
import os
import time

import numpy as np
import torch

from . import _lib
from ._lib import PTG_MLP_NPARAM, check, ptr
from .engine import DeviceTree, device, field_scale, stream, workspace

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "opening_mlp.json")
NIN, NH = 10, 32
RHO_LO, RHO_HI, NRADII = 0.5, 64.0, 48

FEATURES = ["log10(r_Q/size)/2, r_Q=(3.5|Q|/eps)^(1/4)", "delta/size", "|Q|/(M size^2)", "h/size",
            "children/8", "max child mass/M", "sum m_c d_c^2/(M size^2)", "sum m_c d_c^3/(M size^3)",
            "max child distance/size", "log10(M/(eps size^2))/4"]


def init_params(seed=0):
    """Glorot-uniform weights, zero biases (layout of ptg.h: W1, b1, W2, b2, W3, b3)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(PTG_MLP_NPARAM, dtype=np.float32)
    o = 0
    for fan_in, fan_out in ((NIN, NH), (NH, NH), (NH, 1)):
        lim = np.sqrt(6.0 / (fan_in + fan_out))
        p[o:o + fan_in * fan_out] = rng.uniform(-lim, lim, fan_in * fan_out)
        o += fan_in * fan_out + fan_out
    return p


_model_ids = itertools.count(1)


class OpeningMLP:
    """Parameters (float32, PTG_MLP_NPARAM) on the device plus the inference margin (log10 units).
    The parameters are fixed once the model exists (training builds a new model), so caches of
    opening radii and autotuning decisions key a model by the serial number it gets here rather
    than by a device address, which the allocator may hand to another model later."""

    def __init__(self, params, margin=0.0, meta=None):
        p = torch.as_tensor(np.asarray(params, dtype=np.float32))
        if p.numel() != PTG_MLP_NPARAM:
            raise ValueError(f"expected {PTG_MLP_NPARAM} parameters, got {p.numel()}")
        self.params = p.to(device()).contiguous()
        self.margin = float(margin)
        self.meta = dict(meta or {})
        self.uid = next(_model_ids)

    @classmethod
    def load(cls, path=DEFAULT_MODEL):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is missing: train the opening MLP first (python -m pytreegrav_amd.learned)")
        with open(path) as f:
            d = json.load(f)
        return cls(d["params"], d.get("margin", 0.0), d.get("meta"))

    def save(self, path):
        d = {"params": [float(v) for v in self.params.cpu().numpy()], "margin": self.margin, "meta": self.meta}
        with open(path, "w") as f:
            json.dump(d, f)

    def key(self):
        return (self.uid, self.margin)

    def forward(self, x, bf16=False):
        """MLP outputs for an (m, 10) float32 feature tensor on the device."""
        x = torch.as_tensor(x, dtype=torch.float32, device=device()).contiguous()
        y = torch.empty(x.shape[0], dtype=torch.float32, device=device())
        check(_lib.load().ptg_mlp_forward(ptr(self.params), ptr(x), x.shape[0], int(bool(bf16)), ptr(y), stream()),
              "ptg_mlp_forward")
        return y


def node_features(tree, nodes, eps):
    """(m, 10) float32 features of node elements `nodes` (device int32) at tolerance eps."""
    nodes = torch.as_tensor(nodes, dtype=torch.int32, device=device()).contiguous()
    x = torch.empty((nodes.numel(), NIN), dtype=torch.float32, device=device())
    check(_lib.load().ptg_node_features(_lib.ctypes.byref(tree.arrays), tree.n, tree.nnodes, ptr(nodes),
                                        nodes.numel(), float(eps), ptr(x), stream()), "ptg_node_features")
    return x


def error_curves(tree, nodes, cap=4096, nradii=NRADII, rho_lo=RHO_LO, rho_hi=RHO_HI, stat="max"):
    """(m, nradii) float64 monopole error over 32 directions at rho_j * size (units M/size^2),
    their maximum (stat "max") or root mean square ("rms"); rows of skipped nodes are -1."""
    nodes = torch.as_tensor(nodes, dtype=torch.int32, device=device()).contiguous()
    err = torch.empty((nodes.numel(), nradii), dtype=torch.float64, device=device())
    check(_lib.load().ptg_node_errcurve(_lib.ctypes.byref(tree.arrays), tree.n, tree.nnodes, tree.forest_ref,
                                        ptr(nodes), nodes.numel(), int(cap), int(nradii), float(rho_lo),
                                        float(rho_hi), {"max": 0, "rms": 1}[stat], ptr(err), stream()),
          "ptg_node_errcurve")
    return err


def radii_ladder(nradii=NRADII, rho_lo=RHO_LO, rho_hi=RHO_HI):
    return rho_lo * np.exp(np.log(rho_hi / rho_lo) * np.arange(nradii) / (nradii - 1))


def samples(tree, nodes, err, rows, eps, nradii=NRADII, rho_lo=RHO_LO, rho_hi=RHO_HI):
    """Training samples (x (m,10), y (m,), valid (m,)) pairing curve rows with tolerances."""
    rows = torch.as_tensor(rows, dtype=torch.int32, device=device()).contiguous()
    eps = torch.as_tensor(eps, dtype=torch.float64, device=device()).contiguous()
    m = rows.numel()
    x = torch.empty((m, NIN), dtype=torch.float32, device=device())
    y = torch.empty(m, dtype=torch.float32, device=device())
    valid = torch.empty(m, dtype=torch.int32, device=device())
    check(_lib.load().ptg_mlp_samples(_lib.ctypes.byref(tree.arrays), tree.n, tree.nnodes, ptr(nodes), ptr(rows),
                                      ptr(eps), m, ptr(err), int(nradii), float(rho_lo), float(rho_hi), ptr(x),
                                      ptr(y), ptr(valid), stream()), "ptg_mlp_samples")
    return x, y, valid


def grad(params, x, y, index=None, seed=0, batch=None, under=4.0):
    """Loss gradient (PTG_MLP_NPARAM,) and loss, through ptg_mlp_grad."""
    n = x.shape[0]
    batch = int(index.numel() if index is not None else (batch or n))
    g = torch.empty(PTG_MLP_NPARAM + 1, dtype=torch.float32, device=device())
    ws = workspace(_lib.load().ptg_mlp_grad_workspace_bytes(batch))
    check(_lib.load().ptg_mlp_grad(ptr(params), ptr(x), ptr(y), n, ptr(index), int(seed), batch, float(under), ptr(g),
                                   ptr(ws), ws.numel(), stream()), "ptg_mlp_grad")
    return g[:PTG_MLP_NPARAM], g[PTG_MLP_NPARAM]


def train(x, y, steps=4000, batch=4096, lr=3e-3, lr_end=2e-4, under=4.0, seed=0, init=None, log_every=0):
    """Adam on the asymmetric squared loss with the project's kernels; returns (params, losses)."""
    L = _lib.load()
    dev = device()
    x = x.to(dev, torch.float32).contiguous()
    y = y.to(dev, torch.float32).contiguous()
    p = torch.as_tensor(init if init is not None else init_params(seed), dtype=torch.float32).to(dev).contiguous()
    m = torch.zeros_like(p)
    v = torch.zeros_like(p)
    g = torch.empty(PTG_MLP_NPARAM + 1, dtype=torch.float32, device=dev)
    ws = workspace(L.ptg_mlp_grad_workspace_bytes(batch))
    losses = torch.empty(steps, dtype=torch.float32, device=dev)
    s = stream()
    for t in range(1, steps + 1):
        lr_t = lr_end + 0.5 * (lr - lr_end) * (1 + np.cos(np.pi * (t - 1) / max(steps - 1, 1)))
        check(L.ptg_mlp_grad(ptr(p), ptr(x), ptr(y), x.shape[0], None, int(seed) * 1000003 + t, batch, float(under),
                             ptr(g), ptr(ws), ws.numel(), s), "ptg_mlp_grad")
        check(L.ptg_mlp_adam(ptr(p), ptr(m), ptr(v), ptr(g), t, float(lr_t), 0.9, 0.999, 1e-8, s), "ptg_mlp_adam")
        losses[t - 1:t].copy_(g[PTG_MLP_NPARAM:PTG_MLP_NPARAM + 1])
        if log_every and t % log_every == 0:
            print(f"[train] step {t} loss {float(losses[t - 1]):.5f}", flush=True)
    return p, losses.cpu().numpy()


def open_radii(tree, pack, model, eps, bf16=False, scale=None):
    """Write the learned opening radius of every node into the FP32 pack (ptg_open_radii).  The
    tolerance is eps, or eps * scale for a device float64 `scale` (no host round trip): one value,
    or one per system of a forest tree."""
    if scale is not None and scale.numel() < (tree.nsys if tree.forest_ref is not None else 1):
        raise ValueError("scale needs one value per system of the forest")
    check(_lib.load().ptg_open_radii(_lib.ctypes.byref(tree.arrays), tree.n, tree.nnodes, tree.forest_ref,
                                     _lib.ctypes.byref(pack["wt"]), ptr(model.params), float(model.margin),
                                     float(eps), ptr(scale), int(bool(bf16)), stream()), "ptg_open_radii")


def training_set(tree, alphas, nodes_per_tree=20000, cap=4096, seed=0, fs=None, stat="max"):
    """Samples from one tree: `nodes_per_tree` random nodes, each paired with every tolerance
    alpha * (RMS field of the tree's particles).  Returns (x, y, valid) device tensors."""
    rng = np.random.default_rng(seed)
    k = min(nodes_per_tree, tree.nnodes)
    nodes = np.sort(rng.choice(tree.nnodes, size=k, replace=False)).astype(np.int32) + tree.n
    nodes_d = torch.from_numpy(nodes).to(device())
    err = error_curves(tree, nodes_d, cap=cap, stat=stat)
    fs = field_scale(tree) if fs is None else fs
    rows = np.repeat(np.arange(k, dtype=np.int32), len(alphas))
    eps = np.tile(np.asarray(alphas, dtype=np.float64) * fs, k)
    return samples(tree, nodes_d, err, rows, eps)


def _datasets(n, seed):
    from .sampling import gaussian_mixture, plummer_truncated

    x1, m1 = plummer_truncated(n, seed=seed)
    x2, m2 = gaussian_mixture(n, seed=seed + 1, ncomp=64)
    return [(x1, m1, np.full(n, 0.01)), (x2, m2, np.full(n, 0.01))]


def main(argv=None):
    ap = argparse.ArgumentParser(description="train the opening-radius MLP on the GPU")
    ap.add_argument("--out", default=DEFAULT_MODEL)
    ap.add_argument("-N", type=int, default=1000000)
    ap.add_argument("--nodes", type=int, default=40000)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--under", type=float, default=4.0)
    ap.add_argument("--cap", type=int, default=8192)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--stat", choices=("max", "rms"), default="max",
                    help="label statistic over the 32 directions: worst case or root mean square")
    args = ap.parse_args(argv)
    alphas = np.geomspace(1e-4, 1e-2, 7)
    t0 = time.time()
    X, Y, V = [], [], []
    for i, (x, m, h) in enumerate(_datasets(args.N, args.seed)):
        tree = DeviceTree(x, m, h, quadrupole=True)
        xs, ys, vs = training_set(tree, alphas, nodes_per_tree=args.nodes, cap=args.cap, seed=args.seed + i,
                                  stat=args.stat)
        X.append(xs)
        Y.append(ys)
        V.append(vs)
    X, Y, V = torch.cat(X), torch.cat(Y), torch.cat(V)
    keep = V > 0
    print(f"[learned] samples {int(keep.sum())} of {V.numel()} (clipped {int((V == 2).sum())}) "
          f"in {time.time() - t0:.1f}s", flush=True)
    X, Y = X[keep].contiguous(), Y[keep].contiguous()
    perm = torch.randperm(X.shape[0], generator=torch.Generator().manual_seed(args.seed)).to(device())
    nval = X.shape[0] // 10
    Xv, Yv = X[perm[:nval]], Y[perm[:nval]]
    Xt, Yt = X[perm[nval:]].contiguous(), Y[perm[nval:]].contiguous()
    t1 = time.time()
    p, losses = train(Xt, Yt, steps=args.steps, batch=args.batch, under=args.under, seed=args.seed,
                      log_every=max(1, args.steps // 10))
    torch.cuda.synchronize()
    model = OpeningMLP(p.cpu().numpy())
    pv = model.forward(Xv)
    res = (pv - Yv).cpu().numpy()
    # margin: raise predictions so that 98% of validation labels are not under-predicted
    margin = float(max(0.0, np.quantile(-res, 0.98)))
    model.margin = margin
    model.meta = {"train_seconds": round(time.time() - t1, 2), "samples": int(X.shape[0]), "N": args.N,
                  "alphas": [float(a) for a in alphas], "steps": args.steps, "batch": args.batch,
                  "under_weight": args.under, "label_statistic": args.stat, "final_loss": float(losses[-1]),
                  "val_rmse": float(np.sqrt(np.mean(res ** 2))), "val_under_frac_at_margin":
                  float(np.mean(res + margin < 0)), "features": FEATURES,
                  "data": "Plummer + 64-component Gaussian mixture, uniform node sample, h=0.01"}
    model.save(args.out)
    print(json.dumps({"out": args.out, **model.meta, "margin": margin}), flush=True)


if __name__ == "__main__":
    main()
