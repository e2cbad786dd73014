"""Drop-in gravity API: the names, signatures, defaults, return shapes and error behaviour of
pytreegrav's frontend (src/pytreegrav/frontend.py:20-1044), computed on an MI355X by libptg_hip.

Numerics.  ``precision="float64"`` reproduces the reference's CPU path (FP64, the reference's
branch set per pair, reference group_size default 8).  ``precision="float32"`` is the fast walk:
FP32 node-relative coordinates with FP32 accumulation, FP64 keys and moments.  The self-gravity
functions keep the reference's ``device`` argument: ``device="cpu"`` (default) selects float64,
``device="cuda"`` selects float32 -- the precisions of the reference's two backends -- and both run
on the GPU.  ``parallel`` is accepted and ignored: every path is parallel.

Inputs may be array-likes (results come back as float64 numpy arrays in input order, like the
reference) or torch tensors on the GPU (results stay on the GPU as float64 tensors).
"""

import warnings

import numpy as np
import torch

from . import engine
from .engine import PTG_ACC, PTG_POT, PTG_TIDAL

# group size of the FP32 walk when the caller does not choose one (one target per lane)
FP32_DEFAULT_GROUP = 64


# Accel / Potential with method="adaptive" in float32 walk the tree above this many particles and sum
# directly below it: on one MI355X a whole float32 Accel call from numpy inputs (tree build + walk vs
# the direct sum, Plummer spheres, h = 0.01) takes 1.01 vs 0.62 ms at N = 16000 and 1.10 vs 1.48 ms
# at N = 32000 (0.82 vs 0.42 and 0.84 vs 1.25 ms from device tensors)
FP32_TREE_MIN_N = 24000


class _DefaultTheta(float):
    """The reference's default opening angle, 0.7 (frontend.py:149, :425).  It is an ordinary float
    everywhere; only its identity tells a theta the caller left at the default from an explicit 0.7,
    so the learned and error-bounded criteria can replace the default by their own guard while an
    explicit theta=0.7 is kept."""

    def __repr__(self):
        return "0.7"


THETA_DEFAULT = _DefaultTheta(0.7)
# geometric guard of the learned and error-bounded criteria when the caller gives no theta: the
# reference's test at this opening angle is applied on top of the per-node error bound, so an
# opening radius the network under-predicts on an unfamiliar particle distribution still cannot
# accept a node nearer than size / GUARD_THETA
GUARD_THETA = 1.0


def _f64(a):
    """Cast to float64 without collapsing a length-1 array to a scalar (frontend.py:20-25)."""
    if isinstance(a, torch.Tensor):
        return a.to(torch.float64)
    return np.asarray(a, dtype=np.float64)


def _atleast_2d(a):
    if isinstance(a, torch.Tensor):
        return a.reshape(1, -1) if a.dim() == 1 else a
    return np.atleast_2d(a)


def _atleast_1d(a):
    if isinstance(a, torch.Tensor):
        return a.reshape(1) if a.dim() == 0 else a
    return np.atleast_1d(a)


def _is_dev(a):
    return isinstance(a, torch.Tensor) and a.is_cuda


def _finish(t, like_device):
    """Device tensor -> numpy (host callers) or the tensor itself (GPU callers)."""
    return t if like_device else t.cpu().numpy()


def _zeros_like_len(a):
    if isinstance(a, torch.Tensor):
        return torch.zeros(a.shape[0], dtype=torch.float64, device=a.device)
    return np.zeros(len(a))


def _finite(*arrays):
    for a in arrays:
        if isinstance(a, torch.Tensor):
            if not bool(torch.isfinite(a).all()):
                return False
        elif not np.all(np.isfinite(a)):
            return False
    return True


def _precision(precision, device=None):
    if precision is None:
        return "float32" if device == "cuda" else "float64"
    if precision in ("float32", "f32", np.float32, torch.float32):
        return "float32"
    if precision in ("float64", "f64", np.float64, torch.float64):
        return "float64"
    raise ValueError(f"precision must be 'float64' or 'float32', got {precision!r}")


def _group(group_size, precision):
    if isinstance(group_size, str):
        if group_size != "auto":
            raise ValueError(f"group_size must be an integer, None or 'auto', got {group_size!r}")
        return group_size  # timed per problem by engine.autotune_group_size
    if group_size is None:
        return FP32_DEFAULT_GROUP if precision == "float32" else 8
    group_size = int(group_size)
    if not 1 <= group_size <= 256:
        raise ValueError(f"group_size must be between 1 and 256, got {group_size}")
    return group_size


def checkTreeQuadrupoles(tree, quadrupole):
    """Raise ValueError if a quadrupole walk is requested on a tree without them (frontend.py:28-37)."""
    if quadrupole and not tree.HasQuads:
        raise ValueError(
            "quadrupole=True requires a tree built with quadrupole moments: pass a tree from "
            "ConstructTree(..., quadrupole=True), or evaluate with quadrupole=False."
        )


def valueTestMethod(method):
    """Raise TypeError/ValueError unless method is 'adaptive', 'bruteforce' or 'tree' (frontend.py:40-50)."""
    methods = ["adaptive", "bruteforce", "tree"]
    if type(method) != str:
        raise TypeError("Invalid method type %s, must be str" % type(method))
    if method not in methods:
        raise ValueError("Invalid method %s. Must be one of: %s" % (method, str(methods)))


def warn_if_coincident_positions(tree, softening=None):
    """Warn when the build found bit-identical positions (frontend.py:53-70)."""
    if not tree.HasCoincidentPoints:
        return
    if softening is not None and bool((softening > 0).any()):
        warnings.warn(
            "Warning: Particle positions are non-unique. Softening will determine the answer for overlapping particles."
        )
        return
    warnings.warn(
        "Warning: Particle positions are non-unique. The answer will be singular or garbage for overlapping particles."
    )


def _check_device(device):
    if device not in ("cpu", "cuda"):
        raise ValueError(f"device must be 'cpu' or 'cuda', got {device!r}")


class Octree:
    """GPU-resident octree with the reference Octree's attributes (octree.py:8-28).

    Particles occupy [0, NumParticles) in Morton order, the root is NumParticles, nodes follow in
    depth-first preorder.  Array attributes are host numpy views materialised on first access;
    the device arrays stay in ``self.dev`` for the walk.  ``NumNodes`` is the exact element count
    (the reference over-allocates node slots; every slot here is a built node).
    """

    def __init__(self, points, masses, softening, quadrupole=False, compute_moments=True):
        self.dev = engine.DeviceTree(points, masses, softening, quadrupole=quadrupole,
                                     compute_moments=compute_moments)
        self.NumParticles = self.dev.n
        self.NumNodes = self.dev.ne
        self.HasQuads = bool(quadrupole)
        self.HasCoincidentPoints = self.dev.has_coincident
        self._host = {}

    def _get(self, name, fn):
        if name not in self._host:
            self._host[name] = fn()
        return self._host[name]

    @property
    def Coordinates(self):
        return self._get("Coordinates", lambda: self.dev.coordinates.cpu().numpy())

    @property
    def Masses(self):
        return self._get("Masses", lambda: self.dev.masses.cpu().numpy())

    @property
    def Softenings(self):
        return self._get("Softenings", lambda: self.dev.softenings.cpu().numpy())

    def _node_rows(self, t):
        a = t.cpu().numpy()
        a[: self.dev.n] = 0.0  # particle rows are not written on the device; a particle has zero extent
        return a

    @property
    def Sizes(self):
        return self._get("Sizes", lambda: self._node_rows(self.dev.sizes))

    @property
    def Deltas(self):
        return self._get("Deltas", lambda: self._node_rows(self.dev.deltas))

    @property
    def Quadrupoles(self):
        if not self.HasQuads:
            return np.zeros((0, 3, 3))
        def host():
            q = self.dev.quadrupoles.cpu().numpy().reshape(-1, 3, 3)
            q[: self.dev.n] = 0.0  # particle rows are not written on the device; a particle has none
            return q

        return self._get("Quadrupoles", host)

    @property
    def FirstSubnode(self):
        return self._get("FirstSubnode", lambda: self.dev.first_subnode.cpu().numpy().astype(np.int64))

    @property
    def NextBranch(self):
        return self._get("NextBranch", lambda: self.dev.next_branch.cpu().numpy().astype(np.int64))

    @property
    def TreewalkIndices(self):
        return self._get("TreewalkIndices", lambda: self.dev.order.cpu().numpy())


def ConstructTree(pos, m=None, softening=None, quadrupole=False, vel=None, compute_moments=True, morton_order=True,
                  radix=True):
    """Build an octree on the GPU (frontend.py:73-142).

    Always the radix build: ``morton_order``/``radix`` are accepted for compatibility (the
    reference's insertion build is not part of this library), and ``vel`` is rejected because the
    velocity-carrying DynamicOctree serves the correlation functions, which are not provided.
    """
    if vel is not None:
        raise NotImplementedError("ConstructTree(vel=...) builds the reference's DynamicOctree, which serves the "
                                  "correlation functions; it is not provided by pytreegrav_amd")
    pos = _atleast_2d(_f64(pos))
    if m is None:
        m = _zeros_like_len(pos)
        compute_moments = False
    m = _atleast_1d(_f64(m))
    if softening is None:
        softening = m * 0 if isinstance(m, torch.Tensor) else np.zeros_like(m)
    softening = _atleast_1d(_f64(softening))
    msg = "Invalid input detected - aborting treebuild to avoid going into an infinite loop!"
    # host arrays are checked here; device tensors by the build itself (no extra passes)
    if not all(isinstance(a, torch.Tensor) for a in (pos, m, softening)) and not _finite(pos, m, softening):
        print(msg)
        raise RuntimeError(msg)
    if len(pos) == 0:
        raise ValueError("ConstructTree needs at least one particle")
    try:
        tree = Octree(pos, m, softening, quadrupole=quadrupole, compute_moments=compute_moments)
    except engine.NonFiniteInput:
        print(msg)
        raise RuntimeError(msg) from None
    warn_if_coincident_positions(tree, softening)
    return tree


def _morton_order(points):
    """Morton permutation of arbitrary points (grouped_treewalk.py:269-275); device int64 tensor."""
    return engine.morton_order(points)


def FieldsTarget_grouped(pos, soft, tree, potential=False, accel=False, tidal=False, group_size=8, G=1.0, theta=0.7,
                         quadrupole=False, parallel=True, precision="float64"):
    """Any subset of potential/accel/tidal at ``pos`` (already in a spatially compact order) from
    ``tree`` in one traversal (grouped_treewalk.py:350-397).  Returns a dict of arrays."""
    if not (potential or accel or tidal):
        raise ValueError("request at least one of potential=True, accel=True, tidal=True")
    checkTreeQuadrupoles(tree, quadrupole)
    on_dev = _is_dev(pos)
    pos = _atleast_2d(_f64(pos))
    mask = engine.fields_mask(potential, accel, tidal)
    perm = torch.arange(len(pos), dtype=torch.int64, device=engine.device())
    out = engine.walk(tree.dev, pos, _atleast_1d(_f64(soft)), perm, mask, quadrupole=quadrupole,
                      group_size=_group(group_size, _precision(precision)), theta=theta, G=G,
                      precision=_precision(precision))
    return {k: _finish(v, on_dev) for k, v in engine.unpack_fields(out, mask).items()}


def AccelTarget_grouped(pos, soft, tree, group_size=8, G=1.0, theta=0.7, quadrupole=False, parallel=True,
                        precision="float64"):
    """grouped_treewalk.py:285-311"""
    return FieldsTarget_grouped(pos, soft, tree, accel=True, group_size=group_size, G=G, theta=theta,
                                quadrupole=quadrupole, precision=precision)["accel"]


def PotentialTarget_grouped(pos, soft, tree, group_size=8, G=1.0, theta=0.7, quadrupole=False, parallel=True,
                            precision="float64"):
    """grouped_treewalk.py:314-329"""
    return FieldsTarget_grouped(pos, soft, tree, potential=True, group_size=group_size, G=G, theta=theta,
                                quadrupole=quadrupole, precision=precision)["potential"]


def TidalTensorTarget_grouped(pos, soft, tree, group_size=8, G=1.0, theta=0.7, quadrupole=False, parallel=True,
                              precision="float64"):
    """grouped_treewalk.py:332-347"""
    return FieldsTarget_grouped(pos, soft, tree, tidal=True, group_size=group_size, G=G, theta=theta,
                                quadrupole=quadrupole, precision=precision)["tidal"]


# ---------------------------------------------------------------------------------------------
# shared drivers
# ---------------------------------------------------------------------------------------------
def _self_fields(pos, m, softening, mask, G, theta, tree, return_tree, parallel, method, quadrupole, group_size,
                 precision, adaptive_n, criterion="theta", alpha=None, expansion=None, expansion_rel=3.0):
    """Potential/Accel/TidalTensor (frontend.py:145-293, 421-560, 688-789).

    criterion "error"/"learned" (float32 tree walks only, no reference counterpart): accept a node
    only when its estimated (analytic or learned) monopole error is below alpha x the RMS field;
    theta acts as the reference's geometric guard (GUARD_THETA unless the caller gave one).  expansion (float32
    monopole walks of the potential and/or acceleration, no reference counterpart): fold accepted
    elements far enough from a target group into a third-order expansion about the group's centre
    when its truncation bound is below expansion x the RMS field and below expansion_rel x the
    element's own field there (0: no such cap)."""
    valueTestMethod(method)
    if expansion is not None:
        if precision != "float32" or quadrupole or (mask & PTG_TIDAL):
            raise ValueError("expansion runs on the float32 walk with monopole interactions, for the potential "
                             "and/or the acceleration")
        if not expansion > 0:
            raise ValueError("expansion must be a positive tolerance (relative to the RMS field)")
        if not expansion_rel >= 0:
            raise ValueError("expansion_rel must be >= 0 (0: no cap)")
        method = "tree"
    if criterion not in ("theta", "error", "learned"):
        raise ValueError(f"criterion must be 'theta', 'error' or 'learned', got {criterion!r}")
    if criterion != "theta":
        if precision != "float32":
            raise ValueError(f"criterion={criterion!r} runs on the float32 walk (device='cuda' or precision='float32')")
        if alpha is None or not alpha > 0:
            raise ValueError(f"criterion={criterion!r} needs alpha > 0 (per-node tolerance relative to the RMS field)")
        method = "tree"
    on_dev = _is_dev(pos)
    pos = _atleast_2d(_f64(pos))
    m = _atleast_1d(_f64(m))
    if softening is None:
        softening = m * 0 if isinstance(m, torch.Tensor) else np.zeros_like(m)
    softening = _atleast_1d(_f64(softening))
    if method == "adaptive":
        # the reference crosses over where its 32-thread CPU curves do (frontend.py:210-220); float64
        # keeps those thresholds, since the choice changes the answer and float64 reproduces the
        # reference, while float32 uses the crossover measured for the GPU walk against the GPU
        # direct sum (FP32_TREE_MIN_N)
        n_min = FP32_TREE_MIN_N if precision == "float32" else (4000 if parallel else 1000)
        if tree is not None or len(pos) > n_min:
            method = "tree"
        else:
            method = "bruteforce"
    if method == "bruteforce":
        out = engine.direct(pos, softening, pos, m, softening, mask, G=G, self_mode=True, precision=precision)
        if return_tree:
            tree = None
    elif criterion != "theta":
        if tree is None:
            tree = ConstructTree(pos, m, softening, quadrupole=True)  # |Q| drives the criterion
            out = engine.self_fields(tree.dev, pos, softening, mask, theta=theta, G=G, quadrupole=quadrupole,
                                     group_size=64 if group_size == "auto" else group_size, criterion=criterion,
                                     alpha=alpha, expansion=expansion, expansion_rel=expansion_rel)
        else:
            checkTreeQuadrupoles(tree, True)
            fs = engine.field_scale(tree.dev)
            out = engine.walk(tree.dev, pos, softening, _morton_order(pos), mask, quadrupole=quadrupole,
                              group_size=group_size, theta=theta, G=G, precision=precision, criterion=criterion,
                              eps=alpha * fs, model=_learned_model() if criterion == "learned" else None,
                              expansion=None if expansion is None else expansion * fs, expansion_rel=expansion_rel)
    elif expansion is not None:
        if tree is None:
            tree = ConstructTree(pos, m, softening)
            out = engine.self_fields(tree.dev, pos, softening, mask, theta=theta, G=G,
                                     group_size=64 if group_size == "auto" else group_size, expansion=expansion,
                                     expansion_rel=expansion_rel)
        else:
            out = engine.walk(tree.dev, pos, softening, _morton_order(pos), mask, group_size=group_size, theta=theta,
                              G=G, precision=precision, expansion=expansion * engine.field_scale(tree.dev),
                              expansion_rel=expansion_rel)
    else:
        if tree is None:
            tree = ConstructTree(pos, m, softening, quadrupole=quadrupole)
            idx = tree.dev.order  # built from pos, so its walk order already indexes pos
        else:
            idx = _morton_order(pos)  # sort, do not reuse the stored permutation (frontend.py:253-258)
        checkTreeQuadrupoles(tree, quadrupole)
        out = engine.walk(tree.dev, pos, softening, idx, mask, quadrupole=quadrupole, group_size=group_size,
                          theta=theta, G=G, precision=precision)
    return out, tree, on_dev


def _target_fields(pos_target, pos_source, m_source, softening_target, softening_source, mask, G, theta, tree,
                   return_tree, method, quadrupole, group_size, precision):
    """PotentialTarget/AccelTarget/TidalTensorTarget (frontend.py:296-418, 563-685, 792-901)."""
    valueTestMethod(method)
    if tree is None and (pos_source is None or m_source is None):
        raise ValueError("Must pass either pos_source & m_source or source tree.")
    on_dev = _is_dev(pos_target)
    pos_target = _atleast_2d(_f64(pos_target))
    if softening_target is None:
        softening_target = _zeros_like_len(pos_target)
    softening_target = _atleast_1d(_f64(softening_target))
    if pos_source is not None:
        pos_source = _atleast_2d(_f64(pos_source))
        m_source = _atleast_1d(_f64(m_source))
        if softening_source is None:
            softening_source = _zeros_like_len(pos_source)
        softening_source = _atleast_1d(_f64(softening_source))
    if method == "adaptive":
        if pos_source is None or len(pos_target) * len(pos_source) > 10 ** 6:
            method = "tree"
        else:
            method = "bruteforce"
    if method == "bruteforce":
        out = engine.direct(pos_target, softening_target, pos_source, m_source, softening_source, mask, G=G,
                            self_mode=False, precision=precision)
        if return_tree:
            tree = None
    else:
        if tree is None:
            tree = ConstructTree(pos_source, m_source, softening_source, quadrupole=quadrupole)
        checkTreeQuadrupoles(tree, quadrupole)
        tsort = _morton_order(pos_target)
        out = engine.walk(tree.dev, pos_target, softening_target, tsort, mask, quadrupole=quadrupole,
                          group_size=group_size, theta=theta, G=G, precision=precision)
    return out, tree, on_dev


def _ret(val, tree, return_tree):
    return (val, tree) if return_tree else val


_model = {}


def _learned_model():
    if "m" not in _model:
        from .learned import OpeningMLP

        _model["m"] = OpeningMLP.load()
    return _model["m"]


# ---------------------------------------------------------------------------------------------
# public API (frontend.py:145-901)
# ---------------------------------------------------------------------------------------------
def Potential(pos, m, softening=None, G=1.0, theta=THETA_DEFAULT, tree=None, return_tree=False, parallel=False,
              method="adaptive", quadrupole=False, group_size=None, device="cpu", precision=None, criterion="theta",
              alpha=None, expansion=None, expansion_rel=3.0):
    """Gravitational potential at the particle positions, shape (N,) (frontend.py:145-293).

    criterion/alpha/expansion/expansion_rel (float32 only): see Accel."""
    valueTestMethod(method)
    _check_device(device)
    prec = _precision(precision, device)
    if criterion != "theta" and theta is THETA_DEFAULT:
        theta = GUARD_THETA
    out, tree, on_dev = _self_fields(pos, m, softening, PTG_POT, G, theta, tree, return_tree, parallel, method,
                                     quadrupole, _group(group_size, prec), prec, None, criterion, alpha, expansion,
                                     expansion_rel)
    return _ret(_finish(out[:, 0], on_dev), tree, return_tree)


def Accel(pos, m, softening=None, G=1.0, theta=THETA_DEFAULT, tree=None, return_tree=False, parallel=False,
          method="adaptive", quadrupole=False, group_size=None, device="cpu", precision=None, criterion="theta",
          alpha=None, expansion=None, expansion_rel=3.0):
    """Gravitational acceleration at the particle positions, shape (N,3) (frontend.py:421-560).

    criterion="learned" / "error" with alpha (float32 only): the learned or analytic per-node error
    bound on top of the geometric criterion, which then acts as a guard at theta (GUARD_THETA = 1.0
    unless the caller passes a theta).
    expansion (float32, monopole): fold far accepted elements into a third-order expansion about
    each target group's centre when its truncation bound is below expansion x the RMS field and
    below expansion_rel x the element's own field there (default 3; 0 = no such cap: faster, with
    a larger worst-case error)."""
    valueTestMethod(method)
    _check_device(device)
    prec = _precision(precision, device)
    if criterion != "theta" and theta is THETA_DEFAULT:
        theta = GUARD_THETA
    out, tree, on_dev = _self_fields(pos, m, softening, PTG_ACC, G, theta, tree, return_tree, parallel, method,
                                     quadrupole, _group(group_size, prec), prec, None, criterion, alpha, expansion,
                                     expansion_rel)
    return _ret(_finish(out, on_dev), tree, return_tree)


def TidalTensor(pos, m, softening=None, G=1.0, theta=0.7, tree=None, return_tree=False, parallel=False,
                method="adaptive", quadrupole=False, group_size=None, precision="float64"):
    """Tidal tensor T_ij = dg_i/dx_j at the particle positions, shape (N,3,3) (frontend.py:688-789)."""
    prec = _precision(precision)
    out, tree, on_dev = _self_fields(pos, m, softening, PTG_TIDAL, G, theta, tree, return_tree, parallel, method,
                                     quadrupole, _group(group_size, prec), prec, None)
    return _ret(_finish(out.reshape(-1, 3, 3), on_dev), tree, return_tree)


def PotentialTarget(pos_target, pos_source, m_source, softening_target=None, softening_source=None, G=1.0, theta=0.7,
                    tree=None, return_tree=False, parallel=False, method="adaptive", quadrupole=False,
                    group_size=None, precision="float64"):
    """Potential at arbitrary targets from sources, shape (N,) (frontend.py:296-418)."""
    prec = _precision(precision)
    out, tree, on_dev = _target_fields(pos_target, pos_source, m_source, softening_target, softening_source, PTG_POT,
                                       G, theta, tree, return_tree, method, quadrupole, _group(group_size, prec), prec)
    return _ret(_finish(out[:, 0], on_dev), tree, return_tree)


def AccelTarget(pos_target, pos_source, m_source, softening_target=None, softening_source=None, G=1.0, theta=0.7,
                tree=None, return_tree=False, parallel=False, method="adaptive", quadrupole=False, group_size=None,
                precision="float64"):
    """Acceleration at arbitrary targets from sources, shape (N,3) (frontend.py:563-685)."""
    prec = _precision(precision)
    out, tree, on_dev = _target_fields(pos_target, pos_source, m_source, softening_target, softening_source, PTG_ACC,
                                       G, theta, tree, return_tree, method, quadrupole, _group(group_size, prec), prec)
    return _ret(_finish(out, on_dev), tree, return_tree)


def TidalTensorTarget(pos_target, pos_source, m_source, softening_target=None, softening_source=None, G=1.0,
                      theta=0.7, tree=None, return_tree=False, parallel=False, method="adaptive", quadrupole=False,
                      group_size=None, precision="float64"):
    """Tidal tensor at arbitrary targets, shape (N,3,3) (frontend.py:792-901)."""
    prec = _precision(precision)
    out, tree, on_dev = _target_fields(pos_target, pos_source, m_source, softening_target, softening_source,
                                       PTG_TIDAL, G, theta, tree, return_tree, method, quadrupole,
                                       _group(group_size, prec), prec)
    return _ret(_finish(out.reshape(-1, 3, 3), on_dev), tree, return_tree)


class Field:
    """A reusable gravitational field: build once, evaluate any field subset in one traversal
    (frontend.py:904-1044).  The tree stays resident on the GPU between calls."""

    def __init__(self, pos, m, softening=None, G=1.0, theta=0.7, quadrupole=False, group_size=None, parallel=False,
                 precision="float64"):
        pos = _atleast_2d(_f64(pos))
        m = _atleast_1d(_f64(m))
        if softening is None:
            softening = m * 0 if isinstance(m, torch.Tensor) else np.zeros_like(m)
        softening = _atleast_1d(_f64(softening))
        self.G = G
        self.theta = theta
        self.quadrupole = quadrupole
        self.precision = _precision(precision)
        self.group_size = _group(group_size, self.precision)
        self.parallel = parallel
        self.tree = ConstructTree(pos, m, softening, quadrupole=quadrupole)
        self._pos = engine.to_dev(pos, (-1, 3))
        self._soft = engine.to_dev(softening, (-1,))

    def __repr__(self):
        return (f"<pytreegrav_amd.Field: {self.tree.NumParticles} particles, "
                f"{'quadrupole' if self.quadrupole else 'monopole'}, theta={self.theta}, "
                f"G={self.G}, group_size={self.group_size}, precision={self.precision}>")

    def evaluate(self, potential=False, accel=False, tidal=False, pos_target=None, softening_target=None, theta=None,
                 parallel=None, group_size=None):
        """Any combination of fields from one traversal; dict in the order of pos_target (or the sources)."""
        if not (potential or accel or tidal):
            raise ValueError("request at least one of potential=True, accel=True, tidal=True")
        theta = self.theta if theta is None else theta
        group_size = self.group_size if group_size is None else _group(group_size, self.precision)
        mask = engine.fields_mask(potential, accel, tidal)
        if pos_target is None:
            on_dev = False
            pos, soft, perm = self._pos, self._soft, self.tree.dev.order
        else:
            on_dev = _is_dev(pos_target)
            pos = _atleast_2d(_f64(pos_target))
            if softening_target is None:
                softening_target = _zeros_like_len(pos)
            soft = _atleast_1d(_f64(softening_target))
            perm = _morton_order(pos)
        out = engine.walk(self.tree.dev, pos, soft, perm, mask, quadrupole=self.quadrupole, group_size=group_size,
                          theta=theta, G=self.G, precision=self.precision)
        return {k: _finish(v, on_dev) for k, v in engine.unpack_fields(out, mask).items()}

    def potential(self, pos_target=None, **kwargs):
        return self.evaluate(potential=True, pos_target=pos_target, **kwargs)["potential"]

    def accel(self, pos_target=None, **kwargs):
        return self.evaluate(accel=True, pos_target=pos_target, **kwargs)["accel"]

    def tidal(self, pos_target=None, **kwargs):
        return self.evaluate(tidal=True, pos_target=pos_target, **kwargs)["tidal"]


# ---------------------------------------------------------------------------------------------
# column density (frontend.py:1321-1457)
# ---------------------------------------------------------------------------------------------
# below this many targets the reference walks each target alone; at or above it, groups of
# group_size targets share a traversal (treewalk.py:19); only the binned estimator depends on it
COLUMN_GROUP_MIN_TARGETS = 1000


def ColumnDensity(pos, m, radii, rays=None, randomize_rays=False, healpix=False, tree=None, theta=0.5,
                  return_tree=False, parallel=False, group_size=16, device="cpu"):
    """Column density from each particle's position to infinity, the particles being uniform
    spheres of the given radii (frontend.py:1321-1457).

    rays: None (angle-binned estimator, 6 equal-area sky bins, opening angle theta), an int (6 = the
    six axis directions, otherwise that many random directions from numpy's global generator), or an
    (N_rays, 3) array of directions (normalised here).  randomize_rays rotates the ray grid of each
    target by misc.random_rotation(its Morton rank).  healpix=NSIDE needs healpy.  Returns (N, N_rays)
    (or (N, 6)) float64 in input order.  The ray-traced columns are exact (no opening-angle
    truncation); the binned ones follow the reference's grouped estimator (group_size consecutive
    Morton targets per traversal when there are at least COLUMN_GROUP_MIN_TARGETS of them).
    ``parallel`` is accepted and ignored; ``device`` is validated as the reference does, and both
    values run the FP64 walks on the GPU."""
    if np.any(np.asarray(radii.cpu() if isinstance(radii, torch.Tensor) else radii) <= 0):
        warnings.warn(
            "Some particle radii are <= 0; these particles are point-like and contribute no column "
            "density. Supply a nonzero radius if their mass should be included."
        )
    on_dev = _is_dev(pos)
    pos = _atleast_2d(_f64(pos))
    if tree is None:
        tree = ConstructTree(pos, _f64(m), _f64(radii))
        idx = tree.dev.order  # the tree was built from pos: its walk order is pos's Morton order
    else:
        idx = _morton_order(pos)  # a supplied tree may hold other particles: sort the targets alone
    pos_dev = engine.to_dev(pos, (-1, 3))
    pos_sorted = pos_dev.index_select(0, idx)

    if type(rays) == int:
        if rays == 6:
            rays = np.vstack([np.eye(3), -np.eye(3)])
        else:
            rays = np.random.normal(size=(rays, 3))
    elif type(rays) == np.ndarray:
        if not len(rays.shape) == 2:
            raise Exception("rays array argument must be 2D.")
        elif rays.shape[1] != 3:
            raise Exception("rays array argument is not an array of 3D vectors.")
        rays = np.array(rays, dtype=np.float64)
    elif rays is not None:
        raise Exception("rays argument type is not supported")
    if healpix:
        try:
            import healpy as hp
        except ImportError as e:
            raise ImportError("healpix=... needs healpy: pip install healpy") from e
        rays = np.array(hp.pix2vec(healpix, np.arange(hp.nside2npix(healpix)))).T
    if rays is not None:
        rays = rays / np.sqrt((rays * rays).sum(1))[:, None]

    if device not in ("cpu", "cuda"):
        raise ValueError(f"device must be 'cpu' or 'cuda', got {device!r}")
    if device == "cuda" and (rays is None or randomize_rays):
        raise ValueError("device='cuda' supports only the ray-traced path: pass rays, and not randomize_rays")

    nt = pos_sorted.shape[0]
    if rays is None:
        gs = group_size if nt >= COLUMN_GROUP_MIN_TARGETS else 1
        columns = engine.column_binned(tree.dev, pos_sorted, gs, theta)
    elif randomize_rays:
        # target k (in Morton order) sees the grid rotated by random_rotation(k) (treewalk.py:1061-1078)
        rot = engine.random_rotations(rays, nt)
        columns = engine.column_rays(tree.dev, pos_sorted, rot, per_target=True)
    else:
        columns = engine.column_rays(tree.dev, pos_sorted, rays)
    if bool(torch.isnan(columns).any()):
        print("WARNING some column densities are NaN!")
    unsorted = torch.empty_like(columns)
    unsorted[idx] = columns  # undo the permutation
    columns = _finish(unsorted, on_dev)
    if return_tree:
        return columns, tree
    return columns
