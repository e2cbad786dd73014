"""ctypes wrapper of the CPU oracle (``ptg_oracle.c``).

TEST INFRASTRUCTURE ONLY: the checker the HIP library is compared against, and the CPU baseline
timed beside it.  Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may
import this package; the product (``pytreegrav_amd``) never does.

Each wrapper mirrors one reference function (file:line in the C source).  Arrays come back as
float64/int64 numpy arrays in the reference's layout.
"""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libptg_oracle.so")
_lib = None

_d = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p


def build_library(force=False):
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build_library()
        L = ctypes.CDLL(_LIB_PATH)
        L.ptgo_force_kernel.restype = ctypes.c_double
        L.ptgo_force_kernel.argtypes = [ctypes.c_double, ctypes.c_double]
        L.ptgo_potential_kernel.restype = ctypes.c_double
        L.ptgo_potential_kernel.argtypes = [ctypes.c_double, ctypes.c_double]
        L.ptgo_tidal_kernel.restype = ctypes.c_double
        L.ptgo_tidal_kernel.argtypes = [ctypes.c_double, ctypes.c_double]
        L.ptgo_spread_bits.restype = ctypes.c_int64
        L.ptgo_spread_bits.argtypes = [ctypes.c_int64]
        L.ptgo_morton_keys.argtypes = [_d, ctypes.c_int64, _d, ctypes.c_double, _i64]
        L.ptgo_radix_argsort.argtypes = [_i64, ctypes.c_int64, _i64]
        L.ptgo_bounding_cube.argtypes = [_d, ctypes.c_int64, _d, _d]
        L.ptgo_morton_order.argtypes = [_d, ctypes.c_int64, _i64]
        L.ptgo_column_rays.argtypes = [_vp, _d, ctypes.c_int64, _d, ctypes.c_int64, ctypes.c_int, _d, ctypes.c_int]
        L.ptgo_column_binned.argtypes = [_vp, _d, ctypes.c_int64, ctypes.c_int64, ctypes.c_double, _d, ctypes.c_int]
        L.ptgo_build.restype = _vp
        L.ptgo_build.argtypes = [_d, _d, _d, ctypes.c_int64, ctypes.c_int, ctypes.c_int]
        L.ptgo_tree_free.argtypes = [_vp]
        L.ptgo_tree_info.argtypes = [_vp, _i64]
        L.ptgo_tree_get.restype = ctypes.c_int
        L.ptgo_tree_get.argtypes = [_vp, ctypes.c_int, ctypes.c_int64, _vp]
        L.ptgo_fields_grouped.argtypes = [_vp, _d, _d, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_int, _d, _i64]
        L.ptgo_fields_target_bruteforce.argtypes = [_d, _d, ctypes.c_int64, _d, _d, _d, ctypes.c_int64,
                                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                    ctypes.c_int, _d]
        L.ptgo_fields_self_bruteforce.argtypes = [_d, _d, _d, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                                  ctypes.c_int, ctypes.c_double, ctypes.c_int, _d]
        _lib = L
    return _lib


def _f64(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if shape is not None:
        a = a.reshape(shape)
    return a


def _p(a, t=_d):
    return a.ctypes.data_as(t)


# ---- kernels (kernel.py:4-75) ----
def force_kernel(r, h):
    return lib().ptgo_force_kernel(float(r), float(h))


def potential_kernel(r, h):
    return lib().ptgo_potential_kernel(float(r), float(h))


def tidal_kernel(r, h):
    return lib().ptgo_tidal_kernel(float(r), float(h))


# ---- keys / sort (octree.py:54-132, 539-574; grouped_treewalk.py:269-275) ----
def spread_bits(x):
    return lib().ptgo_spread_bits(int(x))


def bounding_cube(pts):
    pts = _f64(pts, (-1, 3))
    c = np.zeros(3)
    s = np.zeros(1)
    lib().ptgo_bounding_cube(_p(pts), len(pts), _p(c), _p(s))
    return c, float(s[0])


def morton_keys(pts, center, size):
    pts = _f64(pts, (-1, 3))
    c = _f64(center, (3,))
    out = np.empty(len(pts), np.int64)
    lib().ptgo_morton_keys(_p(pts), len(pts), _p(c), float(size), _p(out, _i64))
    return out


def radix_argsort(keys):
    keys = np.ascontiguousarray(keys, dtype=np.int64)
    out = np.empty(len(keys), np.int64)
    lib().ptgo_radix_argsort(_p(keys, _i64), len(keys), _p(out, _i64))
    return out


def morton_order(pts):
    pts = _f64(pts, (-1, 3))
    out = np.empty(len(pts), np.int64)
    lib().ptgo_morton_order(_p(pts), len(pts), _p(out, _i64))
    return out


# ---- tree (octree.py:664-1298) ----
class OracleTree:
    """The reference Octree's attributes, restated (serial radix split + rollups).

    ``node_hi`` is the first unused node slot; arrays are truncated to it, so ``NumNodes`` here is
    the number of elements actually built (the reference keeps slack capacity beyond it).  The C
    tree stays alive for walks until the object is collected.
    """

    def __init__(self, pos, m, h, quadrupole=False, compute_moments=True):
        pos = _f64(pos, (-1, 3))
        n = len(pos)
        m = _f64(m, (n,))
        h = _f64(h, (n,))
        L = lib()
        self._free = L.ptgo_tree_free  # kept so the destructor still works at interpreter exit
        self.ptr = L.ptgo_build(_p(pos), _p(m), _p(h), n, int(quadrupole), int(compute_moments))
        info = np.zeros(5, np.int64)
        L.ptgo_tree_info(self.ptr, _p(info, _i64))
        self.NumParticles = int(info[0])
        self.Capacity = int(info[1])
        self.node_hi = int(info[2])
        self.NumNodes = self.node_hi
        self.HasQuads = bool(info[3])
        self.HasCoincidentPoints = bool(info[4])
        k = self.node_hi

        def get(which, shape, dtype):
            a = np.empty(shape, dtype)
            assert L.ptgo_tree_get(self.ptr, which, shape[0], a.ctypes.data_as(_vp)) == 0
            return a

        self.Coordinates = get(0, (k, 3), np.float64)
        self.Masses = get(1, (k,), np.float64)
        self.Softenings = get(2, (k,), np.float64)
        self.Sizes = get(3, (k,), np.float64)
        self.Deltas = get(4, (k,), np.float64)
        self.Quadrupoles = get(5, (k, 3, 3), np.float64) if self.HasQuads else np.zeros((0, 3, 3))
        self.FirstSubnode = get(6, (k,), np.int64)
        self.NextBranch = get(7, (k,), np.int64)
        self.TreewalkIndices = get(8, (n,), np.int64)

    def __del__(self):
        if getattr(self, "ptr", None) and getattr(self, "_free", None):
            self._free(self.ptr)
            self.ptr = None


def build_tree(pos, m, h, quadrupole=False, compute_moments=True):
    return OracleTree(pos, m, h, quadrupole=quadrupole, compute_moments=compute_moments)


def fields_grouped(tree, pos, soft, potential=False, accel=False, tidal=False, quadrupole=False, group_size=8,
                   G=1.0, theta=0.7, nthreads=0, return_counters=False):
    """grouped_treewalk.FieldsTarget_grouped (:350-397); pos/soft in the order given (no sorting)."""
    pos = _f64(pos, (-1, 3))
    n = len(pos)
    soft = _f64(soft, (n,))
    W = int(potential) + 3 * int(accel) + 9 * int(tidal)
    out = np.zeros((n, W))
    ctr = np.zeros(2, np.int64)
    if n:
        lib().ptgo_fields_grouped(tree.ptr, _p(pos), _p(soft), n, int(potential), int(accel), int(tidal),
                                  int(quadrupole), int(group_size), float(G), float(theta), int(nthreads),
                                  _p(out), _p(ctr, _i64))
    res = _unpack(out, potential, accel, tidal)
    if return_counters:
        return res, {"interactions": int(ctr[0]), "visits": int(ctr[1])}
    return res


def _unpack(out, potential, accel, tidal):
    res = {}
    o = 0
    if potential:
        res["potential"] = out[:, 0].copy()
        o = 1
    if accel:
        res["accel"] = out[:, o:o + 3].copy()
        o += 3
    if tidal:
        res["tidal"] = out[:, o:o + 9].reshape(-1, 3, 3).copy()
    return res


def fields_target_bruteforce(xt, ht, xs, ms, hs, potential=False, accel=False, tidal=False, G=1.0, nthreads=0):
    """bruteforce.py target forms (:7-50, :209-256, :353-386)."""
    xt = _f64(xt, (-1, 3))
    xs = _f64(xs, (-1, 3))
    nt, ns = len(xt), len(xs)
    ht = _f64(ht, (nt,))
    ms = _f64(ms, (ns,))
    hs = _f64(hs, (ns,))
    W = int(potential) + 3 * int(accel) + 9 * int(tidal)
    out = np.zeros((nt, W))
    if nt:
        lib().ptgo_fields_target_bruteforce(_p(xt), _p(ht), nt, _p(xs), _p(ms), _p(hs), ns, int(potential),
                                            int(accel), int(tidal), float(G), int(nthreads), _p(out))
    return _unpack(out, potential, accel, tidal)


def fields_self_bruteforce(x, m, h, potential=False, accel=False, tidal=False, G=1.0, nthreads=0):
    """bruteforce.py self forms (:86-116, :164-206, :319-350)."""
    x = _f64(x, (-1, 3))
    n = len(x)
    m = _f64(m, (n,))
    h = _f64(h, (n,))
    W = int(potential) + 3 * int(accel) + 9 * int(tidal)
    out = np.zeros((n, W))
    if n:
        lib().ptgo_fields_self_bruteforce(_p(x), _p(m), _p(h), n, int(potential), int(accel), int(tidal),
                                          float(G), int(nthreads), _p(out))
    return _unpack(out, potential, accel, tidal)


def self_fields_tree(pos, m, h, potential=False, accel=False, tidal=False, quadrupole=False, group_size=8, G=1.0,
                     theta=0.7, nthreads=0, return_counters=False):
    """frontend.Potential/Accel tree path (:243-288, :513-555): build, walk in tree order, scatter back."""
    pos = _f64(pos, (-1, 3))
    h = _f64(h, (len(pos),))
    t = OracleTree(pos, m, h, quadrupole=quadrupole)
    idx = t.TreewalkIndices
    res = fields_grouped(t, pos[idx], h[idx], potential=potential, accel=accel, tidal=tidal, quadrupole=quadrupole,
                         group_size=group_size, G=G, theta=theta, nthreads=nthreads, return_counters=return_counters)
    ctr = None
    if return_counters:
        res, ctr = res
    out = {}
    for k, v in res.items():
        o = np.empty_like(v)
        o[idx] = v
        out[k] = o
    return (out, ctr) if return_counters else out


# ---- column density (treewalk.py:821-1276) ----
def column_rays(tree, pos, rays, per_target=False, nthreads=0):
    """Ray-traced column density of every target along every ray: (nt, nr); rays (nr, 3) shared or
    (nt, nr, 3) per target (ColumnDensityWalk_singleray per pair, treewalk.py:904-973)."""
    pos = _f64(pos, (-1, 3))
    nt = len(pos)
    rays = np.ascontiguousarray(rays, dtype=np.float64)
    nr = rays.shape[-2] if rays.ndim == 3 else rays.shape[0]
    out = np.zeros((nt, nr))
    lib().ptgo_column_rays(tree.ptr, _p(pos), nt, _p(rays), nr, int(bool(per_target)), _p(out),
                           nthreads or os.cpu_count() or 1)
    return out


def column_binned(tree, pos, group_size=16, theta=0.5, nthreads=0):
    """Angular-binned column density, groups of group_size consecutive targets
    (ColumnDensityBinned_grouped, treewalk.py:1197-1276): (nt, 6)."""
    pos = _f64(pos, (-1, 3))
    nt = len(pos)
    out = np.zeros((nt, 6))
    lib().ptgo_column_binned(tree.ptr, _p(pos), nt, int(group_size), float(theta), _p(out),
                             nthreads or os.cpu_count() or 1)
    return out
