"""ctypes binding of libptg_hip.so (declared in include/ptg.h).

The library is built in-tree by ``python -m pytreegrav_amd.build`` (or ``__graft_entry__.build()``).
There is no fallback: if the library or a GPU is missing, every compute call raises.
"""

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libptg_hip.so")

PTG_POT, PTG_ACC, PTG_TIDAL = 1, 2, 4
PTG_F32, PTG_F64 = 0, 1
PTG_ERR_OVERFLOW = -3

_c = ctypes
_vp = _c.c_void_p
_i64 = _c.c_int64
_i32 = _c.c_int32
_dbl = _c.c_double


class BuildInfo(_c.Structure):
    _fields_ = [
        ("n", _i64),
        ("nnodes", _i64),
        ("has_collisions", _i32),
        ("has_coincident", _i32),
        ("max_level", _i32),
        ("nonfinite", _i32),
        ("center", _dbl * 3),
        ("size", _dbl),
    ]


class TreeArrays(_c.Structure):
    _fields_ = [
        ("coordinates", _vp),
        ("masses", _vp),
        ("softenings", _vp),
        ("sizes", _vp),
        ("deltas", _vp),
        ("quadrupoles", _vp),
        ("first_subnode", _vp),
        ("next_branch", _vp),
        ("treewalk_indices", _vp),
        ("nchild", _vp),
        ("child_base", _vp),
        ("slot", _vp),
        ("keys", _vp),
        ("node_stats", _vp),
        ("node_count", _vp),
    ]


class WalkTree(_c.Structure):
    _fields_ = [
        ("xyzm", _vp),
        ("aux", _vp),
        ("kids", _vp),
        ("quad", _vp),
        ("nslots", _i64),
        ("prec", _i32),
        ("has_quad", _i32),
        ("origin", _dbl * 3),
    ]


class BuildCounts(_c.Structure):
    _fields_ = [
        ("nnodes", _i32),
        ("max_level", _i32),
        ("has_collisions", _i32),
        ("has_coincident", _i32),
        ("nonfinite", _i32),
        ("overflow", _i32),
        ("pad", _i32 * 2),
        ("center", _dbl * 3),
        ("size", _dbl),
    ]


PTG_OVERFLOW_NODES, PTG_OVERFLOW_LEVELS, PTG_OVERFLOW_COLLISIONS = 1, 2, 4
PTG_DEFERRED_COLLISION_CAP = 8192


class Forest(_c.Structure):
    _fields_ = [("nsys", _i64), ("offsets", _vp), ("roots", _vp), ("cubes", _vp)]


class Criterion(_c.Structure):
    _fields_ = [("criterion", _i32), ("group_expansion", _i32), ("eps", _dbl), ("expansion_eps", _dbl),
                ("eps_scale", _c.c_void_p), ("variant", _i32), ("expansion_rel", _c.c_float)]


PTG_CRIT_THETA, PTG_CRIT_QERR, PTG_CRIT_RADIUS = 0, 1, 2
PTG_MLP_NPARAM = 1441
PTG_GROUPS_RUNS, PTG_GROUPS_CELLS = 0, 1
# build phase boundaries recorded by ptg_build_phase_events, in order
PTG_PHASES = ("start", "bbox", "keys", "sort", "count_nodes", "tree_start", "permute", "nodes", "links", "moments",
              "pack")
_F = _c.POINTER(Forest)

# name -> (restype, argtypes)
SIGNATURES = {
    "ptg_version": (_c.c_int, []),
    "ptg_last_error": (_c.c_char_p, []),
    "ptg_build_workspace_bytes": (_c.c_size_t, [_i64, _i64]),
    "ptg_build_phase_events": (_c.c_int, [_c.POINTER(_vp), _c.c_int]),
    "ptg_build_sort": (_c.c_int, [_vp, _vp, _vp, _i64, _F, _vp, _c.c_size_t, _c.POINTER(BuildInfo), _vp]),
    "ptg_node_workspace_bytes": (_c.c_size_t, [_i64, _i64, _c.c_int]),
    "ptg_groups_workspace_bytes": (_c.c_size_t, [_i64, _i64]),
    "ptg_target_groups": (_c.c_int, [_vp, _vp, _vp, _i64, _F, _c.c_int, _c.c_int, _vp, _vp, _vp, _c.c_size_t,
                                     _vp]),
    "ptg_build_tree": (_c.c_int, [_vp, _vp, _vp, _c.POINTER(BuildInfo), _F, _c.c_int, _c.c_int, _vp, _c.c_size_t,
                                  _vp, _c.c_size_t, _c.POINTER(TreeArrays), _c.POINTER(WalkTree), _vp, _vp]),
    "ptg_morton_order": (_c.c_int, [_vp, _i64, _vp, _vp, _c.c_size_t, _vp]),
    "ptg_build_deferred": (_c.c_int, [_vp, _vp, _vp, _i64, _F, _c.c_int, _c.c_int, _i64, _c.c_int, _vp, _c.c_size_t,
                                      _vp, _c.c_size_t, _c.POINTER(TreeArrays), _c.POINTER(WalkTree), _vp, _vp,
                                      _vp]),
    "ptg_field_sample": (_c.c_int, [_vp, _vp, _F, _i64, _c.c_int, _c.c_int, _vp, _vp, _vp, _vp, _vp]),
    "ptg_pack_walk": (_c.c_int, [_c.POINTER(TreeArrays), _i64, _i64, _F, _c.POINTER(WalkTree), _vp]),
    "ptg_walk_workspace_bytes": (_c.c_size_t, [_i64, _c.c_int]),
    "ptg_column_rays": (_c.c_int, [_c.POINTER(TreeArrays), _i64, _i64, _c.POINTER(WalkTree), _vp, _i64, _vp, _i64,
                                   _c.c_int, _vp, _vp, _vp, _vp, _vp]),
    "ptg_column_binned": (_c.c_int, [_c.POINTER(TreeArrays), _i64, _i64, _c.POINTER(WalkTree), _vp, _i64, _i64, _dbl,
                                     _vp, _vp, _vp, _c.c_size_t, _vp, _vp]),
    "ptg_column_workspace_bytes": (_c.c_size_t, [_i64]),
    "ptg_random_rotations": (_c.c_int, [_i64, _i64, _vp, _i64, _vp, _vp]),
    "ptg_walk": (_c.c_int, [_c.POINTER(WalkTree), _F, _vp, _vp, _vp, _c.c_int, _i64, _c.c_int, _c.c_int, _c.c_int,
                            _vp, _vp,
                            _dbl, _c.POINTER(Criterion), _dbl, _c.c_int, _vp, _vp, _vp, _c.c_size_t, _vp]),
    "ptg_walk_check": (_c.c_int, [_vp, _vp]),
    "ptg_walk_subtree": (_c.c_int, [_c.POINTER(WalkTree), _vp, _vp, _vp, _vp, _i64, _c.c_int, _c.c_int, _dbl, _dbl,
                                    _vp, _vp, _vp, _c.c_size_t, _vp]),
    "ptg_walk_noalloc": (_c.c_int, [_c.POINTER(WalkTree), _F, _vp, _vp, _vp, _c.c_int, _i64, _c.c_int, _c.c_int,
                                    _c.c_int, _vp, _vp, _dbl, _c.POINTER(Criterion), _dbl, _c.c_int, _vp, _vp, _vp,
                                    _vp]),
    "ptg_node_features": (_c.c_int, [_c.POINTER(TreeArrays), _i64, _i64, _vp, _i64, _dbl, _vp, _vp]),
    "ptg_node_errcurve": (_c.c_int, [_c.POINTER(TreeArrays), _i64, _i64, _F, _vp, _i64, _i64, _c.c_int, _dbl, _dbl,
                                     _c.c_int, _vp, _vp]),
    "ptg_mlp_samples": (_c.c_int, [_c.POINTER(TreeArrays), _i64, _i64, _vp, _vp, _vp, _i64, _vp, _c.c_int, _dbl,
                                   _dbl, _vp, _vp, _vp, _vp]),
    "ptg_mlp_grad_workspace_bytes": (_c.c_size_t, [_i64]),
    "ptg_mlp_grad": (_c.c_int, [_vp, _vp, _vp, _i64, _vp, _c.c_uint64, _i64, _c.c_float, _vp, _vp, _c.c_size_t,
                                _vp]),
    "ptg_mlp_adam": (_c.c_int, [_vp, _vp, _vp, _vp, _i64, _c.c_float, _c.c_float, _c.c_float, _c.c_float, _vp]),
    "ptg_mlp_forward": (_c.c_int, [_vp, _vp, _i64, _c.c_int, _vp, _vp]),
    "ptg_open_radii": (_c.c_int, [_c.POINTER(TreeArrays), _i64, _i64, _F, _c.POINTER(WalkTree), _vp, _c.c_float,
                                  _dbl, _vp, _c.c_int, _vp]),
    "ptg_rms_rows": (_c.c_int, [_vp, _i64, _c.c_int, _c.c_int, _c.c_int, _i64, _vp, _vp]),
    "ptg_direct": (_c.c_int, [_vp, _vp, _i64, _vp, _vp, _vp, _i64, _c.c_int, _c.c_int, _c.c_int, _dbl, _vp, _vp]),
}

_lib = None


class PtgError(RuntimeError):
    """A libptg_hip call returned a nonzero status."""


def load():
    """Load libptg_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP library first (python -m pytreegrav_amd.build)")
        L = _c.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status, what):
    if status != 0:
        msg = load().ptg_last_error().decode(errors="replace")
        raise PtgError(f"{what} failed with status {status}: {msg}")


def ptr(t):
    """Device pointer of a torch tensor (or None -> NULL)."""
    return None if t is None else _vp(t.data_ptr())
