"""pytreegrav_amd: pytreegrav's tree-gravity API on AMD Instinct MI355X (gfx950).

The public names mirror pytreegrav's (src/pytreegrav/__init__.py re-exports octree and frontend):
ConstructTree, Potential, Accel, PotentialTarget, AccelTarget, TidalTensor, TidalTensorTarget,
Field, ColumnDensity, the grouped-walk functions, and the per-target walks (treewalk.py) and brute-force
sums (bruteforce.py) that the reference's frontend re-exports.  All compute runs in libptg_hip.so (hand-written HIP for
gfx950); there is no CPU fallback.
"""

from .frontend import (  # noqa: F401
    Accel,
    AccelTarget,
    ColumnDensity,
    ConstructTree,
    Field,
    FieldsTarget_grouped,
    AccelTarget_grouped,
    Octree,
    Potential,
    PotentialTarget,
    PotentialTarget_grouped,
    TidalTensor,
    TidalTensorTarget,
    TidalTensorTarget_grouped,
    _morton_order,
    checkTreeQuadrupoles,
    valueTestMethod,
    warn_if_coincident_positions,
)
from .forest import BatchedAccelPotential, Forest  # noqa: F401
from .treewalk import (  # noqa: F401
    ColumnDensity_grouped,
    ColumnDensity_grouped_parallel,
    ColumnDensity_tree,
    ColumnDensity_tree_parallel,
    ColumnDensityBinned_grouped,
    ColumnDensityBinned_grouped_parallel,
    AccelTarget_tree,
    AccelTarget_tree_parallel,
    AccelWalk,
    AccelWalk_quad,
    PotentialTarget_tree,
    PotentialTarget_tree_parallel,
    PotentialWalk,
    PotentialWalk_quad,
    acceptance_criterion,
    subtree_end,
)
from .bruteforce import (  # noqa: F401
    SYMMETRIC_NMIN,
    AccelTarget_bruteforce,
    AccelTarget_bruteforce_parallel,
    Accel_bruteforce,
    Accel_bruteforce_parallel,
    Accel_bruteforce_symmetric,
    PotentialTarget_bruteforce,
    PotentialTarget_bruteforce_parallel,
    Potential_bruteforce,
    Potential_bruteforce_parallel,
    Potential_bruteforce_symmetric,
    TidalTensorTarget_bruteforce,
    TidalTensorTarget_bruteforce_parallel,
    TidalTensor_bruteforce,
    TidalTensor_bruteforce_parallel,
)

__version__ = "0.1.0"
