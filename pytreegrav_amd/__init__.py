"""pytreegrav_amd: pytreegrav's tree-gravity API on AMD Instinct MI355X (gfx950).

The public names mirror pytreegrav's (src/pytreegrav/__init__.py re-exports octree and frontend):
ConstructTree, Potential, Accel, PotentialTarget, AccelTarget, TidalTensor, TidalTensorTarget,
Field, ColumnDensity, and the grouped-walk functions.  All compute runs in libptg_hip.so (hand-written HIP for
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

__version__ = "0.1.0"
