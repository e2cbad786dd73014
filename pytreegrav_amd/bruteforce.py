"""The reference's brute-force sums (bruteforce.py:7-386, bruteforce_symmetric.py:97-158) on the GPU.

Every function runs the LDS-tiled FP64 direct sum of libptg_hip.so (ptg_direct) with the
reference's pair rules: the larger of the two softenings and the spline kernels inside it; a
particle's own term, a target sitting on a source and the force of two coincident particles are
skipped, while the self potential of two distinct coincident particles keeps their softened term
m PotentialKernel(0, h) (bruteforce.py:70-82; nothing when h = 0).  The
reference's ``_parallel`` and ``_symmetric`` forms (thread-parallel CPU schedules of the same sums,
equal to the serial ones to 1e-12 in its own tests) are the same GPU sum here.  Host arrays give
numpy results, device tensors give device tensors.
"""

from . import engine
from .frontend import _atleast_1d, _atleast_2d, _f64, _finish, _is_dev, _zeros_like_len

# the reference's particle count above which its frontend takes the symmetric CPU sum
# (bruteforce_symmetric.py:27); kept for callers that read it
SYMMETRIC_NMIN = 1000


def _self(x, m, softening, G, mask):
    on_dev = _is_dev(x)
    x = _atleast_2d(_f64(x))
    m = _atleast_1d(_f64(m))
    softening = _zeros_like_len(x) if softening is None else _atleast_1d(_f64(softening))
    out = engine.direct(x, softening, x, m, softening, mask, G=G, self_mode=True)
    return _finish(out, on_dev)


def _target(x_target, softening_target, x_source, m_source, softening_source, G, mask):
    on_dev = _is_dev(x_target)
    x_target = _atleast_2d(_f64(x_target))
    x_source = _atleast_2d(_f64(x_source))
    m_source = _atleast_1d(_f64(m_source))
    softening_target = (_zeros_like_len(x_target) if softening_target is None
                        else _atleast_1d(_f64(softening_target)))
    softening_source = (_zeros_like_len(x_source) if softening_source is None
                        else _atleast_1d(_f64(softening_source)))
    out = engine.direct(x_target, softening_target, x_source, m_source, softening_source, mask, G=G,
                        self_mode=False)
    return _finish(out, on_dev)


_POT = engine.fields_mask(potential=True)
_ACC = engine.fields_mask(accel=True)
_TID = engine.fields_mask(tidal=True)


def Potential_bruteforce(x, m, softening, G=1.0):
    """Mutual potential (N,) of the particles (bruteforce.py:54-83)."""
    return _self(x, m, softening, G, _POT)[:, 0]


def Accel_bruteforce(x, m, softening, G=1.0):
    """Mutual accelerations (N, 3) of the particles (bruteforce.py:120-161)."""
    return _self(x, m, softening, G, _ACC)


def TidalTensor_bruteforce(x, m, softening, G=1.0):
    """Mutual tidal tensors (N, 3, 3) of the particles (bruteforce.py:286-316)."""
    return _self(x, m, softening, G, _TID).reshape(-1, 3, 3)


def PotentialTarget_bruteforce(x_target, softening_target, x_source, m_source, softening_source, G=1.0):
    """Potential (N,) at the targets from the sources (bruteforce.py:7-42)."""
    return _target(x_target, softening_target, x_source, m_source, softening_source, G, _POT)[:, 0]


def AccelTarget_bruteforce(x_target, softening_target, x_source, m_source, softening_source, G=1.0):
    """Acceleration (N, 3) at the targets from the sources (bruteforce.py:209-248)."""
    return _target(x_target, softening_target, x_source, m_source, softening_source, G, _ACC)


def TidalTensorTarget_bruteforce(x_target, softening_target, x_source, m_source, softening_source, G=1.0):
    """Tidal tensors (N, 3, 3) at the targets from the sources (bruteforce.py:353-381)."""
    return _target(x_target, softening_target, x_source, m_source, softening_source, G, _TID).reshape(-1, 3, 3)


Potential_bruteforce_parallel = Potential_bruteforce
Accel_bruteforce_parallel = Accel_bruteforce
TidalTensor_bruteforce_parallel = TidalTensor_bruteforce
PotentialTarget_bruteforce_parallel = PotentialTarget_bruteforce
AccelTarget_bruteforce_parallel = AccelTarget_bruteforce
TidalTensorTarget_bruteforce_parallel = TidalTensorTarget_bruteforce
Potential_bruteforce_symmetric = Potential_bruteforce
Accel_bruteforce_symmetric = Accel_bruteforce
