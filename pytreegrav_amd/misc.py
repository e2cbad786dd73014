"""Host helpers of the reference's misc module (src/pytreegrav/misc.py)."""

import numpy as np


def random_rotation(seed):
    """Reproducible random rotation matrix for an integer seed (misc.py:6-54): the x axis uniform on
    the sphere, a second uniform direction made orthonormal to it (Gram-Schmidt) as the y axis, and
    z = x cross y.  Drawn from a Mersenne-Twister stream seeded with ``seed`` (uniform(-1,1) for the
    cosine, uniform(0,1) for the azimuth, twice), without touching numpy's global random state."""
    rs = np.random.RandomState(seed)
    R = np.zeros((3, 3))
    for row in range(2):
        cos_t = rs.uniform(-1, 1)
        sin_t = np.sqrt(max(1 - cos_t * cos_t, 0))
        phi = 2 * np.pi * rs.uniform(0, 1)
        R[row] = sin_t * np.cos(phi), sin_t * np.sin(phi), cos_t
    R[1] -= np.dot(R[0], R[1]) * R[0]
    R[1] /= np.sqrt(np.dot(R[1], R[1]))
    R[2] = np.cross(R[0], R[1])
    return R
