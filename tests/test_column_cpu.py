"""Column density oracle (CPU restatement of treewalk.py:821-1276) pinned without the reference:
the ray-traced walk against an independent closed-form direct sum of uniform-sphere chords (the walk
is exact, so the two agree to rounding), hand-computed single-sphere values, and the binned walk
against a brute-force per-particle binning when no node can be accepted.  Plus the host-side
random rotation (misc.py:6-54): orthonormal, right-handed, reproducible."""

import numpy as np
import pytest

import oracle
from pytreegrav_amd.misc import random_rotation


def chord_direct(targets, rays, x, m, h):
    """Column from every target along every ray through uniform spheres, by direct summation:
    density 3M/(4 pi h^3); a ray from p along n passes a sphere at impact parameter b with the sphere
    centre at axial distance z; from outside the sphere (|d| > h) it crosses the whole chord
    2 sqrt(h^2 - b^2) if the sphere is ahead (z > 0), from inside only z + sqrt(h^2 - b^2)."""
    out = np.zeros((len(targets), len(rays)))
    for j, p in enumerate(targets):
        d = x - p
        r2 = np.sum(d * d, 1)
        for i, n in enumerate(rays):
            z = d @ n
            b2 = r2 - z * z
            for k in range(len(x)):
                if h[k] <= 0 or b2[k] < 0 or b2[k] >= h[k] ** 2:
                    continue
                rho = 3 * m[k] / (4 * np.pi * h[k] ** 3)
                half = np.sqrt(h[k] ** 2 - b2[k])
                if r2[k] > h[k] ** 2:
                    if z[k] > 0:
                        out[j, i] += rho * 2 * half
                else:
                    out[j, i] += rho * (z[k] + half)
    return out


def unit(v):
    v = np.asarray(v, dtype=float)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def test_single_sphere_values():
    x = np.array([[0.0, 0.0, 0.0], [50.0, 50.0, 50.0]])
    m, h = np.array([2.0, 1e-9]), np.array([1.0, 1e-6])
    t = oracle.build_tree(x, m, h)
    rho = 3 * 2.0 / (4 * np.pi)
    east = np.array([[1.0, 0.0, 0.0]])
    cases = [([-5.0, 0.0, 0.0], rho * 2), ([0.0, 0.0, 0.0], rho * 1), ([-5.0, 0.6, 0.0], rho * 2 * 0.8),
             ([5.0, 0.0, 0.0], 0.0), ([-5.0, 1.01, 0.0], 0.0), ([-0.5, 0.0, 0.0], rho * 1.5)]
    for p, want in cases:
        got = oracle.column_rays(t, [p], east)[0, 0]
        assert abs(got - want) <= 1e-12 * max(want, 1.0), (p, got, want)


def test_rays_match_direct_chords():
    rng = np.random.default_rng(4)
    n = 300
    x = rng.normal(size=(n, 3))
    m = rng.uniform(0.5, 1.5, n) / n
    h = rng.uniform(0.1, 0.5, n)
    h[:3] = 0.0  # zero radii contribute nothing
    t = oracle.build_tree(x, m, h)
    targets = np.vstack([x[:40], rng.normal(size=(10, 3)) * 2])
    rays = unit(rng.normal(size=(7, 3)))
    got = oracle.column_rays(t, targets, rays)
    want = chord_direct(targets, rays, x, m, h)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    # per-target ray grids give the same columns as the shared grid
    per = oracle.column_rays(t, targets, np.broadcast_to(rays, (len(targets), 7, 3)).copy(), per_target=True)
    assert np.array_equal(per, got)


def binned_direct(targets, x, m, h):
    out = np.zeros((len(targets), 6))
    size = 4 * np.pi / 6
    for j, p in enumerate(targets):
        for k in range(len(x)):
            d = x[k] - p
            r2 = d @ d
            a = np.abs(d)
            b = (0 if d[0] > 0 else 1) if a[0] > a[1] and a[0] > a[2] else \
                ((2 if d[1] > 0 else 3) if a[1] > a[2] else (4 if d[2] > 0 else 5))
            if r2 > h[k] ** 2:
                out[j, b] += m[k] / r2 / size
            elif h[k] > 0:
                c0 = m[k] * 3 / (4 * np.pi * h[k] ** 2)
                f = np.sqrt(r2) / h[k]
                out[j] += (1 - f) * c0
                out[j, b] += 2 * f * c0
    return out


@pytest.mark.parametrize("gs", [1, 16])
def test_binned_all_opened_equals_particle_binning(gs):
    """With theta tiny no node passes the opening test: only particles contribute, each to its sky bin
    (point mass) or spread by the overlap blend."""
    rng = np.random.default_rng(5)
    n = 200
    x = rng.normal(size=(n, 3))
    m = np.full(n, 1.0 / n)
    h = np.full(n, 0.3)
    t = oracle.build_tree(x, m, h)
    targets = x[:48]
    got = oracle.column_binned(t, targets, gs, 1e-9)
    want = binned_direct(targets, x, m, h)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_binned_far_cluster_lands_in_one_bin():
    """A compact cluster far along +y is accepted as one node: all its mass M / r^2 in bin 2."""
    rng = np.random.default_rng(6)
    x = rng.normal(size=(64, 3)) * 0.01 + [0.0, 100.0, 0.0]
    x = np.vstack([x, [[0.0, 0.0, 0.0], [0.0, 0.0, 0.1]]])
    m = np.full(len(x), 1.0)
    h = np.full(len(x), 1e-3)
    t = oracle.build_tree(x, m, h)
    got = oracle.column_binned(t, [[0.0, 0.0, -5.0]], 1, 0.7)[0]
    com = x[:64].mean(0)
    r2 = np.sum((com - [0.0, 0.0, -5.0]) ** 2)
    assert got[2] == pytest.approx(64 / r2 / (4 * np.pi / 6), rel=1e-6)


def test_binned_overlap_blend_is_pinned_by_the_uniform_sphere():
    """The binned estimator's treatment of a target inside a particle's sphere, fixed by geometry
    rather than by restating its formula: at the sphere's centre every direction sees the column
    rho * h = 3 m / (4 pi h^2) (so every sky bin gets exactly that); and at the surface the inside
    form meets the outside point-mass form m / r^2 / (4 pi / 6) in the bin facing the sphere and 0 in
    the others, so the estimate is continuous across r = h."""
    m1, h1 = 2.0, 0.5
    x = np.array([[0.0, 0.0, 0.0], [40.0, 0.0, 0.0]])  # the second particle: a far point mass in bin 0
    m = np.array([m1, 1e-30])
    h = np.array([h1, 1e-3])
    t = oracle.build_tree(x, m, h)
    centre = oracle.column_binned(t, [[0.0, 0.0, 0.0]], 1, 1e-9)[0]
    assert np.allclose(centre, 3 * m1 / (4 * np.pi * h1 ** 2), rtol=1e-12, atol=0)
    # just inside and just outside the surface, below the sphere: the sphere lies in bin 4 (+z)
    size = 4 * np.pi / 6
    for eps in (1e-9, -1e-9):
        z = -h1 * (1 + eps)
        col = oracle.column_binned(t, [[0.0, 0.0, z]], 1, 1e-9)[0]
        assert col[4] == pytest.approx(m1 / z ** 2 / size, rel=1e-6)
        assert np.abs(np.delete(col, 4)).max() < 1e-6 * col[4]


def test_random_rotation_is_a_reproducible_rotation():
    for seed in (0, 1, 17, 123456):
        R = random_rotation(seed)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-12)
        assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-12)
        assert np.array_equal(R, random_rotation(seed))
    assert not np.allclose(random_rotation(1), random_rotation(2))
    state = np.random.get_state()[1].copy()
    random_rotation(5)
    assert np.array_equal(np.random.get_state()[1], state)  # the global generator is untouched
