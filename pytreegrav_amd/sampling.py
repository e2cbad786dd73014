"""Seeded synthetic particle sets for tests and benchmarks.

``plummer`` follows the reference's inverse-CDF convention (examples/error_benchmark.py:28-35,
examples/benchmark_scaling.py:31-38): u ~ U(0,1), r = sqrt(u^(2/3)(1+u^(2/3)+u^(4/3))/(1-u^2)),
which is r = (u^(-2/3) - 1)^(-1/2), isotropic directions from normalised Gaussians, m = 1/N.
``plummer_truncated`` draws u ~ U(0, M(rmax)) with M(r) = r^3/(1+r^2)^(3/2), the exact truncation at
rmax scale radii used by the benchmark workloads.
"""

import numpy as np


def plummer(n, seed=42, legacy=False):
    """Plummer sphere (M=1, a=1) exactly as examples/error_benchmark.py:28-35 draws it.

    legacy=True reproduces tests/test_plummer.py:10-18 (np.random.seed stream) instead of
    default_rng.  Returns (pos (n,3), m (n,)).
    """
    if legacy:
        np.random.seed(seed)
        u = np.random.rand(n)
        r = np.sqrt(u ** (2.0 / 3) * (1 + u ** (2.0 / 3) + u ** (4.0 / 3)) / (1 - u ** 2))
        d = np.random.normal(size=(n, 3))
    else:
        rng = np.random.default_rng(seed)
        u = rng.random(n)
        r = np.sqrt(u ** (2.0 / 3) * (1 + u ** (2.0 / 3) + u ** (4.0 / 3)) / (1 - u ** 2))
        d = rng.normal(size=(n, 3))
    pos = (d.T * r / np.sum(d ** 2, axis=1) ** 0.5).T
    return np.ascontiguousarray(pos, dtype=np.float64), np.repeat(1.0 / n, n)


def plummer_truncated(n, seed=0, rmax=100.0, scale=1.0, mass=1.0, center=(0.0, 0.0, 0.0)):
    """Plummer sphere truncated at rmax scale radii: u ~ U(0, M(rmax)), r = (u^(-2/3)-1)^(-1/2)."""
    rng = np.random.default_rng(seed)
    mmax = rmax ** 3 / (1.0 + rmax ** 2) ** 1.5
    u = rng.random(n) * mmax
    u = np.maximum(u, np.finfo(np.float64).tiny)
    r = (u ** (-2.0 / 3.0) - 1.0) ** -0.5
    d = rng.normal(size=(n, 3))
    pos = (d.T * (scale * r) / np.sqrt(np.sum(d ** 2, axis=1))).T + np.asarray(center, np.float64)
    return np.ascontiguousarray(pos, dtype=np.float64), np.full(n, mass / n)


def uniform_cube(n, seed=0, side=1.0, center=None):
    """n points uniform in a cube of the given side ([0, side)^3 unless a centre is given)."""
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3)) * side
    if center is not None:
        pos += np.asarray(center, np.float64) - 0.5 * side
    return np.ascontiguousarray(pos), np.full(n, 1.0 / n)


def gaussian_mixture(n, seed=0, ncomp=64):
    """Clumpy set: ncomp Gaussian clumps, means U[-1,1]^3, sigma log-uniform in [0.005, 0.1],
    weights Dirichlet(1); m = 1/N."""
    rng = np.random.default_rng(seed)
    means = rng.uniform(-1.0, 1.0, size=(ncomp, 3))
    sig = np.exp(rng.uniform(np.log(0.005), np.log(0.1), size=ncomp))
    w = rng.dirichlet(np.ones(ncomp))
    comp = rng.choice(ncomp, size=n, p=w)
    pos = means[comp] + rng.normal(size=(n, 3)) * sig[comp, None]
    return np.ascontiguousarray(pos), np.full(n, 1.0 / n)


def plummer_clusters(nsys, seed=0, nmin=1000, nmax=100000, amin=0.1, amax=10.0, rmax=100.0):
    """Independent Plummer clusters: N log-uniform in [nmin, nmax], scale radius log-uniform in
    [amin, amax], softening 0.01 x scale radius.  Returns a list of (pos, m, h)."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=nsys)).astype(np.int64)
    scales = np.exp(rng.uniform(np.log(amin), np.log(amax), size=nsys))
    out = []
    for i in range(nsys):
        pos, m = plummer_truncated(int(ns[i]), seed=int(rng.integers(1 << 31)), rmax=rmax, scale=scales[i])
        out.append((pos, m, np.full(len(m), 0.01 * scales[i])))
    return out
