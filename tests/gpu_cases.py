"""Seeded particle sets shared by the GPU parity tests (inputs of the reference's own tests where
they exist: tests/test_radix_build.py, tests/test_nonunique_positions.py)."""

import numpy as np

from pytreegrav_amd.sampling import plummer, plummer_truncated


def case(name):
    if name == "n1":
        return np.array([[0.2, 0.3, 0.4]]), np.array([2.0]), np.array([0.1])
    if name in ("n2", "n3"):
        k = int(name[1])
        rng = np.random.default_rng(13)
        return rng.random((k, 3)), np.full(k, 1.0 / k), np.full(k, 0.1)
    if name == "uniform_1e4":
        rng = np.random.default_rng(42)
        return rng.random((10000, 3)), np.full(10000, 1e-4), np.full(10000, 0.01)
    if name == "plummer_3e4":
        x, m = plummer(30000, seed=42, legacy=True)
        return x, m, np.zeros_like(m)
    if name == "plummer_1e5_h":
        x, m = plummer_truncated(100000, seed=3)
        return x, m, np.full_like(m, 0.01)
    if name == "plummer_1e6":
        x, m = plummer_truncated(1000000, seed=5)
        return x, m, np.full_like(m, 0.01)
    if name == "clustered":  # test_radix_build.py:53-66
        np.random.seed(5)
        x = np.vstack([np.random.normal(size=(5000, 3)) * 1e-3, np.random.rand(5000, 3)])
        return x, np.ones(len(x)) / len(x), np.repeat(1e-4, len(x))
    if name == "rekey":  # test_radix_build.py:82-98
        np.random.seed(7)
        x = np.vstack([[[0.0, 0, 0], [1, 1, 1]], 0.5 + 1e-8 * np.random.rand(300, 3)])
        return x, np.ones(len(x)) / len(x), np.repeat(1e-3, len(x))
    if name == "dups":  # test_radix_build.py:69-79
        np.random.seed(11)
        x = np.random.rand(1000, 3)
        x = np.vstack([x, x[:50]])
        return x, np.ones(len(x)) / len(x), np.repeat(0.02, len(x))
    if name == "dups_origin":  # test_nonunique_positions.py:189-196
        x = np.array([[0.0, 0, 0], [0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
        return x, np.repeat(0.2, 5), np.repeat(0.1, 5)
    if name == "bucket_chain":
        x = np.vstack([np.full((20, 3), 0.25), np.eye(3)])
        return x, np.linspace(0.5, 1.5, len(x)), np.full(len(x), 0.05)
    if name == "dup40_6e4":  # test_radix_build.py:238-250 (dup=40)
        rng = np.random.default_rng(5)
        x = rng.random((60000, 3))
        x[:40] = x[0]
        return x, rng.random(60000) / 60000, rng.random(60000) * 0.01
    if name == "lattice":  # test_nonunique_positions.py:228-238
        g = np.arange(20) * 1.0
        x = np.ascontiguousarray(np.stack(np.meshgrid(g, g, g), -1).reshape(-1, 3))
        return x, np.repeat(1.0 / len(x), len(x)), np.repeat(0.1, len(x))
    if name == "all_identical":
        x = np.full((100, 3), 0.5)
        return x, np.full(100, 0.01), np.full(100, 0.1)
    if name == "zero_mass":
        rng = np.random.default_rng(9)
        x = rng.random((3000, 3))
        m = rng.random(3000)
        m[:1500] = 0.0
        return x, m, rng.random(3000) * 0.02
    raise KeyError(name)


BUILD_CASES = ["n1", "n2", "n3", "uniform_1e4", "plummer_3e4", "plummer_1e5_h", "clustered", "rekey", "dups",
               "dups_origin", "bucket_chain", "dup40_6e4", "lattice", "all_identical", "zero_mass"]
