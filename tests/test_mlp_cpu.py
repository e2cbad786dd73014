"""The numpy MLP restatement used to check the HIP training kernels is itself pinned here: its
backpropagated gradient against central finite differences of its loss (float64), including the
asymmetric weight on under-predictions, and the parameter count/layout of include/ptg.h."""

import re
from pathlib import Path

import numpy as np

import mlp_ref

ROOT = Path(__file__).resolve().parents[1]


def test_parameter_count_matches_header():
    h = (ROOT / "include" / "ptg.h").read_text()
    assert int(re.search(r"#define PTG_MLP_NPARAM (\d+)", h).group(1)) == mlp_ref.NPARAM == 1441


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = rng.normal(scale=0.3, size=mlp_ref.NPARAM)
    x = rng.normal(size=(40, 10))
    y = rng.normal(size=40)
    for under in (1.0, 4.0):
        g = mlp_ref.grad(p, x, y, under)
        idx = rng.choice(mlp_ref.NPARAM, 80, replace=False)
        idx = np.concatenate([idx, [0, 319, 320, 351, 352, 1375, 1376, 1407, 1408, 1439, 1440]])
        for k in idx:
            d = np.zeros_like(p)
            d[k] = 1e-6
            fd = (mlp_ref.loss(p + d, x, y, under) - mlp_ref.loss(p - d, x, y, under)) / 2e-6
            assert abs(fd - g[k]) <= 1e-7 + 1e-5 * abs(fd), (under, k, fd, g[k])


def test_init_layout():
    from pytreegrav_amd.learned import init_params

    p = init_params(0)
    W1, b1, W2, b2, W3, b3 = mlp_ref.unpack(p)
    assert np.all(b1 == 0) and np.all(b2 == 0) and b3 == 0
    assert np.abs(W1).max() <= np.sqrt(6 / 42) and np.abs(W2).max() <= np.sqrt(6 / 64)
    assert np.abs(W3).max() <= np.sqrt(6 / 33) and np.all(W3 != 0)
