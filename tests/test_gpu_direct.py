"""LDS-tiled direct sum against the oracle restatement of bruteforce.py: FP64 to 1e-13, FP32 to
the sqrt(N) accumulation bound the reference uses for its own device brute force
(tests/test_cuda_gravity.py:275-288); ragged tiles (N not a multiple of 256) included."""

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _cloud(n, soft, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 3))
    m = rng.random(n) / n
    h = {"zero": np.zeros(n), "uniform": np.full(n, 0.05), "varying": rng.random(n) * 0.2}[soft]
    return x, m, h


@pytest.mark.parametrize("n", [1000, 4097])
@pytest.mark.parametrize("soft", ["zero", "uniform", "varying"])
def test_direct_self_matches_oracle(n, soft):
    from pytreegrav_amd import engine

    x, m, h = _cloud(n, soft)
    print(f"[direct] n={n} soft={soft}", flush=True)
    ref = oracle.fields_self_bruteforce(x, m, h, potential=True, accel=True, tidal=True)
    g64 = engine.unpack_fields(engine.direct(x, h, x, m, h, 7, self_mode=True).cpu(), 7)
    g32 = engine.unpack_fields(engine.direct(x, h, x, m, h, 7, self_mode=True, precision="float32").cpu(), 7)
    tol32 = 1e-5 * np.sqrt(n / 1000)
    for k in ref:
        r = ref[k]
        assert np.abs(g64[k].numpy() - r).max() / np.abs(r).max() < 1e-13, k
        assert np.abs(g32[k].numpy() - r).max() / np.abs(r).max() < 10 * tol32, k


def test_direct_target_form_skips_coincident():
    from pytreegrav_amd import engine

    x, m, h = _cloud(2000, "uniform", seed=4)
    xt = np.vstack([x[:100], np.random.default_rng(5).normal(size=(57, 3))])
    ht = np.full(len(xt), 0.01)
    ref = oracle.fields_target_bruteforce(xt, ht, x, m, h, potential=True, accel=True)
    got = engine.unpack_fields(engine.direct(xt, ht, x, m, h, 3, self_mode=False).cpu(), 3)
    for k in ref:
        assert np.abs(got[k].numpy() - ref[k]).max() / np.abs(ref[k]).max() < 1e-13


def test_frontend_bruteforce_methods():
    from pytreegrav_amd import Accel, Potential, PotentialTarget

    x, m, h = _cloud(700, "varying", seed=8)
    ref = oracle.fields_self_bruteforce(x, m, h, potential=True, accel=True)
    assert np.allclose(Potential(x, m, h, method="bruteforce"), ref["potential"], rtol=1e-13, atol=0)
    assert np.allclose(Accel(x, m, h, method="bruteforce"), ref["accel"], rtol=1e-12, atol=1e-15)
    # PotentialTarget(x, x) is the self potential (tests/test_target_self_potential.py:44-51)
    assert np.allclose(PotentialTarget(x, x, m, h, h, method="bruteforce"), ref["potential"], rtol=1e-12)
