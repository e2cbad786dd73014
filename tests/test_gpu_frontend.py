"""The drop-in API contract on the GPU, modelled on the reference's own tests:
tests/test_radix_build.py:101-147 (N = 1, 2, 3, lists), test_tree_reuse_and_ordering.py (supplied trees,
pre-sorted and permuted targets, idempotent Morton order), test_tree_quadrupole_mismatch.py,
test_nonunique_positions.py (warnings, the one-self-pair potential offset), test_field.py (Field equals
the functional API), test_tidal_tensor.py (point mass, tracelessness, trace = -4 pi G rho) and
test_plummer.py / tree_test.py (accuracy against brute force)."""

import warnings

import numpy as np
import pytest

import oracle
from gpu_cases import case

pytestmark = pytest.mark.gpu


def cloud(n, seed=42):
    rng = np.random.default_rng(seed)
    x = np.ascontiguousarray(rng.normal(size=(n, 3)))
    return x, np.ones(n) / n, np.full(n, 2.0 * n ** (-1 / 3))


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("n", [1, 2, 3])
@pytest.mark.parametrize("method", ["tree", "bruteforce"])
def test_small_n_and_lists(n, method):
    from pytreegrav_amd import Accel, Potential

    rng = np.random.default_rng(n)
    x = rng.random((n, 3))
    m = np.repeat(1.0 / n, n)
    h = np.repeat(0.1, n)
    phi = Potential(x, m, h, method=method)
    acc = Accel(x, m, h, method=method)
    assert phi.shape == (n,) and acc.shape == (n, 3)
    assert np.all(np.isfinite(phi)) and np.all(np.isfinite(acc))
    if n == 1:
        assert phi[0] == 0.0 and np.array_equal(acc[0], np.zeros(3))
    assert np.allclose(Potential(x.tolist(), m.tolist(), h.tolist(), method=method), phi)
    ref = oracle.fields_self_bruteforce(x, m, h, potential=True, accel=True)
    if n > 1:
        assert np.allclose(phi, ref["potential"], rtol=1e-12)
        assert np.allclose(acc, ref["accel"], rtol=1e-12, atol=1e-15)


def test_single_particle_tree():
    from pytreegrav_amd import ConstructTree

    tree = ConstructTree(np.array([[0.2, 0.3, 0.4]]), np.array([2.0]), np.array([0.1]))
    assert tree.NumParticles == 1 and np.isclose(tree.Masses[tree.NumParticles], 2.0)
    collapsed = ConstructTree(np.array([[0.2, 0.3, 0.4]]), np.float64(2.0), np.float64(0.1))
    assert collapsed.NumParticles == 1


def test_morton_order_is_idempotent():
    from pytreegrav_amd import ConstructTree, _morton_order

    x, m, h = cloud(20000)
    tree = ConstructTree(x, m, h)
    xs = np.ascontiguousarray(np.take(x, tree.TreewalkIndices, axis=0))
    assert np.array_equal(_morton_order(xs).cpu().numpy(), np.arange(len(xs)))
    assert np.array_equal(_morton_order(x).cpu().numpy(), oracle.morton_order(x))


@pytest.mark.parametrize("fname", ["Potential", "Accel"])
def test_supplied_tree_and_target_order(fname):
    import pytreegrav_amd as ptg

    fn = getattr(ptg, fname)
    x, m, h = cloud(20000)
    tree = ptg.ConstructTree(x, m, h)
    ref = fn(x, m, h, parallel=True)
    assert rel(fn(x, m, h, tree=tree, parallel=True), ref) < 1e-14
    idx = tree.TreewalkIndices
    xs, hs = np.ascontiguousarray(x[idx]), np.ascontiguousarray(h[idx])
    assert rel(fn(xs, m, hs, tree=tree), np.take(ref, idx, axis=0)) < 1e-14
    perm = np.random.default_rng(0).permutation(len(x))
    assert rel(fn(np.ascontiguousarray(x[perm]), m, np.ascontiguousarray(h[perm]), tree=tree),
               np.take(ref, perm, axis=0)) < 1e-14


@pytest.mark.parametrize("fname", ["Potential", "Accel"])
def test_supplied_tree_larger_than_targets(fname):
    import pytreegrav_amd as ptg

    fn = getattr(ptg, fname)
    x, m, h = cloud(5000)
    tree = ptg.ConstructTree(x, m, h)
    sub = np.ascontiguousarray(x[:200])
    got = fn(sub, m[:200], np.ascontiguousarray(h[:200]), tree=tree)
    assert len(got) == 200 and np.all(np.isfinite(got))
    sub_ref = fn(sub, m[:200], np.ascontiguousarray(h[:200]), tree=tree, group_size=1)
    full_ref = fn(x, m, h, tree=tree, group_size=1)
    assert rel(sub_ref, full_ref[:200]) < 1e-14


def test_quadrupole_mismatch_raises_and_match_works():
    from pytreegrav_amd import Accel, AccelTarget, ConstructTree, Potential, PotentialTarget

    rng = np.random.default_rng(7)
    x = rng.normal(size=(500, 3))
    m = rng.random(500) / 500
    h = np.repeat(0.05, 500)
    tree = ConstructTree(x, m, h)
    assert not tree.HasQuads
    for call in (lambda: Potential(x, m, h, tree=tree, quadrupole=True, method="tree"),
                 lambda: Accel(x, m, h, tree=tree, quadrupole=True, method="tree"),
                 lambda: PotentialTarget(x[:50], None, None, h[:50], tree=tree, quadrupole=True, method="tree"),
                 lambda: AccelTarget(x[:50], None, None, h[:50], tree=tree, quadrupole=True, method="tree")):
        with pytest.raises(ValueError, match="quadrupole"):
            call()
    qtree = ConstructTree(x, m, h, quadrupole=True)
    assert np.all(np.isfinite(Accel(x, m, h, tree=qtree, quadrupole=True, method="tree")))
    assert np.all(np.isfinite(PotentialTarget(x[:50], None, None, h[:50], tree=qtree, quadrupole=True,
                                              method="tree")))


def duplicated(soft):
    rng = np.random.default_rng(3)
    base = rng.random((2000, 3))
    x = np.vstack([base, base[:200]])
    is_dup = np.zeros(len(x), bool)
    is_dup[:200] = True
    is_dup[2000:] = True
    return x, np.repeat(1.0 / len(x), len(x)), np.repeat(soft, len(x)), is_dup


def test_duplicates_warn_and_potential_offset_is_one_self_pair():
    from pytreegrav_amd import Accel, ConstructTree, Potential

    x, m, h, is_dup = duplicated(0.02)
    with pytest.warns(UserWarning, match="non-unique") as rec:
        ConstructTree(x, m, h)
    assert any("Softening will" in str(w.message) for w in rec)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        a_tree = Accel(x, m, h, method="tree")
        a_bf = Accel(x, m, h, method="bruteforce")
        d = Potential(x, m, h, method="tree") - Potential(x, m, h, method="bruteforce")
    assert np.sqrt(np.mean(np.sum((a_tree - a_bf) ** 2, 1))) / np.sqrt(np.mean(np.sum(a_bf ** 2, 1))) < 0.02
    expected = m[0] * abs(oracle.potential_kernel(0.0, 0.02))
    assert np.mean(d[is_dup]) == pytest.approx(expected, rel=0.05)
    assert abs(np.mean(d[~is_dup])) < 0.02 * expected
    x0, m0, h0, _ = duplicated(0.0)
    with pytest.warns(UserWarning, match="non-unique") as rec:
        ConstructTree(x0, m0, h0)
    assert any("singular" in str(w.message) or "garbage" in str(w.message) for w in rec)


def test_lattice_does_not_warn():
    from pytreegrav_amd import ConstructTree

    g = np.arange(20) * 1.0
    x = np.ascontiguousarray(np.stack(np.meshgrid(g, g, g), -1).reshape(-1, 3))
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        tree = ConstructTree(x, np.repeat(1.0 / len(x), len(x)), np.repeat(0.1, len(x)))
    assert not tree.HasCoincidentPoints


@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_field_matches_functional_api(precision):
    from pytreegrav_amd import Accel, Field, Potential, TidalTensor

    x, m, h = cloud(6000, seed=5)
    f = Field(x, m, h, theta=0.6, precision=precision)
    fused = f.evaluate(potential=True, accel=True, tidal=True)
    # fused vs separate walks: same accepted set, sums compiled per field set (the reference allows
    # 64 ulp, tests/test_field.py:46-58)
    assert rel(fused["potential"], f.potential()) < (1e-6 if precision == "float32" else 1e-13)
    assert rel(fused["accel"], f.accel()) < (1e-6 if precision == "float32" else 1e-13)
    if precision == "float64":
        assert np.array_equal(f.potential(), Potential(x, m, h, theta=0.6))
        assert np.array_equal(f.accel(), Accel(x, m, h, theta=0.6))
        assert np.array_equal(f.tidal(), TidalTensor(x, m, h, theta=0.6))
    xt = np.random.default_rng(2).normal(size=(300, 3))
    got = f.accel(pos_target=xt)
    ref = oracle.fields_target_bruteforce(xt, np.zeros(300), x, m, h, accel=True)["accel"]
    assert np.sqrt(np.mean(np.sum((got - ref) ** 2, 1))) / np.sqrt(np.mean(np.sum(ref ** 2, 1))) < 0.01


def test_tidal_tensor_contract():
    from pytreegrav_amd import TidalTensor, TidalTensorTarget

    x = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 2.0]])
    T = TidalTensor(x, np.array([1.0, 0.0]), method="bruteforce")[1]
    assert np.allclose(T, np.diag([-0.125, -0.125, 0.25]), rtol=1e-12, atol=0)
    rng = np.random.default_rng(1)
    xc, mc = rng.normal(size=(400, 3)), rng.random(400) / 400
    for method, quad in (("bruteforce", False), ("tree", False), ("tree", True)):
        Tt = TidalTensor(xc, mc, np.zeros(400), method=method, theta=0.7, quadrupole=quad)
        assert np.abs(np.trace(Tt, axis1=1, axis2=2)).max() / np.abs(Tt).max() < 1e-13
    tg = rng.normal(size=(12, 3)) * 1.5
    Tg = TidalTensorTarget(tg, xc, mc, softening_source=np.full(400, 0.6), method="bruteforce")
    ref = oracle.fields_target_bruteforce(tg, np.zeros(12), xc, mc, np.full(400, 0.6), tidal=True)["tidal"]
    assert rel(Tg, ref) < 1e-13


def test_plummer_accuracy_both_precisions():
    from pytreegrav_amd import Accel, Potential
    from pytreegrav_amd.sampling import plummer

    pos, m = plummer(30000, seed=42, legacy=True)
    h = np.zeros_like(m)
    a_bf = Accel(pos, m, h, method="bruteforce")
    p_bf = Potential(pos, m, h, method="bruteforce")
    for dev in ("cpu", "cuda"):
        a = Accel(pos, m, h, method="tree", device=dev)
        p = Potential(pos, m, h, method="tree", device=dev)
        assert np.sqrt(np.mean(np.sum((a - a_bf) ** 2, 1))) / np.sqrt(np.mean(np.sum(a_bf ** 2, 1))) < 0.02
        assert np.std(p - p_bf) / np.std(p_bf) < 0.02


def test_torch_inputs_stay_on_device():
    import torch

    from pytreegrav_amd import Accel, AccelTarget

    x, m, h = cloud(3000)
    dx = torch.from_numpy(x).cuda()
    out = Accel(dx, torch.from_numpy(m).cuda(), torch.from_numpy(h).cuda(), method="tree")
    assert isinstance(out, torch.Tensor) and out.is_cuda
    assert rel(out.cpu().numpy(), Accel(x, m, h, method="tree")) < 1e-14
    g = AccelTarget(dx[:100], dx, torch.from_numpy(m).cuda(), method="tree", precision="float32")
    assert isinstance(g, torch.Tensor) and g.is_cuda and g.shape == (100, 3)


def test_adaptive_dispatch_thresholds():
    """frontend.py:210-220, 364-368: float64 walks the tree above 1000 (serial) / 4000 (parallel)
    particles or with a supplied tree, as the reference does; float32 above the crossover measured on
    the GPU (FP32_TREE_MIN_N)."""
    from pytreegrav_amd import Accel, Potential, PotentialTarget, frontend

    x, m, h = cloud(800)
    assert np.array_equal(Potential(x, m, h), Potential(x, m, h, method="bruteforce"))
    x, m, h = cloud(2000)
    assert np.array_equal(Accel(x, m, h), Accel(x, m, h, method="tree"))  # float64: the reference's 1000
    assert np.array_equal(Accel(x, m, h, parallel=True), Accel(x, m, h, method="bruteforce"))  # its 4000
    assert 2000 < frontend.FP32_TREE_MIN_N
    assert np.array_equal(Accel(x, m, h, device="cuda"), Accel(x, m, h, device="cuda", method="bruteforce"))
    x, m, h = cloud(frontend.FP32_TREE_MIN_N + 1)
    assert np.array_equal(Accel(x, m, h, device="cuda"), Accel(x, m, h, device="cuda", method="tree"))
    x, m, h = cloud(800)
    xt = x[:100]
    assert np.array_equal(PotentialTarget(xt, x, m, h[:100], h), PotentialTarget(xt, x, m, h[:100], h,
                                                                                  method="bruteforce"))


def test_non_finite_device_input_aborts_the_build(capsys):
    import torch

    from pytreegrav_amd import ConstructTree

    x, m, h = cloud(1000)
    for bad in ("pos", "m", "h"):
        X, M, H = (torch.from_numpy(a.copy()).cuda() for a in (x, m, h))
        {"pos": X, "m": M, "h": H}[bad].view(-1)[7] = float("inf") if bad == "m" else float("nan")
        with pytest.raises(RuntimeError):
            ConstructTree(X, M, H)
        assert "Invalid input detected" in capsys.readouterr().out


@pytest.mark.parametrize("quadrupole", [False, True])
def test_grouped_target_api_matches_oracle(quadrupole):
    """FieldsTarget_grouped and the three *Target_grouped wrappers (grouped_treewalk.py:285-397) on
    Morton-sorted external targets: equal to the oracle's grouped walk to 1e-10 (same groups of
    group_size consecutive targets), for every field subset, with G and theta passed through."""
    from pytreegrav_amd import (AccelTarget_grouped, ConstructTree, FieldsTarget_grouped, PotentialTarget_grouped,
                                TidalTensorTarget_grouped)

    x, m, h = case("plummer_1e5_h")
    rng = np.random.default_rng(77)
    xt = rng.normal(size=(5000, 3)) * 2.0
    ht = rng.random(5000) * 0.02
    order = oracle.morton_order(xt)
    xs, hs = np.ascontiguousarray(xt[order]), np.ascontiguousarray(ht[order])
    tree = ConstructTree(x, m, h, quadrupole=quadrupole)
    ot = oracle.build_tree(x, m, h, quadrupole=quadrupole)
    for gs, theta, G in ((8, 0.7, 1.0), (32, 0.5, 2.5), (1, 1.0, 1.0)):
        print(f"[grouped api] quad={quadrupole} gs={gs} theta={theta}", flush=True)
        ref = oracle.fields_grouped(ot, xs, hs, potential=True, accel=True, tidal=True, quadrupole=quadrupole,
                                    group_size=gs, theta=theta, G=G)
        got = FieldsTarget_grouped(xs, hs, tree, potential=True, accel=True, tidal=True, group_size=gs, G=G,
                                   theta=theta, quadrupole=quadrupole)
        assert set(got) == {"potential", "accel", "tidal"}
        for k in got:
            assert np.abs(got[k] - ref[k]).max() < 1e-10 * np.abs(ref[k]).max(), k
        a = AccelTarget_grouped(xs, hs, tree, group_size=gs, G=G, theta=theta, quadrupole=quadrupole)
        p = PotentialTarget_grouped(xs, hs, tree, group_size=gs, G=G, theta=theta, quadrupole=quadrupole)
        t = TidalTensorTarget_grouped(xs, hs, tree, group_size=gs, G=G, theta=theta, quadrupole=quadrupole)
        assert a.shape == (5000, 3) and p.shape == (5000,) and t.shape == (5000, 3, 3)
        assert np.abs(a - ref["accel"]).max() < 1e-10 * np.abs(ref["accel"]).max()
        assert np.abs(p - ref["potential"]).max() < 1e-10 * np.abs(ref["potential"]).max()
        assert np.abs(t - ref["tidal"]).max() < 1e-10 * np.abs(ref["tidal"]).max()
    with pytest.raises(ValueError):
        FieldsTarget_grouped(xs, hs, tree)  # no field requested
