"""Frontend error contract checked before any device work (frontend.py:28-50, 222-225, 355-356,
123-125), and the rule that nothing computes without the GPU library and a GPU."""

import numpy as np
import pytest

import pytreegrav_amd as ptg


def _x(n=10):
    rng = np.random.default_rng(0)
    return rng.random((n, 3)), np.full(n, 1.0 / n), np.full(n, 0.01)


@pytest.mark.parametrize("fn", [ptg.Potential, ptg.Accel, ptg.TidalTensor])
def test_method_validation(fn):
    x, m, h = _x()
    with pytest.raises(TypeError, match="Invalid method type"):
        fn(x, m, h, method=3)
    with pytest.raises(ValueError, match="Invalid method"):
        fn(x, m, h, method="fmm")


@pytest.mark.parametrize("fn", [ptg.Potential, ptg.Accel])
def test_device_validation(fn):
    x, m, h = _x()
    with pytest.raises(ValueError, match="must be 'cpu' or 'cuda'"):
        fn(x, m, h, device="gpu")


@pytest.mark.parametrize("fn", [ptg.PotentialTarget, ptg.AccelTarget, ptg.TidalTensorTarget])
def test_target_forms_need_sources_or_tree(fn):
    x, m, h = _x()
    with pytest.raises(ValueError, match="Must pass either pos_source & m_source or source tree."):
        fn(x, None, None)


def test_non_finite_input_aborts_the_build(capsys):
    x, m, h = _x()
    x[3, 1] = np.nan
    with pytest.raises(RuntimeError):
        ptg.ConstructTree(x, m, h)
    assert "Invalid input detected" in capsys.readouterr().out


def test_precision_validation():
    x, m, h = _x()
    with pytest.raises(ValueError, match="precision"):
        ptg.AccelTarget(x, x, m, precision="float16")


def test_group_size_validation():
    x, m, h = _x()
    with pytest.raises(ValueError, match="group_size"):
        ptg.Accel(x, m, h, group_size=1000)


def test_velocity_tree_is_not_provided():
    x, m, h = _x()
    with pytest.raises(NotImplementedError):
        ptg.ConstructTree(x, m, h, vel=np.zeros_like(x))


def test_quadrupole_guard_message():
    class T:
        HasQuads = False

    with pytest.raises(ValueError, match="quadrupole"):
        ptg.checkTreeQuadrupoles(T(), True)
    ptg.checkTreeQuadrupoles(T(), False)


def test_no_cpu_fallback():
    """Without a GPU the compute path raises instead of silently computing on the host."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    x, m, h = _x(50)
    with pytest.raises(RuntimeError, match="GPU"):
        ptg.Accel(x, m, h, method="bruteforce")
    with pytest.raises(RuntimeError, match="GPU"):
        ptg.ConstructTree(x, m, h)


def test_coincident_warning_messages():
    class T:
        HasCoincidentPoints = True

    with pytest.warns(UserWarning, match="Softening will"):
        ptg.warn_if_coincident_positions(T(), np.full(3, 0.1))
    with pytest.warns(UserWarning, match="singular or garbage"):
        ptg.warn_if_coincident_positions(T(), np.zeros(3))


@pytest.mark.parametrize("fn", [ptg.Potential, ptg.Accel])
def test_criterion_validation(fn):
    x, m, h = _x()
    with pytest.raises(ValueError, match="criterion must be"):
        fn(x, m, h, device="cuda", criterion="fmm", alpha=1e-3)
    with pytest.raises(ValueError, match="float32 walk"):
        fn(x, m, h, device="cpu", criterion="learned", alpha=1e-3)
    with pytest.raises(ValueError, match="needs alpha"):
        fn(x, m, h, device="cuda", criterion="error")


def test_group_size_validation():
    x, m, h = _x()
    with pytest.raises(ValueError, match="group_size must be"):
        ptg.Accel(x, m, h, device="cuda", group_size="fast")
    with pytest.raises(ValueError, match="group_size must be between"):
        ptg.Accel(x, m, h, device="cuda", group_size=512, method="tree")


def test_forest_offsets_validation():
    x, m, h = _x(10)
    for off in ([0, 1, 10], [0, 5, 9], [1, 5, 10], [0, 10, 5, 10]):
        with pytest.raises(ValueError, match="offsets must rise"):
            ptg.Forest(x, m, h, off)


def test_shipped_opening_model_layout():
    import json
    import os

    from pytreegrav_amd import learned

    with open(learned.DEFAULT_MODEL) as f:
        d = json.load(f)
    assert len(d["params"]) == 1441 and all(np.isfinite(d["params"]))
    assert 0.0 <= d["margin"] < 0.5
    assert d["meta"]["features"] == learned.FEATURES
    assert os.path.getsize(learned.DEFAULT_MODEL) < 64 * 1024
