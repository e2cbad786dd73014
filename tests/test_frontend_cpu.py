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
