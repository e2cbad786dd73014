"""Oracle softening kernels (kernel.py:4-75) against closed forms and identities.

Pins: ForceKernel(0,h)=32/(3h^3) [10.666.../h^3], PotentialKernel(0,h)=-2.8/h
(tests/test_cuda_gravity.py:122), TidalKernel(0,h)=384/(5h^5) (kernel.py:33), TidalKernel(1,1)=3 and
continuity at q=0.5, 1 (tests/test_tidal_tensor.py:243-253), TidalKernel = -(1/r) dForceKernel/dr
(:256-262), and the enclosed-mass relation K = M(<r)/r^3 of the M4 spline against quadrature of
its density.
"""

import numpy as np
import pytest

import oracle


@pytest.mark.parametrize("h", [0.1, 1.0, 3.0])
def test_values_at_origin(h):
    assert oracle.potential_kernel(0.0, h) == pytest.approx(-2.8 / h, rel=1e-15)
    assert oracle.force_kernel(0.0, h) == pytest.approx(32.0 / (3.0 * h ** 3), rel=1e-15)
    assert oracle.tidal_kernel(0.0, h) == pytest.approx(384.0 / (5.0 * h ** 5), rel=1e-15)


def test_tidal_kernel_at_support_radius():
    assert oracle.tidal_kernel(1.0, 1.0) == pytest.approx(3.0, rel=1e-9)


@pytest.mark.parametrize("fn", ["force_kernel", "potential_kernel", "tidal_kernel"])
@pytest.mark.parametrize("q", [0.5, 1.0])
def test_continuity(fn, q):
    f = getattr(oracle, fn)
    h = 0.7
    eps = 1e-9
    a, b = f(q * h * (1 - eps), h), f(q * h * (1 + eps), h)
    assert abs(a - b) <= 1e-6 * max(abs(a), abs(b))


def test_point_mass_outside_support():
    for r in (1.01, 2.0, 7.5):
        assert oracle.force_kernel(r, 1.0) == pytest.approx(r ** -3, rel=1e-15)
        assert oracle.potential_kernel(r, 1.0) == pytest.approx(-1.0 / r, rel=1e-15)
        assert oracle.tidal_kernel(r, 1.0) == pytest.approx(3.0 * r ** -5, rel=1e-15)
    assert oracle.potential_kernel(0.3, 0.0) == pytest.approx(-1.0 / 0.3, rel=1e-15)


def test_tidal_is_minus_dK_dr_over_r():
    h = 1.3
    for r in np.linspace(0.05, 1.25, 25):
        d = 1e-6 * h
        dK = (oracle.force_kernel(r + d, h) - oracle.force_kernel(r - d, h)) / (2 * d)
        assert oracle.tidal_kernel(r, h) == pytest.approx(-dK / r, rel=2e-6)


def test_force_kernel_is_enclosed_spline_mass():
    """K(r) r^3 = M(<r) for the unit-mass M4 spline of support h (density 8/(pi h^3) w(q))."""
    h = 1.0
    rr = np.linspace(0, 1, 200001)
    q = rr / h
    w = np.where(q <= 0.5, 1 - 6 * q ** 2 + 6 * q ** 3, 2 * (1 - q) ** 3)
    rho = 8 / (np.pi * h ** 3) * w
    integrand = 4 * np.pi * rr ** 2 * rho
    cum = np.concatenate([[0], np.cumsum(0.5 * (integrand[1:] + integrand[:-1]) * np.diff(rr))])
    for r in (0.2, 0.45, 0.5, 0.75, 0.99):
        i = int(round(r / (rr[1] - rr[0])))
        assert oracle.force_kernel(rr[i], h) * rr[i] ** 3 == pytest.approx(cum[i], rel=1e-6)


def test_potential_kernel_is_integral_of_force():
    """phi(r) = -int_r^inf K(s) s ds, i.e. dphi/dr = K r."""
    h = 0.8
    for r in np.linspace(0.05, 0.79, 15):
        d = 1e-6
        dphi = (oracle.potential_kernel(r + d, h) - oracle.potential_kernel(r - d, h)) / (2 * d)
        assert dphi == pytest.approx(oracle.force_kernel(r, h) * r, rel=1e-6)
