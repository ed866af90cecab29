"""The coarse-to-fine solve's step-size rule on the CPU: the stability bound's pieces
against the stencils' own symbols and the RK3 stability polynomial, and the levels it
picks."""
import math

import numpy as np
import pytest

from trapbec_amd import initial
from trapbec_amd import warmstart as ws

# the reference's 6th-order stencils (src/deriv.f90:17,24), in units of 1/h^2 and 1/h
D2 = np.array([2.0, -27.0, 270.0, -490.0, 270.0, -27.0, 2.0]) / 180.0
D1 = np.array([-1.0, 9.0, -45.0, 0.0, 45.0, -9.0, 1.0]) / 60.0
THETA = np.linspace(0.0, math.pi, 200001)


def symbol(c):
    k = np.arange(-3, 4)
    return np.exp(1j * np.outer(THETA, k)) @ c


def rk3(z):
    return 1 + z + z * z / 2 + z ** 3 / 6


def test_rk3_real_limit_is_where_the_polynomial_leaves_the_unit_disc():
    assert abs(abs(rk3(-ws.RK3_REAL_LIMIT)) - 1.0) < 1e-3
    assert abs(rk3(-2.50)) < 1.0 < abs(rk3(-2.53))
    assert all(abs(rk3(z)) <= 1.0 for z in np.linspace(-2.5, 0.0, 1001))


def test_kinetic_bound_is_half_the_laplacian_symbol_at_nyquist():
    """Two grids of one non-rotating problem differ in max_rate by the -1/2 Laplacian
    bound alone: 1/2 max|D2(theta)| (= 1088/360 at theta = pi) times sum 1/h^2."""
    d2max = float(np.abs(symbol(D2)).max())
    assert d2max == pytest.approx(1088.0 / 180.0, rel=1e-12)
    pr = initial.Problem((40, 48, 56), (10.0, 12.0, 14.0), (1.0, 0.9, 1.3), 500.0, 0.0, -1e-3j)
    a, b = (40, 48, 56), (20, 30, 28)
    diff = ws.max_rate(pr, a) - ws.max_rate(pr, b)
    inv_h2 = [(n / L) ** 2 - (m / L) ** 2 for n, m, L in zip(a, b, pr.box)]
    assert diff == pytest.approx(0.5 * d2max * sum(inv_h2), rel=1e-12)


def test_frame_term_bound_covers_the_first_derivative_symbol():
    """The rotation part is |Omega| r_max (bound of |D1|) / h; the bound used (110/60,
    the sum of |coefficients|) is at least the symbol's maximum (1.586)."""
    d1max = float(np.abs(symbol(D1)).max())
    assert 1.58 < d1max <= float(np.abs(D1).sum()) == pytest.approx(110.0 / 60.0)
    p0 = initial.Problem((32, 32, 32), (8.0, 8.0, 8.0), (1.0, 1.0, 1.0), 100.0, 0.0, -1e-3j)
    p1 = initial.Problem((32, 32, 32), (8.0, 8.0, 8.0), (1.0, 1.0, 1.0), 100.0, 0.6, -1e-3j)
    rot = ws.max_rate(p1) - ws.max_rate(p0)
    assert rot == pytest.approx(0.6 * math.hypot(4.0, 4.0) * (110.0 / 60.0) / 0.25, rel=1e-12)
    assert rot >= 0.6 * math.hypot(4.0, 4.0) * d1max / 0.25


def test_trap_and_interaction_terms():
    """At one grid, max_rate minus the kinetic bound is the trap at the box corner plus
    2 mu_TF, mu_TF = (1/2)(15 g w_x w_y w_z / 4 pi)^(2/5) for norm 1."""
    pr = initial.Problem((24, 24, 24), (6.0, 8.0, 10.0), (1.0, 0.5, 2.0), 300.0, 0.0, -1e-3j)
    kin = sum(1088.0 / 360.0 * (n / L) ** 2 for n, L in zip(pr.shape, pr.box))
    v = 0.5 * ((1.0 * 3.0) ** 2 + (0.5 * 4.0) ** 2 + (2.0 * 5.0) ** 2)
    mu = 0.5 * (15.0 * 300.0 * 1.0 / (4.0 * math.pi)) ** 0.4
    assert ws.max_rate(pr) - kin == pytest.approx(v + 2.0 * mu, rel=1e-12)


def test_levels_at_the_c5_problem():
    pr = ws.fine_problem(np.array([1.0, 0.8, 3000.0, 0.5, 0.0]), 384, 8e-4)
    assert ws.max_rate(pr) * 8e-4 < ws.RK3_REAL_LIMIT     # the configured dtau is stable
    levels = ws.continuation_levels(pr)
    assert [s for s, _ in levels] == [(128, 128, 128), (192, 192, 192)]
    for shape, dtau in levels:
        assert dtau * ws.max_rate(pr, shape) == pytest.approx(0.95 * ws.RK3_REAL_LIMIT)
        assert dtau > 8e-4
    # never below the fine dtau
    slow = ws.fine_problem(np.array([1.0, 0.8, 3000.0, 0.5, 0.0]), 384, 0.5)
    assert all(d == 0.5 for _, d in ws.continuation_levels(slow))
