"""Pins of oracle functions that the weak-form and golden-vector checks of
test_oracle.py leave open.  CPU only.

  mass        closed forms: a constant state integrates to U times the
              mesh's area (shoelace over the element vertices, computed
              here); the Gaussian pulse exp(-40 |x|^2) on [-1, 1]^2
              integrates to pi/40 (the tails beyond the square are below
              e^-40).  A wrong scale factor (the orthonormal basis's
              phi_0 = sqrt(2), the reference triangle's area 1/2, J versus
              J/2) fails both.
  ic_point    the isentropic-vortex superposition of the benchmark
              configurations evaluated here from its definition: velocity
              perturbations and temperature deficits of every vortex add,
              each measured to the nearest periodic image, with
              rho = T^(1/(gamma-1)), p = rho^gamma on the unit background.
"""
import math

import numpy as np
import pytest

import dgcuda_amd as dg
import oracle as orc
from helpers import golden_mesh, mesh


def shoelace_area(m):
    V = m.arrays()[0]
    x1, y1, x2, y2, x3, y3 = V
    return float(np.sum(0.5 * np.abs((x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1))))


@pytest.mark.parametrize("p", [0, 2, 4])
@pytest.mark.parametrize("which", ["sv1", "jittered"])
def test_mass_of_a_constant_is_state_times_area(which, p):
    m = golden_mesh("sv1") if which == "sv1" else mesh(7, seed=3)
    area = shoelace_area(m)
    if which == "jittered":
        assert abs(area - 100.0) <= 1e-12 * 100.0
    U = (1.3, -0.4, 0.25, 2.9)
    prob = dg.euler_constant(U)
    o = orc.Oracle.from_mesh(m, p, prob)
    u = o.project_ic(orc.ic_array(prob, m.period))
    got = o.mass(u)
    want = np.array(U) * area
    assert np.all(np.abs(got - want) <= 1e-13 * np.abs(want)), (got, want)


def test_mass_of_the_gaussian_pulse_is_pi_over_k():
    m = dg.Mesh.periodic_jittered(64, 64, -1.0, 1.0, -1.0, 1.0, 0.1, 1)
    prob = dg.advection_gaussian(a=(1.0, 1.0), center=(0.0, 0.0), k=40.0)
    o = orc.Oracle.from_mesh(m, 4, prob)
    u = o.project_ic(orc.ic_array(prob, m.period))
    got = o.mass(u)[0]
    assert abs(got - math.pi / 40.0) <= 1e-13 * (math.pi / 40.0), got


def test_mass_is_linear_and_reads_mode_zero_only():
    """Higher modes integrate to zero over each element (orthogonal to the
    constant): changing them leaves the mass alone."""
    m = mesh(5, seed=2)
    prob = dg.euler_constant((1.0, 0.0, 0.0, 2.5))
    o = orc.Oracle.from_mesh(m, 3, prob)
    rng = np.random.default_rng(0)
    u = rng.standard_normal(o.shape)
    v = u.copy()
    v[:, 1:, :] = rng.standard_normal(v[:, 1:, :].shape)
    assert np.array_equal(o.mass(u), o.mass(v))
    w = rng.standard_normal(o.shape)
    np.testing.assert_allclose(o.mass(2.0 * u - 3.0 * w), 2.0 * o.mass(u) - 3.0 * o.mass(w), rtol=1e-13, atol=1e-12)


def vortex_state(x, y, vort, L=10.0, gamma=1.4):
    """Conserved state of superposed isentropic vortices on the background
    rho = p = 1, u = v = 1, from the configurations' definition."""
    du = np.zeros_like(x)
    dv = np.zeros_like(x)
    dT = np.zeros_like(x)
    for beta, xc, yc in vort:
        dx = (x - xc + L / 2) % L - L / 2  # nearest periodic image
        dy = (y - yc + L / 2) % L - L / 2
        r2 = dx ** 2 + dy ** 2
        du -= beta / (2 * math.pi) * dy * np.exp((1 - r2) / 2)
        dv += beta / (2 * math.pi) * dx * np.exp((1 - r2) / 2)
        dT -= (gamma - 1) * beta ** 2 / (8 * gamma * math.pi ** 2) * np.exp(1 - r2)
    T = 1.0 + dT
    rho = T ** (1 / (gamma - 1))
    p = rho ** gamma
    u, v = 1.0 + du, 1.0 + dv
    return np.stack([rho, rho * u, rho * v, p / (gamma - 1) + 0.5 * rho * (u ** 2 + v ** 2)])


@pytest.mark.parametrize("nv", [1, 2, 3, 4])
def test_vortex_superposition_matches_its_definition(nv):
    rng = np.random.default_rng(100 + nv)
    vort = [(float(rng.uniform(2, 5)), float(rng.uniform(0, 10)), float(rng.uniform(0, 10))) for _ in range(nv)]
    prob = dg.euler_vortex(tuple(vort))
    ic = orc.ic_array(prob, (10.0, 10.0))
    # random points, plus points near the edges where the nearest image of
    # a vortex is across the periodic boundary
    pts = np.concatenate([rng.uniform(0, 10, (200, 2)),
                          np.column_stack([rng.uniform(0, 0.3, 50), rng.uniform(0, 10, 50)]),
                          np.column_stack([rng.uniform(0, 10, 50), rng.uniform(9.7, 10, 50)])])
    want = vortex_state(pts[:, 0], pts[:, 1], vort)
    got = np.array([orc.ic_point(ic, 4, 1.4, float(x), float(y)) for x, y in pts]).T
    err = np.abs(got - want).max() / np.abs(want).max()
    assert err <= 1e-14, err


def test_vortex_superposition_is_not_a_sum_of_states():
    """The vortices superpose in velocity and temperature, not in conserved
    variables: adding three single-vortex states (less twice the
    background) is far from the combined state where they overlap."""
    vort = [(4.0, 4.2, 5.0), (3.0, 5.6, 5.3), (5.0, 4.9, 6.4)]
    ic = orc.ic_array(dg.euler_vortex(tuple(vort)), (10.0, 10.0))
    bg = orc.ic_point(orc.ic_array(dg.euler_constant((1.0, 1.0, 1.0, 1.0 / 0.4 + 1.0))), 4, 1.4, 0.0, 0.0)
    x, y = 4.9, 5.4
    parts = sum(orc.ic_point(orc.ic_array(dg.euler_vortex((v,)), (10.0, 10.0)), 4, 1.4, x, y) for v in vort)
    got = orc.ic_point(ic, 4, 1.4, x, y)
    np.testing.assert_allclose(got, vortex_state(np.array([x]), np.array([y]), vort)[:, 0], rtol=1e-14)
    assert np.abs(got - (parts - 2 * bg)).max() > 1e-2
