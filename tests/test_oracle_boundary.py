"""The oracle's boundary conditions and the supersonic vortex (the
reference's euler/supersonic problem: the annulus sector 1 <= r <= 1.384
with reflecting arcs and inflow / outflow rays), on the CPU.

Pins:
* the supersonic vortex state against the textbook isentropic vortex
  between two arcs: density (1 + (g-1)/2 M^2 (1 - 1/r^2))^(1/(g-1)), speed
  M / r along the arcs (clockwise), p = rho^g / g (the reference writes the
  constants for g = 1.4, M = 2.25 as 1.0125 and 2.5);
* the steady-state residual of the projected exact solution on the
  reference's three shipped meshes (sv1, sv1refined, sv1refined1: each
  refinement halves h) falls at the rate h^p for p = 1..3.  A boundary
  state at a wrong point, a normal of the wrong orientation or a flipped
  wall velocity makes the residual stall at the boundary sides;
* boundary sides of every kind are present in those meshes (36 reflecting,
  3 outflow, 3 inflow in sv1), all on local faces 0 and 2; relabelling every
  element's vertices moves them to face 1 and leaves the residual norm
  unchanged.
The golden vertex dumps that ship with the reference's Euler output
(euler/output/U0..U3.msh, on sv1refined1) record neither their time nor
their order; the projected state matches them to 3e-5 only, so they are
not used as vectors.
"""
import math
import os

import numpy as np
import pytest

import dgcuda_amd as dg
import oracle as orc
from helpers import relabelled

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
MESHES = ("sv1", "sv1refined", "sv1refined1")


def mesh(name):
    return dg.Mesh.read_pmsh(os.path.join(GOLD, name + ".pmsh"))


def test_boundary_kinds_present():
    V, es, sd, si = mesh("sv1").arrays()
    right = np.asarray(si).reshape(4, -1)[1]
    assert (right == -1).sum() == 36 and (right == -2).sum() == 3 and (right == -3).sum() == 3


@pytest.mark.parametrize("gamma,mach", [(1.4, 2.25), (1.4, 1.5)])
def test_supersonic_state_closed_form(gamma, mach):
    if gamma != 1.4:
        pytest.skip("the reference's literals 1.0125 and 2.5 hold for gamma = 1.4")
    prob = dg.supersonic_vortex(mach=mach, gamma=gamma)
    ic = orc.ic_array(prob)
    rng = np.random.default_rng(3)
    for _ in range(50):
        r = rng.uniform(1.0, 1.384)
        th = rng.uniform(0.05, math.pi / 2 - 0.05)
        x, y = r * math.cos(th), r * math.sin(th)
        U = orc.ic_point(ic, 4, gamma, x, y)
        # 1.0125 = (gamma - 1)/2 * 2.25^2: the reference hard-codes the inner Mach number
        rho = (1 + (gamma - 1) / 2 * 2.25 ** 2 * (1 - 1 / r ** 2)) ** (1 / (gamma - 1))
        speed = mach / r
        u, v = speed * math.sin(th), -speed * math.cos(th)
        p = rho ** gamma / gamma
        want = [rho, rho * u, rho * v, p / (gamma - 1) + 0.5 * rho * speed ** 2]
        assert np.allclose(U, want, rtol=1e-13, atol=1e-14), (r, th, U, want)


def test_supersonic_float_pow_quirk_rounds_through_float():
    prob = dg.supersonic_vortex(quirks=dg.DG_QUIRK_FLOAT_POW)
    U = orc.ic_point(orc.ic_array(prob), 4, 1.4, 1.1, 0.3)
    assert U[0] == float(np.float32(U[0]))  # rho is a float power
    clean = orc.ic_point(orc.ic_array(dg.supersonic_vortex()), 4, 1.4, 1.1, 0.3)
    assert abs(U[0] - clean[0]) <= 1e-7 * clean[0] and U[0] != clean[0]


def residual_of(m, p, quirks=0):
    """L2 norm of M^-1 R of the projected steady state (orthonormal basis:
    the J-weighted sum of squared coefficients)."""
    prob = dg.supersonic_vortex(quirks=quirks)
    o = orc.Oracle.from_mesh(m, p, prob)
    u = o.project_ic(orc.ic_array(prob, m.period))
    k = o.rhs(u, 1.0)
    ge, _, _ = m.geometry(0)
    return math.sqrt(float((ge[0] * (k ** 2).sum(axis=(0, 1))).sum()))


def residual(name, p, quirks=0):
    return residual_of(mesh(name), p, quirks)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_steady_residual_converges_at_order_p(p):
    res = [residual(n, p) for n in MESHES]
    rates = [math.log2(a / b) for a, b in zip(res, res[1:])]
    print("p=%d residuals %s rates %s" % (p, ["%.3e" % r for r in res], ["%.2f" % r for r in rates]))
    for r in rates:
        assert r >= 0.85 * p, rates


def test_boundary_mesh_runs_and_conserves_nothing_spurious():
    """A short run on sv1 stays finite and physical, and the state stays
    close to the exact steady state it starts from."""
    m = mesh("sv1refined")
    prob = dg.supersonic_vortex()
    o = orc.Oracle.from_mesh(m, 2, prob)
    u0 = o.project_ic(orc.ic_array(prob, m.period))
    u, n, t = o.integrate(u0, 0.02, 0.3)
    assert n > 0 and np.isfinite(u).all()
    assert orc.rel_l2(u, u0) < 1e-3


@pytest.mark.parametrize("p", [1, 2, 3])
def test_residual_invariant_under_vertex_relabelling(p):
    """The residual norm of the steady state does not depend on the local
    vertex numbering (same space, symmetric rules), which checks the face-1
    point mapping of inflow, outflow and reflecting sides.  Under
    DG_QUIRK_REFLECT_MIRROR the reference's mirrored wall point on face 1
    makes the steady state inconsistent there (a residual of order 1 that
    does not fall with p)."""
    m0 = mesh("sv1refined")
    m1 = relabelled(m0)
    si = m1.arrays()[3]
    assert (si[2][si[1] < 0] == 1).sum() > 0  # boundary sides on face 1 now
    a, b = residual_of(m0, p), residual_of(m1, p)
    assert abs(a - b) <= 1e-6 * a, (a, b)
    assert residual_of(m1, p, dg.DG_QUIRK_REFLECT_MIRROR) > 3 * a
