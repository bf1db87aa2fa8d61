"""End-to-end checks of the GPU path that need no reference vectors:

  * p + 1 convergence on the stationary isentropic vortex (an exact steady
    solution) in FP64 on the matrix-core kernel at p = 3 and 4, on the
    periodic square [0, 20]^2 with the vortex at its centre (the vortex's
    tail at the periodic boundary, below e^-49, leaves no kink that would
    cap the rate at fine meshes), meshes of 32, 64 and 128 squares a side:
    the L2 distance of the state at t = 0.5 from its own projection falls at
    rate >= p + 1/2 (measured on MI355X: 3.84 and 4.32 at p = 3, 5.05 and
    4.60 at p = 4; the same errors at CFL 0.1, so not the time step's); at
    32 the error equals the FP64 oracle's (the states agree to 1e-12, so
    errors of order 1e-5 agree to 1e-6 relative);
  * dg_eval_vertices (the reference's eval_u) against the oracle's vertex
    values, every variable, FP64 and FP32 states.
"""
import math

import numpy as np
import pytest

import dgcuda_amd as dg
import oracle as orc
from helpers import mesh, perturbed_state, problem_for

pytestmark = pytest.mark.gpu


def stationary_vortex(n, L=20.0):
    m = dg.Mesh.periodic_jittered(n, n, 0.0, L, 0.0, L, 0.0, 1)
    return m, dg.euler_vortex(((5.0, L / 2, L / 2),), background=(1.0, 0.0, 0.0, 1.0))


CFL = 0.3


def l2_rho(J, u, u0):
    return float(np.sqrt(np.sum(J * np.sum((u - u0)[0] ** 2, axis=0))))


@pytest.mark.parametrize("p", [3, 4])
def test_p_plus_one_convergence_fp64(p):
    errs = []
    for n in (32, 64, 128):
        print("convergence p=%d n=%d" % (p, n), flush=True)
        m, prob = stationary_vortex(n)
        s = dg.Solver(m, prob, p, dg.DG_FP64)
        s.set_kernel(2)
        s.project_ic()
        u0 = s.get_state()
        s.integrate(0.5, CFL)
        u = s.get_state()
        J = m.geometry()[0][0]
        errs.append(l2_rho(J, u, u0))
        if n == 32:
            o = orc.Oracle.from_mesh(m, p, prob)
            ou0 = o.project_ic(orc.ic_array(prob, m.period))
            ou, _, _ = o.integrate(ou0, 0.5, CFL)
            e_or = l2_rho(J, ou, ou0)
            print("  GPU error %.6e, oracle %.6e" % (errs[-1], e_or), flush=True)
            assert abs(errs[-1] - e_or) <= 1e-6 * e_or
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    print("p=%d errors %s rates %s" % (p, errs, rates))
    assert min(rates) >= p + 0.5, rates


@pytest.mark.parametrize("prec", [dg.DG_FP64, dg.DG_FP32])
@pytest.mark.parametrize("N", [1, 4])
def test_eval_vertices_vs_oracle(N, prec):
    m = mesh(9, seed=4)
    prob = problem_for(N)
    for p in (1, 2, 4):
        o = orc.Oracle.from_mesh(m, p, prob)
        u = perturbed_state(o, prob, m.period, seed=p)
        s = dg.Solver(m, prob, p, prec)
        s.set_state(u)
        held = s.get_state()  # FP32: the state rounded to the bulk precision
        for var in range(N):
            want = o.eval_vertices(held, var)
            got = s.eval_vertices(var)
            err = np.abs(got - want).max() / np.abs(want).max()
            assert err <= 1e-14, (p, var, err)


@pytest.mark.parametrize("prec", [dg.DG_FP64, dg.DG_FP32])
def test_velocity_and_pressure_views_vs_oracle(prec):
    """dg_eval_vertices_velocity / _pressure (the reference CPU driver's
    eval_u_velocity and eval_p): from the oracle's vertex values of the held
    state, u = (rho u)/rho, v = (rho v)/rho, p = (gamma - 1)(E - (u^2 + v^2)/2 rho)."""
    m = mesh(9, seed=6)
    prob = problem_for(4)
    for p in (1, 3, 4):
        o = orc.Oracle.from_mesh(m, p, prob)
        s = dg.Solver(m, prob, p, prec)
        s.set_state(perturbed_state(o, prob, m.period, seed=p))
        held = s.get_state()
        rho, mu, mv, E = (o.eval_vertices(held, v) for v in range(4))
        u, v = mu / rho, mv / rho
        pw = 0.4 * (E - (u * u + v * v) / 2.0 * rho)
        uv = s.eval_vertices_velocity()
        assert np.abs(uv[..., 0] - u).max() <= 1e-13 * np.abs(u).max()
        assert np.abs(uv[..., 1] - v).max() <= 1e-13 * np.abs(v).max()
        assert np.abs(s.eval_vertices_pressure() - pw).max() <= 1e-13 * np.abs(pw).max()
