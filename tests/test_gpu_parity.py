"""GPU parity: the HIP path through the C ABI against the FP64 oracle.

Bars: FP64 bulk within 1e-12 relative L2 of the oracle (single RHS
evaluations within 1e-13 relative max-norm); FP32 bulk within the tolerance
written in each test.
"""
import numpy as np
import pytest

import dgcuda_amd as dg
import oracle as orc
from helpers import mesh, perturbed_state, problem_for, rel_max

pytestmark = pytest.mark.gpu

RHS_TOL = {dg.DG_FP64: 1e-13, dg.DG_FP32: 2e-5}


@pytest.mark.parametrize("prec", [dg.DG_FP64, dg.DG_FP32])
@pytest.mark.parametrize("p", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("N", [1, 4])
def test_rhs_parity(N, p, prec):
    print("rhs parity N=%d p=%d prec=%d" % (N, p, prec), flush=True)
    m = mesh(6)
    prob = problem_for(N)
    o = orc.Oracle.from_mesh(m, p, prob)
    u = perturbed_state(o, prob, m.period, seed=p)
    ref = o.rhs(u, dt=0.01)
    assert np.isfinite(ref).all(), "oracle reference not finite: test data out of the physical range"
    s = dg.Solver(m, prob, p, prec)
    got = s.rhs(u, dt=0.01)
    err = rel_max(got, ref)
    print("  rel max err %.3e" % err)
    assert err <= RHS_TOL[prec]


@pytest.mark.parametrize("N", [1, 4])
def test_project_ic_parity(N):
    m = mesh(10)
    prob = problem_for(N)
    for p in range(6):
        o = orc.Oracle.from_mesh(m, p, prob)
        ref = o.project_ic(orc.ic_array(prob, m.period))
        s = dg.Solver(m, prob, p, dg.DG_FP64)
        s.project_ic()
        got = s.get_state()
        # device exp / pow may differ from glibc's by an ulp
        assert rel_max(got, ref) < 1e-13, p


@pytest.mark.parametrize("scheme", [dg.DG_RK4, dg.DG_RK2])
@pytest.mark.parametrize("N,p", [(1, 1), (4, 2), (4, 4)])
def test_steps_parity_fp64(N, p, scheme):
    m = mesh(8)
    prob = problem_for(N)
    o = orc.Oracle.from_mesh(m, p, prob)
    u0 = o.project_ic(orc.ic_array(prob, m.period))
    ref, t_ref = o.run_steps(u0, 10, 0.3, scheme)
    s = dg.Solver(m, prob, p, dg.DG_FP64)
    s.set_state(u0)
    s.run_steps(10, 0.3, scheme, use_graph=False)
    got = s.get_state()
    err = orc.rel_l2(got, ref)
    print("N=%d p=%d scheme=%d rel L2 %.3e" % (N, p, scheme, err))
    assert err <= 1e-12
    assert abs(s.audit()["t"] - t_ref) <= 1e-14 * t_ref


def test_cfl_dt_parity():
    m = mesh(12)
    prob = problem_for(4)
    for p in (1, 3):
        o = orc.Oracle.from_mesh(m, p, prob)
        u0 = o.project_ic(orc.ic_array(prob, m.period))
        s = dg.Solver(m, prob, p, dg.DG_FP64)
        s.set_state(u0)
        assert abs(s.cfl_dt(0.3) - o.cfl_dt(u0, 0.3)) <= 1e-15 * o.cfl_dt(u0, 0.3)


def test_c1_advection_full_run_fp64():
    """C1: advection p=1 on 32x32 jittered squares of [-1,1]^2 to t=2."""
    m = dg.Mesh.periodic_jittered(32, 32, -1.0, 1.0, -1.0, 1.0, 0.1, 1)
    prob = dg.advection_gaussian(a=(1.0, 1.0), center=(0.0, 0.0), k=40.0)
    o = orc.Oracle.from_mesh(m, 1, prob)
    u0 = o.project_ic(orc.ic_array(prob, m.period))
    ref, n_ref, t_ref = o.integrate(u0, 2.0, 0.3)
    s = dg.Solver(m, prob, 1, dg.DG_FP64)
    s.project_ic()
    n, t = s.integrate(2.0, 0.3)
    got = s.get_state()
    err = orc.rel_l2(got, ref)
    print("C1 steps %d/%d rel L2 vs oracle %.3e" % (n, n_ref, err))
    assert n == n_ref and t == 2.0
    assert err <= 1e-12
    mass0, mass1 = o.mass(u0)[0], s.audit()["mass"][0]
    assert abs(mass1 - mass0) <= 1e-13 * abs(mass0)


def test_c2_vortex_coarse_full_run():
    """C2 physics at 32x32: FP64 within 1e-12 and FP32 within 1e-5 of the oracle at t=2."""
    m = mesh(32)
    prob = problem_for(4)
    o = orc.Oracle.from_mesh(m, 2, prob)
    u0 = o.project_ic(orc.ic_array(prob, m.period))
    ref, n_ref, _ = o.integrate(u0, 2.0, 0.3)
    for prec, tol in ((dg.DG_FP64, 1e-12), (dg.DG_FP32, 1e-5)):
        s = dg.Solver(m, prob, 2, prec)
        s.project_ic()
        n, t = s.integrate(2.0, 0.3)
        err = orc.rel_l2(s.get_state(), ref)
        print("prec %d steps %d/%d rel L2 %.3e" % (prec, n, n_ref, err))
        assert err <= tol


@pytest.mark.parametrize("kernel", [1, 2])
def test_integrate_continues(kernel):
    """A run to t_end followed by another run (to a later t_end, or a number
    of steps) starts from the wave speed of the state the first run left:
    FP64 within 1e-12 of the oracle doing the same, with the same step
    counts.  The steps a run enqueues past t_end are not taken and must not
    consume that wave speed (they once did: the next run's first step
    divided by a zero maximum and jumped to its end time)."""
    m = mesh(16)
    prob = problem_for(4)
    o = orc.Oracle.from_mesh(m, 2, prob)
    u0 = o.project_ic(orc.ic_array(prob, m.period))
    u1, n1, _ = o.integrate(u0, 0.1, 0.3)
    u2, n2, _ = o.integrate(u1, 0.1, 0.3)
    u3, _ = o.run_steps(u2, 3, 0.3)
    s = dg.Solver(m, prob, 2, dg.DG_FP64)
    s.set_kernel(kernel)
    s.project_ic()
    g1, t1 = s.integrate(0.1, 0.3)
    e1 = orc.rel_l2(s.get_state(), u1)
    g2, t2 = s.integrate(0.2, 0.3)
    e2 = orc.rel_l2(s.get_state(), u2)
    s.run_steps(3, 0.3)
    e3 = orc.rel_l2(s.get_state(), u3)
    print("kernel %d steps %d+%d (oracle %d+%d) rel L2 %.2e %.2e %.2e" % (kernel, g1, g2, n1, n2, e1, e2, e3))
    assert (g1, g2) == (n1, n2) and t1 == 0.1 and t2 == 0.2
    assert e1 <= 1e-12 and e2 <= 1e-12 and e3 <= 1e-12


def test_graph_replay_bitwise():
    m = mesh(16)
    prob = problem_for(4)
    a = dg.Solver(m, prob, 3, dg.DG_FP32)
    b = dg.Solver(m, prob, 3, dg.DG_FP32)
    a.project_ic()
    b.project_ic()
    a.run_steps(7, 0.3, use_graph=True)
    b.run_steps(7, 0.3, use_graph=False)
    assert np.array_equal(a.get_state(), b.get_state())
