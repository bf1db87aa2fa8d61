"""The advection prototype's path on the GPU against the FP64 oracle: the
central flux (every kernel, both precisions), its exact-boundary meshes
read with their sides built at read time (lane kernel up to p = 3,
matrix-core kernels up to p = 4), fixed-step RK4 as the reference's
time-step test runs it, and a central-flux solve in a batch.

Bars: FP64 1e-12 relative (RHS max-norm, runs L2), FP32 2e-5 per RHS and
1e-5 over runs.
"""
import os

import numpy as np
import pytest

import dgcuda_amd as dg
import oracle as orc
from helpers import GOLD_DIR, mesh, perturbed_state, rel_max

pytestmark = pytest.mark.gpu


def central_gaussian():
    prob = dg.advection_gaussian(a=(0.8, -0.5), center=(4.0, 6.0), k=0.4)
    prob.flux = dg.DG_FLUX_CENTRAL
    return prob


@pytest.mark.parametrize("prec", [dg.DG_FP64, dg.DG_FP32])
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_central_flux_rhs_every_kernel(p, prec):
    print("central rhs p=%d prec=%d" % (p, prec), flush=True)
    m = mesh(12, seed=9)
    prob = central_gaussian()
    o = orc.Oracle.from_mesh(m, p, prob)
    u = perturbed_state(o, prob, m.period, seed=p)
    ref = o.rhs(u, dt=0.01)
    s = dg.Solver(m, prob, p, prec)
    for kind in (1, 2, 3):
        s.set_kernel(kind)
        err = rel_max(s.rhs(u, dt=0.01), ref)
        print("  kernel %d rel max err %.3e" % (kind, err))
        assert err <= (1e-12 if prec == dg.DG_FP64 else 2e-5)


@pytest.mark.parametrize("prec", [dg.DG_FP64, dg.DG_FP32])
def test_central_flux_run_and_batch(prec):
    m = mesh(16, seed=3)
    prob = central_gaussian()
    o = orc.Oracle.from_mesh(m, 3, prob)
    u0 = o.project_ic(orc.ic_array(prob, m.period))
    ref, n_ref, _ = o.integrate(u0, 0.4, 0.3)
    s = dg.Solver(m, prob, 3, prec)
    s.set_kernel(2)
    s.project_ic()
    n, t = s.integrate(0.4, 0.3)
    err = orc.rel_l2(s.get_state(), ref)
    print("central run prec=%d steps %d/%d rel L2 %.3e" % (prec, n, n_ref, err))
    assert t == 0.4 and err <= (1e-12 if prec == dg.DG_FP64 else 1e-5)
    b = dg.Batch(prec)
    b.add(m, prob, 3, 0.4)
    b.add(mesh(8, seed=4), dg.euler_vortex(), 2, 0.4)
    b.project_ic()
    b.integrate(0.3)
    assert np.array_equal(b.get_state(0), s.get_state())


@pytest.mark.parametrize("prec", [dg.DG_FP64, dg.DG_FP32])
@pytest.mark.parametrize("name,dt", [("canonical", 0.001), ("box", 0.01), ("supersimple", 0.001)])
def test_exact_boundary_time_steps_vs_oracle(name, dt, prec):
    """(x - y)^alpha on the prototype's meshes, every boundary side exact,
    10 RK4 steps of the test's dt; alpha > p too (not in the space, so the
    state moves).  The lane kernel up to p = 3, the matrix-core kernels'
    boundary variant (kinds 2 and 3) up to p = 4."""
    m = dg.Mesh.read_vertex_pmsh(os.path.join(GOLD_DIR, "t0", name + ".pmsh"))
    for p in (1, 2, 3, 4):
        for alpha in (p, p + 1):
            prob = dg.advection_poly(alpha)
            o = orc.Oracle.from_mesh(m, p, prob)
            u0 = o.project_ic(orc.ic_array(prob, m.period))
            ref = u0
            for _ in range(10):
                ref = o.rk_step(ref, dt, 1)
            for kind in ((1, 2, 3) if p <= 3 else (2, 3)):
                print("exact boundary %s p=%d alpha=%d prec=%d kernel %d" % (name, p, alpha, prec, kind), flush=True)
                s = dg.Solver(m, prob, p, prec)
                s.set_kernel(kind)
                s.set_state(u0)
                for _ in range(10):
                    s.rk_step(dt)
                err = orc.rel_l2(s.get_state(), ref)
                assert err <= (1e-12 if prec == dg.DG_FP64 else 1e-5), err
                assert np.abs(s.eval_vertices(0) - o.eval_vertices(ref, 0)).max() <= (
                    1e-12 if prec == dg.DG_FP64 else 1e-5) * max(1.0, np.abs(ref).max())
                s.close()
