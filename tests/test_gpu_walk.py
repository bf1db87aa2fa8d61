"""The persistent matrix-core walk against the FP64 oracle: workgroups that
process more than one tile per launch, with the next tile's state, geometry
and halo ids in flight while the current tile computes.  That is the form
the 524,288-triangle benchmark runs (8,192 tiles on the resident grid).

Two routes make the walks long:
  * a mesh larger than the resident grid covers: 192x192 jittered squares
    (73,728 triangles, 1,152 tiles) at p = 4, above the 768 (FP32) or 512
    (FP64) workgroups resident on an MI355X;
  * the grid capped at 8 workgroups (dg_solver_set_max_groups), so one
    workgroup walks its XCD's whole range of tiles on meshes the oracle
    runs in well under a second, shuffled and with boundary sides.

Bars: FP64 within 1e-12 relative (max-norm per RHS, L2 over steps with the
same step count), FP32 within 2e-5 per RHS and 1e-5 over steps.  A capped
grid changes which workgroup computes a tile, not the tile's arithmetic:
its results equal the uncapped grid's bit for bit.
"""
import numpy as np
import pytest

import dgcuda_amd as dg
import oracle as orc
from dgcuda_amd.workloads import vortices
from helpers import golden_mesh, mesh, perturbed_state, problem_for, rel_max, shuffled_mesh

pytestmark = pytest.mark.gpu

RHS_TOL = {dg.DG_FP32: 2e-5, dg.DG_FP64: 1e-12}
RUN_TOL = {dg.DG_FP32: 1e-5, dg.DG_FP64: 1e-12}


@pytest.mark.parametrize("prec", [dg.DG_FP32, dg.DG_FP64])
def test_walk_at_full_shape_vs_oracle(prec):
    """73,728 triangles, p = 4, three vortices: more tiles than resident
    workgroups, so every workgroup walks at least two tiles per launch."""
    print("walk full shape prec=%d" % prec, flush=True)
    m = mesh(192, seed=2024)
    prob = dg.euler_vortex(vortices(2024))
    o = orc.Oracle.from_mesh(m, 4, prob)
    u0 = o.project_ic(orc.ic_array(prob, m.period))
    s = dg.Solver(m, prob, 4, prec)
    s.set_kernel(2)
    info = s.launch_info()
    print("  launch %s" % info, flush=True)
    assert info["kind"] == 2 and info["tiles"] == 1152
    assert info["groups"] < info["tiles"] and info["walk"] >= 2
    u = perturbed_state(o, prob, m.period, seed=4)
    err = rel_max(s.rhs(u, dt=0.01), o.rhs(u, dt=0.01))
    print("  rhs rel max err %.3e" % err, flush=True)
    assert err <= RHS_TOL[prec]
    ref, t_ref = o.run_steps(u0, 3, 0.3)
    s.set_state(u0)
    s.run_steps(3, 0.3)
    got = s.get_state()
    err = orc.rel_l2(got, ref)
    print("  3 RK4 steps rel L2 %.3e" % err, flush=True)
    assert err <= RUN_TOL[prec]
    assert s.audit()["steps"] == 3


@pytest.mark.parametrize("prec", [dg.DG_FP32, dg.DG_FP64])
@pytest.mark.parametrize("p", [3, 4])
@pytest.mark.parametrize("N", [1, 4])
def test_capped_grid_walks_every_tile_vs_oracle(N, p, prec):
    """8 workgroups on a shuffled 3,200-triangle mesh (50 tiles, irregular
    halos): each workgroup walks 7 tiles per launch."""
    print("capped walk N=%d p=%d prec=%d" % (N, p, prec), flush=True)
    m = shuffled_mesh(40, seed=13)
    prob = problem_for(N)
    o = orc.Oracle.from_mesh(m, p, prob)
    u = perturbed_state(o, prob, m.period, seed=p + 7)
    ref = o.rhs(u, dt=0.01)
    s = dg.Solver(m, prob, p, prec)
    s.set_kernel(2)
    full = s.rhs(u, dt=0.01)
    s.set_max_groups(8)
    info = s.launch_info()
    assert info["groups"] == 8 and info["tiles"] == 50 and info["walk"] == 7, info
    got = s.rhs(u, dt=0.01)
    assert np.array_equal(got, full)
    err = rel_max(got, ref)
    print("  rhs rel max err %.3e" % err, flush=True)
    assert err <= RHS_TOL[prec]
    u0 = o.project_ic(orc.ic_array(prob, m.period))
    ref, n_ref, _ = o.integrate(u0, 0.2, 0.3)
    s.project_ic()
    n, t = s.integrate(0.2, 0.3)
    err = orc.rel_l2(s.get_state(), ref)
    print("  run to t=0.2: %d/%d steps, rel L2 %.3e" % (n, n_ref, err), flush=True)
    assert t == 0.2 and err <= RUN_TOL[prec]
    if prec == dg.DG_FP64:
        assert n == n_ref


@pytest.mark.parametrize("prec", [dg.DG_FP32, dg.DG_FP64])
@pytest.mark.parametrize("p", [2, 4])
def test_capped_grid_with_boundary_sides_vs_oracle(p, prec):
    """The boundary variant of the walk: the supersonic vortex on the
    reference's sv1refined1 mesh (2,880 triangles, 45 tiles) on 8
    workgroups, 6 tiles each."""
    print("capped walk boundary p=%d prec=%d" % (p, prec), flush=True)
    m = golden_mesh("sv1refined1")
    prob = dg.supersonic_vortex()
    o = orc.Oracle.from_mesh(m, p, prob)
    u = perturbed_state(o, prob, m.period, seed=p, amp=1e-3)
    ref = o.rhs(u, dt=0.01)
    s = dg.Solver(m, prob, p, prec)
    s.set_kernel(2)
    full = s.rhs(u, dt=0.01)
    s.set_max_groups(8)
    info = s.launch_info()
    assert info["groups"] == 8 and info["walk"] == 6, info
    got = s.rhs(u, dt=0.01)
    assert np.array_equal(got, full)
    err = rel_max(got, ref)
    print("  rhs rel max err %.3e" % err, flush=True)
    # near the steady state k is a small difference of O(1) fluxes (FP32)
    assert err <= (1e-13 if prec == dg.DG_FP64 else 1e-4)
    u0 = o.project_ic(orc.ic_array(prob, m.period))
    ref, t_ref = o.run_steps(u0, 5, 0.3)
    s.set_state(u0)
    s.run_steps(5, 0.3)
    err = orc.rel_l2(s.get_state(), ref)
    print("  5 steps rel L2 %.3e" % err, flush=True)
    assert err <= RUN_TOL[prec]


def batch_cases():
    """A 51,200-triangle p = 4 Euler solve (800 tiles, runs of consecutive
    tiles per workgroup) beside small solves of every order and both
    physics."""
    big = (dg.Mesh.periodic_jittered(160, 160, 0.0, 10.0, 0.0, 10.0, 0.1, 31),
           dg.euler_vortex(vortices(31)), 4)
    small = [
        (dg.Mesh.periodic_jittered(20, 20, 0.0, 10.0, 0.0, 10.0, 0.1, 32),
         dg.advection_gaussian(a=(0.6, -0.3), center=(4.0, 6.0), k=0.3), 3),
        (dg.Mesh.periodic_jittered(16, 16, 0.0, 10.0, 0.0, 10.0, 0.1, 33), dg.euler_vortex(vortices(33, n=1)), 2),
        (dg.Mesh.periodic_jittered(12, 12, 0.0, 10.0, 0.0, 10.0, 0.1, 34), dg.euler_vortex(vortices(34, n=2)), 4),
        (dg.Mesh.periodic_jittered(9, 9, 0.0, 10.0, 0.0, 10.0, 0.1, 35),
         dg.advection_gaussian(a=(-0.2, 0.9), center=(2.0, 3.0), k=0.5), 1),
    ]
    return [big] + small


@pytest.mark.parametrize("prec", [dg.DG_FP32, dg.DG_FP64])
def test_batch_runs_of_tiles_vs_single_and_oracle(prec):
    """The batch's pipelined runs (batch_mfma_run) over two or more tiles of
    one solve: equal to the single solve bit for bit at p >= 3, and (FP64)
    within 1e-12 of the oracle after 3 steps.  The same batch on 8
    workgroups walks runs of about 100 tiles with the same results."""
    print("batch runs prec=%d" % prec, flush=True)
    cases = batch_cases()
    b = dg.Batch(prec)
    for m, prob, p in cases:
        b.add(m, prob, p, 1.0)
    b.project_ic()
    b.run_steps(3, 0.3)
    info = b.launch_info()
    print("  launch %s" % info, flush=True)
    assert info["longest_run"] >= 2
    states = [b.get_state(k) for k in range(len(cases))]
    for k, (m, prob, p) in enumerate(cases):
        s = dg.Solver(m, prob, p, prec)
        s.set_kernel(2 if p >= 3 else 1)
        s.project_ic()
        s.run_steps(3, 0.3)
        want = s.get_state()
        if p >= 3:
            assert np.array_equal(states[k], want), k
        else:
            assert rel_max(states[k], want) <= (1e-6 if prec == dg.DG_FP32 else 1e-14), k
        if prec == dg.DG_FP64:
            o = orc.Oracle.from_mesh(m, p, prob)
            ref, _ = o.run_steps(o.project_ic(orc.ic_array(prob, m.period)), 3, 0.3)
            err = orc.rel_l2(states[k], ref)
            print("  slot %d p=%d N=%d rel L2 vs oracle %.3e" % (k, p, prob.physics, err), flush=True)
            assert err <= 1e-12
    bc = dg.Batch(prec)
    for m, prob, p in cases:
        bc.add(m, prob, p, 1.0)
    bc.set_max_groups(8)
    bc.project_ic()
    bc.run_steps(3, 0.3)
    info = bc.launch_info()
    print("  capped launch %s" % info, flush=True)
    assert info["groups"] == 8 and info["longest_run"] >= 90
    for k in range(len(cases)):
        assert np.array_equal(bc.get_state(k), states[k]), k


@pytest.mark.parametrize("prec", [dg.DG_FP32, dg.DG_FP64])
def test_batch_per_class_launches_equal_one_launch(prec):
    """The per-class form (one persistent launch per stage for each kernel
    class: Euler tiles, advection tiles, Euler and advection lane items)
    gives the one-launch form's results bit for bit, also with every class
    kernel capped at 8 workgroups (long runs in each)."""
    print("batch per-class prec=%d" % prec, flush=True)
    cases = batch_cases()
    got = {}
    for mode, cap in ((dg.DG_BATCH_ONE_LAUNCH, 0), (dg.DG_BATCH_PER_CLASS, 0), (dg.DG_BATCH_PER_CLASS, 8)):
        b = dg.Batch(prec)
        b.set_launch_mode(mode)
        for m, prob, p in cases:
            b.add(m, prob, p, 1.0)
        if cap:
            b.set_max_groups(cap)
        b.project_ic()
        b.run_steps(3, 0.3)
        info = b.launch_info()
        print("  mode %d cap %d launch %s" % (mode, cap, info), flush=True)
        assert info["tile_items"] > 0 and info["lane_items"] > 0
        if cap:
            assert info["groups"] == 8 and info["longest_run"] >= 90
        got[(mode, cap)] = [b.get_state(k) for k in range(len(cases))]
        b.close()
    ref = got[(dg.DG_BATCH_ONE_LAUNCH, 0)]
    for key, states in got.items():
        for k in range(len(cases)):
            assert np.array_equal(states[k], ref[k]), (key, k)
