"""Matrix-core stage kernels (FP32 and FP64) against the FP64 oracle and the
VALU kernel.  FP32 tolerance: 2e-5 relative max-norm for one RHS evaluation,
1e-5 relative L2 after a run; FP64: 1e-12 for both."""
import numpy as np
import pytest

import dgcuda_amd as dg
import oracle as orc
from helpers import mesh, perturbed_state, problem_for, rel_max, shuffled_mesh

pytestmark = pytest.mark.gpu


def test_mfma_lane_layout():
    assert dg.selftest_mfma() == 0.0


@pytest.mark.parametrize("prec", [dg.DG_FP32, dg.DG_FP64])
@pytest.mark.parametrize("kind", [2, 3])
@pytest.mark.parametrize("p", [3, 4])
@pytest.mark.parametrize("N", [1, 4])
def test_mfma_rhs_parity(N, p, kind, prec):
    print("mfma rhs N=%d p=%d kind=%d prec=%d" % (N, p, kind, prec), flush=True)
    for m in (mesh(12), shuffled_mesh(12)):
        prob = problem_for(N)
        o = orc.Oracle.from_mesh(m, p, prob)
        u = perturbed_state(o, prob, m.period, seed=p)
        ref = o.rhs(u, dt=0.01)
        assert np.isfinite(ref).all(), "oracle reference not finite: test data out of the physical range"
        s = dg.Solver(m, prob, p, prec)
        s.set_kernel(kind)
        assert s.kernel() == kind
        err = rel_max(s.rhs(u, dt=0.01), ref)
        print("  rel max err %.3e" % err)
        assert err <= (2e-5 if prec == dg.DG_FP32 else 1e-12)


@pytest.mark.parametrize("prec", [dg.DG_FP32, dg.DG_FP64])
@pytest.mark.parametrize("kind", [2, 3])
def test_mfma_run_vs_oracle(kind, prec):
    m = shuffled_mesh(24, seed=3)
    prob = dg.euler_vortex(((3.0, 4.0, 6.0), (2.5, 7.0, 3.0)))
    o = orc.Oracle.from_mesh(m, 4, prob)
    u0 = o.project_ic(orc.ic_array(prob, m.period))
    ref, n_ref, _ = o.integrate(u0, 0.25, 0.3)
    s = dg.Solver(m, prob, 4, prec)
    s.set_kernel(kind)
    s.project_ic()
    n, t = s.integrate(0.25, 0.3)
    err = orc.rel_l2(s.get_state(), ref)
    print("kind %d prec %d steps %d/%d rel L2 %.3e" % (kind, prec, n, n_ref, err))
    assert t == 0.25 and err <= (1e-5 if prec == dg.DG_FP32 else 1e-12)
    if prec == dg.DG_FP64:
        assert n == n_ref
    a0 = o.mass(u0)
    a1 = s.audit()["mass"]
    assert np.all(np.abs(np.array(a1) - a0) <= 2e-6 * np.abs(a0) + 1e-9)


def test_autotuner_picks_matrix_cores_for_p4():
    """Automatic choice times the candidate kernels; at 32,768 triangles and
    p = 4 the matrix-core kernel wins by a wide margin in both precisions
    (the lane kernel spills at this order)."""
    m = mesh(128)
    s = dg.Solver(m, problem_for(4), 4, dg.DG_FP32)
    assert s.kernel() in (2, 3)
    s64 = dg.Solver(m, problem_for(4), 4, dg.DG_FP64)
    assert s64.kernel() in (2, 3)


def test_autotuned_solver_runs_correctly():
    """Whatever kernel the tuner picked, a run from a projected state matches
    the oracle (the tuning launches leave no trace in the solve)."""
    m = mesh(10)
    prob = problem_for(4)
    for p in (1, 2, 3, 4):
        s = dg.Solver(m, prob, p, dg.DG_FP32)
        s.project_ic()
        s.integrate(0.2, 0.3)
        o = orc.Oracle.from_mesh(m, p, prob)
        ref, _, _ = o.integrate(o.project_ic(orc.ic_array(prob, m.period)), 0.2, 0.3)
        err = orc.rel_l2(s.get_state(), ref)
        print("p=%d kernel %d rel L2 %.3e" % (p, s.kernel(), err))
        assert err <= 1e-5


@pytest.mark.parametrize("prec", [dg.DG_FP32, dg.DG_FP64])
@pytest.mark.parametrize("p", [1, 3, 4])
def test_kernel_switch_carries_state(p, prec):
    """The matrix-core kernels keep the FP32 state in their own tile layout
    (modes padded to a multiple of 4; p = 1, 3 pad, p = 4 pads one mode).
    Switching the kernel of a solver that holds a state converts it exactly,
    and steps taken after a switch equal steps on that kernel from the start
    (bitwise).  K = 162: the last tile is partial."""
    print("kernel switch p=%d prec=%d" % (p, prec), flush=True)
    m = shuffled_mesh(9, seed=5)
    prob = problem_for(4)
    o = orc.Oracle.from_mesh(m, p, prob)
    u = perturbed_state(o, prob, m.period, seed=11)
    u32 = u.astype(np.float32).astype(np.float64) if prec == dg.DG_FP32 else u
    s = dg.Solver(m, prob, p, prec)
    s.set_kernel(1)
    s.set_state(u)
    for kind in (2, 1, 3, 2, 1):
        s.set_kernel(kind)
        assert s.kernel() == kind
        assert np.array_equal(s.get_state(), u32), "state changed by the switch to kernel %d" % kind
    s.set_kernel(2)
    s.run_steps(3, 0.3)
    b = dg.Solver(m, prob, p, prec)
    b.set_kernel(2)
    b.set_state(u)
    b.run_steps(3, 0.3)
    got = s.get_state()
    assert np.isfinite(got).all()
    assert np.array_equal(got, b.get_state())
    s.set_kernel(1)
    assert np.array_equal(s.get_state(), got)


@pytest.mark.parametrize("n", [1, 2, 3, 5])
def test_tiny_meshes_every_kernel(n):
    """Meshes of 2 to 50 triangles: one partial tile, neighbours reached
    across several periodic sides, halo lists of tile elements only.  Every
    kernel, at every order the matrix-core kernels take (p = 1, 2 use a
    single mode group of the FP32 tile layout), matches the oracle's RHS:
    FP64 kernels to 1e-12, FP32 kernels to 2e-5 (relative max-norm)."""
    m = dg.Mesh.periodic_jittered(n, n, 0.0, 10.0, 0.0, 10.0, 0.1, 7)
    for N in (1, 4):
        prob = problem_for(N)
        for p in (1, 2, 3, 4):
            print("tiny mesh n=%d N=%d p=%d" % (n, N, p), flush=True)
            o = orc.Oracle.from_mesh(m, p, prob)
            u = perturbed_state(o, prob, m.period, seed=n + p)
            ref = o.rhs(u, dt=0.01)
            assert np.isfinite(ref).all(), "oracle reference not finite: test data out of the physical range"
            for kind in (1, 2, 3):
                s64 = dg.Solver(m, prob, p, dg.DG_FP64)
                s64.set_kernel(kind)
                assert rel_max(s64.rhs(u, dt=0.01), ref) <= 1e-12, kind
            for kind in (1, 2, 3):
                s = dg.Solver(m, prob, p, dg.DG_FP32)
                s.set_kernel(kind)
                err = rel_max(s.rhs(u, dt=0.01), ref)
                assert err <= 2e-5, (kind, err)


@pytest.mark.parametrize("p", [2, 3, 4])
def test_float_sqrt_quirk_fp64_many_states(p):
    """Under DG_QUIRK_FLOAT_SQRT the sound speed goes through a float square
    root, so an argument that differs from the oracle's in its last bits can
    flip the float rounding (a 6e-8 relative change of the wave speed).  The
    FP64 kernels form it in the reference's order without contraction: every
    kernel matches the oracle to 1e-13 on 10 seeded states."""
    m = shuffled_mesh(10, seed=21)
    prob = problem_for(4, quirks=dg.DG_QUIRK_FLOAT_SQRT)
    o = orc.Oracle.from_mesh(m, p, prob)
    s = dg.Solver(m, prob, p, dg.DG_FP64)
    worst = 0.0
    for seed in range(10):
        u = perturbed_state(o, prob, m.period, seed=100 + seed, amp=3e-2)
        ref = o.rhs(u, dt=0.01)
        for kind in (1, 2, 3):
            s.set_kernel(kind)
            worst = max(worst, rel_max(s.rhs(u, dt=0.01), ref))
    print("p=%d worst rel max err %.3e" % (p, worst))
    assert worst <= 1e-13
