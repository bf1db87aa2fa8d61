"""GPU parity on meshes with boundary sides: the supersonic vortex on the
reference's shipped annulus-sector meshes (reflecting arcs, inflow and
outflow rays), through the C ABI against the FP64 oracle.

Bars: FP64 within 1e-13 relative max-norm per RHS and 1e-12 relative L2
over a run with the same step count; FP32 within 1e-4 per RHS (a
near-steady state: k is a small difference of O(1) fluxes) and 1e-5 over a
run.  Boundary meshes run on the lane-per-element kernel up to p = 3 and,
on the matrix-core kernels (kinds 2 and 3, their boundary variant, Euler
and advection) up to p = 4; in a batch such a solve runs its own kernels inside
the batch's captured step.  The first run of the
p = 4 FP64 lane kernel with the boundary path faulted the GPU (memory
aperture violation) and its cause is not found: the lane kernel refuses
boundary meshes at p >= 4 (checked here and on the CPU in test_abi) and no
GPU test runs it.
"""
import numpy as np
import pytest

import dgcuda_amd as dg
import oracle as orc
from helpers import golden_mesh, perturbed_state, rel_max, relabelled

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("quirks", [0, 15])
@pytest.mark.parametrize("prec", [dg.DG_FP64, dg.DG_FP32])
@pytest.mark.parametrize("p", [1, 2, 3])
def test_boundary_rhs_parity(p, prec, quirks):
    print("boundary rhs p=%d prec=%d quirks=%d" % (p, prec, quirks), flush=True)
    m = golden_mesh("sv1refined")
    prob = dg.supersonic_vortex(quirks=quirks)
    o = orc.Oracle.from_mesh(m, p, prob)
    u = perturbed_state(o, prob, m.period, seed=p, amp=1e-3)
    ref = o.rhs(u, dt=0.01)
    s = dg.Solver(m, prob, p, prec)
    for kind in (1, 2, 3):
        s.set_kernel(kind)
        err = rel_max(s.rhs(u, dt=0.01), ref)
        print("  kernel %d rel max err %.3e" % (kind, err))
        # FP32: the state is near the steady solution, so k is a small
        # difference of O(1) fluxes and its relative rounding error is some
        # 300x FP32's
        assert err <= (1e-13 if prec == dg.DG_FP64 else 1e-4)


@pytest.mark.parametrize("quirks", [0, 15])
@pytest.mark.parametrize("prec", [dg.DG_FP64, dg.DG_FP32])
def test_boundary_rhs_parity_p4_matrix_cores(prec, quirks):
    print("boundary rhs p=4 prec=%d quirks=%d" % (prec, quirks), flush=True)
    m = golden_mesh("sv1refined")
    prob = dg.supersonic_vortex(quirks=quirks)
    o = orc.Oracle.from_mesh(m, 4, prob)
    u = perturbed_state(o, prob, m.period, seed=4, amp=1e-3)
    ref = o.rhs(u, dt=0.01)
    s = dg.Solver(m, prob, 4, prec)
    assert s.kernel() in (2, 3)
    for kind in (2, 3):
        s.set_kernel(kind)
        err = rel_max(s.rhs(u, dt=0.01), ref)
        print("  kernel %d rel max err %.3e" % (kind, err))
        assert err <= (1e-13 if prec == dg.DG_FP64 else 1e-4)


@pytest.mark.parametrize("p", [1, 3, 4])
def test_boundary_rhs_parity_face1(p):
    """Boundary sides on local face 1 (vertices relabelled)."""
    m = relabelled(golden_mesh("sv1"))
    prob = dg.supersonic_vortex()
    o = orc.Oracle.from_mesh(m, p, prob)
    u = perturbed_state(o, prob, m.period, seed=5, amp=1e-3)
    ref = o.rhs(u, dt=0.01)
    s = dg.Solver(m, prob, p, dg.DG_FP64)
    for kind in ((1, 2) if p <= 3 else (2,)):
        s.set_kernel(kind)
        assert rel_max(s.rhs(u, dt=0.01), ref) <= 1e-13, kind


@pytest.mark.parametrize("prec", [dg.DG_FP64, dg.DG_FP32])
@pytest.mark.parametrize("p", [2, 3, 4])
def test_boundary_run_vs_oracle(p, prec):
    m = golden_mesh("sv1refined")
    prob = dg.supersonic_vortex()
    o = orc.Oracle.from_mesh(m, p, prob)
    u0 = o.project_ic(orc.ic_array(prob, m.period))
    ref, n_ref, _ = o.integrate(u0, 0.03, 0.3)
    s = dg.Solver(m, prob, p, prec)
    s.project_ic()
    n, t = s.integrate(0.03, 0.3)
    err = orc.rel_l2(s.get_state(), ref)
    print("p=%d prec=%d kernel %d steps %d/%d rel L2 %.3e" % (p, prec, s.kernel(), n, n_ref, err))
    assert t == 0.03
    if prec == dg.DG_FP64:
        assert n == n_ref and err <= 1e-12
    else:
        assert err <= 1e-5


def test_boundary_mesh_kernel_choice():
    m = golden_mesh("sv1")
    prob = dg.supersonic_vortex()
    s = dg.Solver(m, prob, 3, dg.DG_FP32)
    assert s.kernel() in (1, 2, 3)
    s.set_kernel(2)
    s4 = dg.Solver(m, prob, 4, dg.DG_FP32)
    assert s4.kernel() in (2, 3)
    with pytest.raises(dg.DGError) as e:
        s4.set_kernel(1)  # the lane kernel takes boundary sides up to p = 3
    assert e.value.code == dg.STATUS_CODES["DG_ERR_UNSUPPORTED"]
    # a batch takes boundary meshes as solves with their own kernels
    b = dg.Batch(dg.DG_FP32)
    b.add(m, prob, 2, 0.1)
    b.add(m, dg.advection_gaussian(), 4, 0.1)
    with pytest.raises(dg.DGError) as e:
        b.add(m, prob, 5, 0.1)
    assert e.value.code == dg.STATUS_CODES["DG_ERR_UNSUPPORTED"]
    b.close()


@pytest.mark.parametrize("prec", [dg.DG_FP64, dg.DG_FP32])
def test_batch_with_boundary_solves(prec):
    """Supersonic vortices on the reference's sv1refined mesh (p = 2 on the
    lane kernel, p = 3 and 4 on the matrix-core kernel's boundary variant)
    batched with periodic solves: each boundary solve equals the same solve
    alone on that kernel bit for bit, the periodic ones are unaffected, and
    FP64 matches the oracle to 1e-12 with the same step count."""
    m = golden_mesh("sv1refined")
    prob = dg.supersonic_vortex()
    per = dg.Mesh.periodic_jittered(10, 10, 0.0, 10.0, 0.0, 10.0, 0.1, 3)
    vort = dg.euler_vortex(((4.0, 5.0, 5.0),))
    cases = [(m, prob, 2), (per, vort, 3), (m, prob, 3), (per, vort, 1), (m, prob, 4)]
    b = dg.Batch(prec)
    for mm, pr, p in cases:
        b.add(mm, pr, p, 0.03)
    b.project_ic()
    b.integrate(0.3)
    for k, (mm, pr, p) in enumerate(cases):
        s = dg.Solver(mm, pr, p, prec)
        s.set_kernel(2 if p >= 3 else 1)
        s.project_ic()
        n, _ = s.integrate(0.03, 0.3)
        got = b.get_state(k)
        assert b.audit(k)["t"] == 0.03 and b.audit(k)["steps"] == n
        if mm is m or p >= 3:
            assert np.array_equal(got, s.get_state()), k
        else:
            assert rel_max(got, s.get_state()) <= (1e-6 if prec == dg.DG_FP32 else 1e-14), k
        s.close()
        if prec == dg.DG_FP64 and mm is m:
            o = orc.Oracle.from_mesh(mm, p, pr)
            u0 = o.project_ic(orc.ic_array(pr, mm.period))
            ref, n_ref, _ = o.integrate(u0, 0.03, 0.3)
            err = orc.rel_l2(got, ref)
            print("slot %d p=%d steps %d/%d rel L2 vs oracle %.3e" % (k, p, n, n_ref, err))
            assert n == n_ref and err <= 1e-12
    b.close()


def cylinder_mesh(name):
    import os
    from helpers import GOLD_DIR

    return dg.Mesh.read_pmsh(os.path.join(GOLD_DIR, "gmsh", name + ".pmsh"))


@pytest.mark.parametrize("quirks", [0, 15])
@pytest.mark.parametrize("prec", [dg.DG_FP64, dg.DG_FP32])
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_cylinder_rhs_parity(p, prec, quirks):
    """The cylinder problem (euler/cylinder/cylinder.cu): reflecting walls
    about (0.5, 0.5) and the uniform Mach 0.38 far field, on the reference's
    cylinder1 mesh; every kernel that takes boundary sides at this order."""
    print("cylinder rhs p=%d prec=%d quirks=%d" % (p, prec, quirks), flush=True)
    m = cylinder_mesh("cylinder1")
    prob = dg.cylinder(quirks=quirks)
    o = orc.Oracle.from_mesh(m, p, prob)
    u = perturbed_state(o, prob, m.period, seed=p, amp=1e-2)
    ref = o.rhs(u, dt=0.01)
    s = dg.Solver(m, prob, p, prec)
    for kind in ((1, 2, 3) if p <= 3 else (2, 3)):
        s.set_kernel(kind)
        err = rel_max(s.rhs(u, dt=0.01), ref)
        print("  kernel %d rel max err %.3e" % (kind, err))
        assert err <= (1e-13 if prec == dg.DG_FP64 else 1e-4)


@pytest.mark.parametrize("prec", [dg.DG_FP64, dg.DG_FP32])
@pytest.mark.parametrize("name,p", [("cylinder", 2), ("cylinder1", 1), ("cylinder1", 3), ("cylinder1", 4)])
def test_cylinder_rk2_run_vs_oracle(name, p, prec):
    """The reference's run of the cylinder: RK2 from the uniform state (the
    impulsive start around the wall), to t = 0.05."""
    m = cylinder_mesh(name)
    prob = dg.cylinder()
    o = orc.Oracle.from_mesh(m, p, prob)
    u0 = o.project_ic(orc.ic_array(prob, m.period))
    ref, n_ref, _ = o.integrate(u0, 0.05, 0.3, scheme=dg.DG_RK2)
    s = dg.Solver(m, prob, p, prec)
    s.project_ic()
    n, t = s.integrate(0.05, 0.3, scheme=dg.DG_RK2)
    err = orc.rel_l2(s.get_state(), ref)
    print("cylinder %s p=%d prec=%d kernel %d steps %d/%d rel L2 %.3e" % (name, p, prec, s.kernel(), n, n_ref, err))
    assert t == 0.05
    if prec == dg.DG_FP64:
        assert n == n_ref and err <= 1e-12
    else:
        assert err <= 1e-5
    a0, a1 = o.mass(u0), np.array(s.audit()["mass"])
    assert np.isfinite(a1).all() and s.audit()["n_bad"] == 0
