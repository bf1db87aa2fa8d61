"""The benchmark configurations at their full sizes, through properties that
do not need a full oracle run, plus the oracle on the first steps.

  C2  vortex p = 2, 128x128 squares (32,768 triangles), t = 0..2: FP64 within
      1e-12 of the oracle over the first 10 steps; the full run in FP64 and
      FP32, mass conserved to round-off (FP64 1e-13, FP32 audit of an FP32
      state 1e-5), and the state at t = 2 close to the exact solution, the
      vortex carried by the background flow (u = v = 1) to (7, 7): FP64
      relative L2 distance below 1e-5 from its projection (the oracle
      measures 1.3e-4 at 32x32 and 1.9e-5 at 64x64, third order: about
      2.5e-6 expected here; measured 2.7e-6).  FP32 against FP64 is stated,
      not fitted: the FP32 rounding of the fluxes is the same at every step
      where the flow is steady, so the difference grows with the step count
      (measured 1.2e-5, 2.1e-5 and 3.4e-5 after 828, 1,656 and 3,312 steps);
      bars 2.5e-5 at t = 0.5 and 1e-4 at t = 2, growth no faster than the
      step count, and FP32 within 1e-4 of the exact vortex.
  C3  three vortices p = 4, 512x512 squares (524,288 triangles): 2 RK4
      steps against the oracle (FP64 1e-12, FP32 1e-5) on the persistent
      matrix-core kernel the benchmark times; the configuration's 200 steps
      in FP32 within 1e-5 of the same run in FP64 (the GPU FP64 path stands
      for the oracle here: 200 oracle steps take about half an hour on one
      core), mass conserved (FP64 1e-13, FP32 2e-6).
  C4  the 64-solve ragged batch (500 to 200k triangles, p = 1..4, advection
      and Euler): 3 batched steps equal the single solves bit for bit
      (p >= 3) or to 1e-6 / 1e-14 (p <= 2) on a sample of 8 slots spanning
      the sizes.
"""
import numpy as np
import pytest

import dgcuda_amd as dg
import oracle as orc
from dgcuda_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_c2_full_size():
    m, prob, p, t_end = W.vortex_c2()
    assert m.num_elem == 32768
    o = orc.Oracle.from_mesh(m, p, prob)
    u0 = o.project_ic(orc.ic_array(prob, m.period))
    ref10, _ = o.run_steps(u0, 10, 0.3)
    s64 = dg.Solver(m, prob, p, dg.DG_FP64)
    s64.set_state(u0)
    s64.run_steps(10, 0.3)
    err = orc.rel_l2(s64.get_state(), ref10)
    print("C2 10 steps FP64 rel L2 vs oracle %.3e" % err, flush=True)
    assert err <= 1e-12
    s64.project_ic()
    s32 = dg.Solver(m, prob, p, dg.DG_FP32)
    s32.project_ic()
    growth = []
    for t in (0.5, 1.0, t_end):
        n64, t64 = s64.integrate(t, 0.3)
        n32, t32 = s32.integrate(t, 0.3)
        assert t64 == t and t32 == t and n64 == n32
        e32 = orc.rel_l2(s32.get_state(), s64.get_state())
        growth.append((s64.audit()["steps"], e32))
        print("C2 t=%.1f after %d steps: FP32 vs FP64 rel L2 %.3e" % (t, growth[-1][0], e32), flush=True)
    a64, a32 = s64.get_state(), s32.get_state()
    m0 = o.mass(u0)
    mass64 = np.abs(np.array(s64.audit()["mass"]) - m0) / np.abs(m0)
    mass32 = np.abs(np.array(s32.audit()["mass"]) - m0) / np.abs(m0)
    moved = dg.euler_vortex(((5.0, 7.0, 7.0),))
    exact = dg.Solver(m, moved, p, dg.DG_FP64)
    exact.project_ic()
    e64 = orc.rel_l2(a64, exact.get_state())
    e32 = orc.rel_l2(a32, exact.get_state())
    print("C2 t=2 rel L2 vs the carried exact vortex: FP64 %.3e FP32 %.3e; mass drift FP64 %.1e FP32 %.1e"
          % (e64, e32, mass64.max(), mass32.max()), flush=True)
    assert np.all(mass64 <= 1e-13) and np.all(mass32 <= 1e-5)
    assert e64 <= 1e-5 and e32 <= 1e-4
    (n1, g1), _, (n3, g3) = growth
    assert g1 <= 2.5e-5 and g3 <= 1e-4 and g3 / g1 <= n3 / n1


@pytest.mark.parametrize("prec", [dg.DG_FP64, dg.DG_FP32])
def test_c3_full_size_first_steps(prec):
    m, prob, p = W.vortex3_c3()
    assert m.num_elem == 524288
    o = orc.Oracle.from_mesh(m, p, prob)
    u0 = o.project_ic(orc.ic_array(prob, m.period))
    ref, _ = o.run_steps(u0, 2, 0.3)
    s = dg.Solver(m, prob, p, prec)
    s.set_kernel(2)
    assert s.launch_info()["walk"] >= 2
    s.set_state(u0)
    s.run_steps(2, 0.3)
    err = orc.rel_l2(s.get_state(), ref)
    print("C3 2 steps prec=%d rel L2 vs oracle %.3e" % (prec, err), flush=True)
    assert err <= (1e-12 if prec == dg.DG_FP64 else 1e-5)


def test_c3_200_steps_fp32_vs_fp64():
    m, prob, p = W.vortex3_c3()
    s64 = dg.Solver(m, prob, p, dg.DG_FP64)
    s64.project_ic()
    m0 = np.array(s64.audit()["mass"])
    s32 = dg.Solver(m, prob, p, dg.DG_FP32)
    s32.set_kernel(2)
    s32.project_ic()
    s32.run_steps(200, 0.3)
    t = s32.audit()["t"]
    n64, t64 = s64.integrate(t, 0.3)
    err = orc.rel_l2(s32.get_state(), s64.get_state())
    d64 = np.abs(np.array(s64.audit()["mass"]) - m0) / np.abs(m0)
    d32 = np.abs(np.array(s32.audit()["mass"]) - m0) / np.abs(m0)
    print("C3 200 steps (t = %.6f; FP64 %d steps): FP32 vs FP64 rel L2 %.3e, mass drift FP64 %.1e FP32 %.1e"
          % (t, n64, err, d64.max(), d32.max()), flush=True)
    # FP64's own CFL steps to the same time: 200, or 201 with a last short
    # step when its wave speeds round the other way
    assert n64 in (200, 201) and t64 == t
    assert err <= 1e-5
    assert np.all(d64 <= 1e-13) and np.all(d32 <= 2e-6)


@pytest.mark.parametrize("prec", [dg.DG_FP32, dg.DG_FP64])
def test_c4_full_batch_equals_single_solves(prec):
    cases = W.ragged_batch_c4()
    assert len(cases) == 64
    b = dg.Batch(prec)
    for m, prob, p in cases:
        b.add(m, prob, p, 1e30)
    b.project_ic()
    b.run_steps(3, 0.3)
    order = sorted(range(64), key=lambda k: cases[k][0].num_elem)
    sample = [order[i] for i in np.linspace(0, 63, 8).astype(int)]
    for k in sample:
        m, prob, p = cases[k]
        s = dg.Solver(m, prob, p, prec)
        s.set_kernel(2 if p >= 3 else 1)
        s.project_ic()
        s.run_steps(3, 0.3)
        got, want = b.get_state(k), s.get_state()
        print("C4 slot %d: %d triangles p=%d N=%d max rel diff %.3e"
              % (k, m.num_elem, p, prob.physics, np.abs(got - want).max() / np.abs(want).max()), flush=True)
        if p >= 3:
            assert np.array_equal(got, want), k
        else:
            assert np.abs(got - want).max() <= (1e-6 if prec == dg.DG_FP32 else 1e-14) * np.abs(want).max(), k
        s.close()
