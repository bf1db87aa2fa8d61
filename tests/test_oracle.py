"""Pins the FP64 oracle (oracle/dg_oracle.cpp) before anything is compared
against it.  CPU only.

  project_ic   the reference's golden vertex dumps src/2d/tests/t0/*.out
               ((x - y)^alpha projected at p = alpha..5 on three meshes)
  rhs          an independent physical-space evaluation of the weak form
               (tests/weakform.py) on tiny periodic meshes, advection and
               Euler, p = 0..3; free-stream preservation; discrete
               conservation; LLF flux equal and opposite on both sides
  rk_step      the Butcher tableaux in closed form: for the linear advection
               operator L one RK4 step is the degree-4 Taylor polynomial of
               exp(dt L), RK2 (midpoint) the degree-2 one
  cfl_dt       dt = cfl * min_r / max lambda / (2p + 1) with min_r the
               inscribed diameter of a right isosceles triangle and lambda
               from the cell means
  integrate    time advanced before the stages, last step clamped to t_end
  end to end   p + 1 convergence on the isentropic vortex
"""
import glob
import json
import math
import os
import re

import numpy as np
import pytest

import dgcuda_amd as dg
import oracle as orc
import weakform as wf
from helpers import VORTEX_C2, mesh, perturbed_state, problem_for

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def ref_tables():
    with open(os.path.join(GOLD, "ref_tables.json")) as f:
        return json.load(f)


# ---------------------------------------------------------------- project_ic


def read_vertex_mesh(path):
    """Six-field .pmsh (advection prototype): K, then V1x V1y V2x V2y V3x V3y.
    Like the prototype's reader (src/2d/2dadvec/2dadvec.cu:146-154), a
    clockwise element has V1 and V2 swapped."""
    with open(path) as f:
        K = int(f.readline())
        V = np.array([[float(x) for x in f.readline().split()] for _ in range(K)])
    J = (V[:, 2] - V[:, 0]) * (V[:, 5] - V[:, 1]) - (V[:, 4] - V[:, 0]) * (V[:, 3] - V[:, 1])
    cw = J < 0
    V[cw, 0:2], V[cw, 2:4] = V[cw, 2:4].copy(), V[cw, 0:2].copy()
    return V.T.copy()  # (6, K)


def read_t0(path):
    xy, vals = [], []
    for line in open(path):
        m = re.match(r"ST \(([^)]*)\) \{([^}]*)\}", line)
        if m:
            c = [float(x) for x in m.group(1).split(",")]
            xy.append([c[0], c[1], c[3], c[4], c[6], c[7]])
            vals.append([float(x) for x in m.group(2).split(",")])
    return np.array(xy), np.array(vals)


T0 = sorted(glob.glob(os.path.join(GOLD, "t0", "alpha*_*.out")))


def test_t0_fixture_count():
    # alpha = 0..3, p = alpha..5, three meshes
    assert len(T0) == 3 * sum(6 - a for a in range(4))


@pytest.mark.parametrize("path", T0, ids=[os.path.basename(p)[:-4] for p in T0])
def test_t0_golden_projection(path):
    a, mname, p = re.match(r"alpha(\d)_([a-z]+)(\d)\.out", os.path.basename(path)).groups()
    a, p = int(a), int(p)
    V = read_vertex_mesh(os.path.join(GOLD, "t0", mname + ".pmsh"))
    K = V.shape[1]
    o = orc.Oracle(V, np.zeros((3, K), np.int32), np.zeros((4, 0)), np.zeros((4, 0), np.int32), p, 1)
    u = o.project_ic(orc.ic_poly(a))
    got = o.eval_vertices(u, 0)
    xy, want = read_t0(path)
    assert xy.shape[0] == K
    np.testing.assert_allclose(xy, V.T, atol=1e-6)  # same element and vertex order
    scale = max(1.0, float(np.abs(want).max()))
    # the dumps carry 6 significant digits of a single-precision run
    # (src/2d/test_all.cu), hence the tolerance
    assert np.all(np.abs(got - want) <= 5e-6 * np.abs(want) + 2e-6 * scale), (np.abs(got - want).max())


def test_projection_reproduces_polynomials_exactly():
    """(x - y)^alpha with alpha <= p lies in the space: vertex values exact up
    to the precision of the reference's tables (15-digit quadrature literals;
    the basis is orthonormal to ~1e-13 at p = 4 and ~1e-11 at p = 5)."""
    tol = {0: 1e-14, 1: 1e-13, 2: 1e-13, 3: 1e-12, 4: 5e-12, 5: 1e-10}
    m = mesh(6)
    V, es, sd, si = m.arrays()
    for p in range(0, 6):
        for a in range(0, p + 1):
            o = orc.Oracle(V, es, sd, si, p, 1)
            uv = o.eval_vertices(o.project_ic(orc.ic_poly(a)), 0)
            x = V[0::2].T  # (K, 3)
            y = V[1::2].T
            want = (x - y) ** a
            assert np.abs(uv - want).max() <= tol[p] * max(1.0, np.abs(want).max()), (p, a)


# ---------------------------------------------------------------- rhs


def weak_form(m, prob, p, u, dt=1.0):
    t = ref_tables()
    o = orc.Oracle.from_mesh(m, p, prob)
    tb = o.tables()
    V, es, sd, si = m.arrays()
    n_p = (p + 1) * (p + 2) // 2
    # vol weights of the reference rules sum to 1 on the triangle; the
    # oracle halves them (sum 1/2 = area of the reference triangle)
    # (the 6-point Gauss rule of p = 5 carries 8-digit literals: its weights
    # sum to 2 + 3e-8; both evaluations use the same rule)
    assert abs(tb["w"].sum() - 0.5) < 1e-13 and abs(tb["w1"].sum() - 2.0) < (1e-13 if p < 5 else 1e-7)
    return wf.rhs(V, es, sd, si, m.period, u, p, prob.physics, t["basis_coeff"], (tb["r1"], tb["r2"], tb["w"]),
                  (tb["s_r"], tb["w1"]), dt=dt, g=prob.gamma, adv=(prob.ax, prob.ay))[:, :n_p]


@pytest.mark.parametrize("N", [1, 4])
@pytest.mark.parametrize("p", [0, 1, 2, 3, 4, 5])
def test_rhs_matches_independent_weak_form(N, p):
    m = mesh(3, seed=5)
    prob = problem_for(N)
    o = orc.Oracle.from_mesh(m, p, prob)
    u = perturbed_state(o, prob, m.period, seed=p)
    k = o.rhs(u, 0.37)
    ref = weak_form(m, prob, p, u, 0.37)
    err = np.abs(k - ref).max() / np.abs(ref).max()
    assert err <= 1e-12, err


def test_weak_form_catches_a_planted_sign_error():
    """The comparison above is sharp: flipping the LLF dissipation sign in the
    independent evaluation moves it far from the oracle."""
    m = mesh(3, seed=5)
    prob = problem_for(4)
    o = orc.Oracle.from_mesh(m, 2, prob)
    u = perturbed_state(o, prob, m.period, seed=1)
    k = o.rhs(u)
    keep = wf.llf
    try:
        wf.llf = lambda a, b, n, phys, g, adv: keep(b, a, n, phys, g, adv) - 0 * a
        bad = weak_form(m, prob, 2, u)
    finally:
        wf.llf = keep
    assert np.abs(k - bad).max() / np.abs(k).max() > 1e-3


@pytest.mark.parametrize("N", [1, 4])
@pytest.mark.parametrize("p", [1, 3, 5])
def test_free_stream_preserved(N, p):
    m = mesh(5, seed=2)
    if N == 1:
        prob = dg.problem(dg.DG_PHYS_ADVECTION, ic=dg.DG_IC_CONSTANT, constant=(2.5, 0.0, 0.0, 0.0), a=(0.7, -0.4))
    else:
        prob = dg.problem(dg.DG_PHYS_EULER, ic=dg.DG_IC_CONSTANT, constant=(1.2, 0.3, -0.2, 2.7))
    # a constant projects onto mode 0 up to the tables' orthonormality; at
    # p = 5 the face rule is the reference's 6-point Gauss rule, whose
    # literals carry 8 digits (src/2d/quadrature.cu), which leaves a
    # free-stream residual of ~1e-7 |u|
    tol_proj = {1: 1e-14, 3: 1e-13, 5: 1e-11}[p]
    tol_rhs = {1: 1e-13, 3: 1e-12, 5: 2e-7}[p]
    o = orc.Oracle.from_mesh(m, p, prob)
    u = o.project_ic(orc.ic_array(prob, m.period))
    assert np.abs(u[:, 1:, :]).max() <= tol_proj * np.abs(u).max()
    k = o.rhs(u)
    assert np.abs(k).max() <= tol_rhs * np.abs(u).max()


@pytest.mark.parametrize("N", [1, 4])
@pytest.mark.parametrize("p", [1, 2, 4])
def test_discrete_conservation(N, p):
    """sum_K J_K * k[:, 0, K] = 0 on a periodic mesh (mode 0 is constant)."""
    m = mesh(6, seed=3)
    prob = problem_for(N)
    o = orc.Oracle.from_mesh(m, p, prob)
    u = perturbed_state(o, prob, m.period, seed=9)
    k = o.rhs(u)
    J = o.geometry()[0][0]
    tot = np.sum(J * k[:, 0, :], axis=1)
    scale = np.sum(np.abs(J * k[:, 0, :]), axis=1)
    assert np.all(np.abs(tot) <= 1e-13 * scale), tot / scale


@pytest.mark.parametrize("N", [1, 4])
def test_llf_flux_equal_and_opposite(N):
    """Per side, the mode-0 contributions to the left and right element are
    the same flux integral with opposite sign."""
    m = mesh(4, seed=4)
    prob = problem_for(N)
    o = orc.Oracle.from_mesh(m, 3, prob)
    u = perturbed_state(o, prob, m.period, seed=2)
    q, l, r = o.rhs_parts(u)
    assert np.abs(l).max() > 0
    np.testing.assert_allclose(l[:, 0, :], -r[:, 0, :], rtol=0, atol=1e-13 * np.abs(l).max())


def test_hoisted_and_plain_loops_bitwise():
    m = mesh(5, seed=8)
    for N in (1, 4):
        prob = problem_for(N)
        o = orc.Oracle.from_mesh(m, 3, prob, hoist=1)
        u = perturbed_state(o, prob, m.period, seed=4)
        k1 = o.rhs(u, 0.1)
        o.set_hoist(0)
        k0 = o.rhs(u, 0.1)
        assert np.array_equal(k0, k1)


# ---------------------------------------------------------------- time stepping


@pytest.mark.parametrize("scheme,order", [(1, 4), (2, 2)])
def test_rk_butcher_taylor(scheme, order):
    """Advection is linear: k(u) = dt L u.  One step of RK4 equals
    sum_{j<=4} (dt L)^j u / j!, RK2 midpoint the sum to j = 2."""
    m = mesh(4, seed=6)
    prob = problem_for(1)
    o = orc.Oracle.from_mesh(m, 2, prob)
    u = perturbed_state(o, prob, m.period, seed=3, amp=0.3)
    dt = 0.05
    want, term = u.copy(), u.copy()
    for j in range(1, order + 1):
        term = o.rhs(term, dt) / j
        want += term
    got = o.rk_step(u, dt, scheme)
    assert np.abs(got - want).max() <= 1e-13 * np.abs(u).max()
    # and the order is not higher: the next Taylor term is visible
    nxt = o.rhs(term, dt) / (order + 1)
    assert np.abs(got - (want + nxt)).max() > 1e3 * np.abs(got - want).max()


def test_cfl_dt_closed_form():
    n, L = 8, 10.0
    m = dg.Mesh.periodic_jittered(n, n, 0.0, L, 0.0, L, 0.0, 1)  # no jitter: right isosceles triangles
    h = L / n
    for N, p in ((1, 2), (4, 3)):
        prob = problem_for(N)
        o = orc.Oracle.from_mesh(m, p, prob)
        _, _, min_r = o.geometry()
        assert abs(min_r - h * (2 - math.sqrt(2))) <= 1e-13 * h
        u = o.project_ic(orc.ic_array(prob, m.period))
        if N == 1:
            lam = math.hypot(prob.ax, prob.ay)
        else:
            U = u[:, 0, :] * math.sqrt(2.0)  # cell means: phi_0 = sqrt(2) on the reference triangle
            pr = (prob.gamma - 1) * (U[3] - 0.5 * (U[1] ** 2 + U[2] ** 2) / U[0])
            lam = float(np.max(np.hypot(U[1], U[2]) / U[0] + np.sqrt(prob.gamma * pr / U[0])))
        want = 0.3 * min_r / lam / (2 * p + 1)
        assert abs(o.cfl_dt(u, 0.3) - want) <= 1e-13 * want


def test_integrate_clamps_last_step():
    m = mesh(4, seed=1)
    prob = problem_for(4)
    o = orc.Oracle.from_mesh(m, 1, prob)
    u0 = o.project_ic(orc.ic_array(prob, m.period))
    t_end = 0.123
    u, n, t = o.integrate(u0, t_end, 0.3)
    assert t == t_end and n >= 1
    # same number of full steps by hand, then the clamped one
    v, tt = u0.copy(), 0.0
    for _ in range(n):
        dt = o.cfl_dt(v, 0.3)
        if tt + dt > t_end:
            dt = t_end - tt
        tt += dt
        v = o.rk_step(v, dt, 1)
    assert tt == t_end
    assert np.array_equal(u, v)


def test_mass_conserved_over_a_run():
    m = mesh(8, seed=2)
    prob = dg.euler_vortex(VORTEX_C2)
    o = orc.Oracle.from_mesh(m, 2, prob)
    u0 = o.project_ic(orc.ic_array(prob, m.period))
    u, _, _ = o.integrate(u0, 0.5, 0.3)
    m0, m1 = o.mass(u0), o.mass(u)
    assert np.all(np.abs(m1 - m0) <= 1e-12 * np.abs(m0))


def vortex_error(n, p, t_end=0.5):
    """Stationary isentropic vortex (an exact steady solution): L2 error of
    the state at t_end against its own L2 projection."""
    m = dg.Mesh.periodic_jittered(n, n, 0.0, 10.0, 0.0, 10.0, 0.0, 1)
    prob = dg.euler_vortex(((5.0, 5.0, 5.0),), background=(1.0, 0.0, 0.0, 1.0))
    o = orc.Oracle.from_mesh(m, p, prob)
    u0 = o.project_ic(orc.ic_array(prob, m.period))
    u, _, _ = o.integrate(u0, t_end, 0.3)
    J = o.geometry()[0][0]
    return float(np.sqrt(np.sum(J * np.sum((u - u0)[0] ** 2, axis=0))))


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_p_plus_one_convergence(p):
    e = [vortex_error(n, p) for n in (8, 16, 32)]
    rates = [math.log2(e[i] / e[i + 1]) for i in range(2)]
    print("p=%d errors %s rates %s" % (p, e, rates))
    assert rates[-1] >= p + 1 - 0.35, rates
