"""Output and diagnostics on the host (no GPU): the reference's gmsh view
writer and an L2 error against the problem's initial state.

Pins:
* dg_write_msh given the vertex values of the reference's own dump
  (euler/output/U0.msh on sv1refined1, kept as tests/golden/euler_U0.msh)
  and that mesh writes the dump back byte for byte;
* dg_l2_error of the zero state is the L2 norm of U0, which for a constant
  state is |U0| sqrt(area) (area = sum of J/2); the projection of a
  constant is exact (error at rounding level); and the projection error of
  the supersonic vortex on the three shipped meshes falls as h^(p+1).
"""
import math
import os
import re

import numpy as np
import pytest

import dgcuda_amd as dg
import oracle as orc
from helpers import GOLD_DIR, golden_mesh


def dump_values(path):
    text = open(path).read()
    return np.array([[float(t) for t in m.group(1).split(",")] for m in re.finditer(r"\) \{([^}]*)\};", text)])


def test_write_msh_reproduces_reference_dump(tmp_path):
    ref = os.path.join(GOLD_DIR, "euler_U0.msh")
    m = golden_mesh("sv1refined1")
    uv = dump_values(ref)
    assert uv.shape == (m.num_elem, 3)
    out = tmp_path / "U0.msh"
    m.write_msh(0, uv, out)
    assert open(out, "rb").read() == open(ref, "rb").read()


def test_l2_error_of_constants():
    m = golden_mesh("sv1")
    U = (1.3, 0.2, -0.4, 3.1)
    prob = dg.euler_constant(U)
    ge, _, _ = m.geometry(0)
    area = float(ge[0].sum()) / 2
    n_p = 3
    zero = np.zeros((4, n_p, m.num_elem))
    err = m.l2_error(prob, 1, zero)
    assert np.allclose(err, np.abs(U) * math.sqrt(area), rtol=1e-12)
    o = orc.Oracle.from_mesh(m, 1, prob)
    u = o.project_ic(orc.ic_array(prob, m.period))
    assert (m.l2_error(prob, 1, u) <= 1e-13 * np.abs(U).max() * math.sqrt(area)).all()


@pytest.mark.parametrize("p", [1, 2, 3])
def test_l2_error_converges_at_order_p_plus_1(p):
    prob = dg.supersonic_vortex()
    errs = []
    for name in ("sv1", "sv1refined", "sv1refined1"):
        m = golden_mesh(name)
        o = orc.Oracle.from_mesh(m, p, prob)
        u = o.project_ic(orc.ic_array(prob, m.period))
        errs.append(m.l2_error(prob, p, u)[0])
    rates = [math.log2(a / b) for a, b in zip(errs, errs[1:])]
    print("p=%d density L2 errors %s rates %s" % (p, ["%.3e" % e for e in errs], ["%.2f" % r for r in rates]))
    for r in rates:
        assert r >= 0.85 * (p + 1), rates


# ---------------------------------------------------------------- named views

CPU_OUT = os.path.join(GOLD_DIR, "cpu_output")
VIEWS = [("uniform_rho.out", "Density ", 1), ("uniform_E.out", "E ", 1), ("p.out", "E ", 1),
         ("uniform_u.out", "u ", 2)]


def view_values(path, ncomp):
    v = dump_values(path)
    if ncomp == 1:
        return v  # (K, 3)
    return v.reshape(-1, 3, 3)[:, :, :2]  # (u, v, 0) per vertex


@pytest.mark.parametrize("fname,name,ncomp", VIEWS)
def test_named_views_reproduce_the_reference_cpu_output(tmp_path, fname, name, ncomp):
    """The reference's CPU Euler driver wrote these views on sv1 (the density,
    velocity, energy and, under the view name "E ", pressure files of
    euler/cpu/output); given their vertex values and the mesh, the writer
    reproduces each file byte for byte."""
    path = os.path.join(CPU_OUT, fname)
    vals = view_values(path, ncomp)
    m = golden_mesh("sv1")
    assert vals.shape[0] == m.num_elem
    out = tmp_path / fname
    m.write_msh_view(name, vals, out)
    assert open(out, "rb").read() == open(path, "rb").read()


def test_reference_pressure_view_is_the_pressure_of_its_state():
    """p.out against (gamma - 1)(E - (u^2 + v^2)/2 rho) of the same run's
    density, velocity and energy views, vertex by vertex (6 printed digits)."""
    rho = view_values(os.path.join(CPU_OUT, "uniform_rho.out"), 1)
    E = view_values(os.path.join(CPU_OUT, "uniform_E.out"), 1)
    uv = view_values(os.path.join(CPU_OUT, "uniform_u.out"), 2)
    p = view_values(os.path.join(CPU_OUT, "p.out"), 1)
    want = 0.4 * (E - (uv[..., 0] ** 2 + uv[..., 1] ** 2) / 2 * rho)
    assert np.abs(p - want).max() <= 2e-5 * np.abs(p).max()
