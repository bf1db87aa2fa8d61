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
