"""The gmsh (.msh 2.2) to .pmsh converter, restated from the reference's
genmesh.py, on the CPU.

Pins:
* every shipped .msh / .pmsh pair the converter made (supersonic sv1 and
  sv1refined, cylinder and cylinder1) converts to the shipped .pmsh text
  exactly, printed as genmesh.py prints it (%lf), with and without the
  converter's quirks (neither bug fires on these meshes);
* a mesh with a clockwise triangle: the clean mode reorders that element's
  own vertices (every Jacobian positive, the side structure consistent),
  the quirk mode swaps the coordinates of the two shared nodes, moving the
  neighbour's vertices as genmesh.py does;
* an open side without a boundary tag is an error in the clean mode and
  gets the converter's default right element 0 in the quirk mode;
* the gmsh 1.0 file among the reference's meshes (cyl16.msh, whose .pmsh
  the reference left empty) is refused with a parse error;
* periodic pairing (not part of genmesh.py): the elements of a generated
  periodic mesh written as a .msh without its wrap and read back paired
  give the generated mesh's period, side count 3K/2 and, for seeded states,
  the oracle's RHS of the generated mesh (a side paired with the wrong
  partner changes it at O(1)); a channel periodic in x with tagged outflow
  (top) and inflow (bottom) sides keeps them as boundary sides in front,
  outflow first, and preserves a constant state (the inflow / outflow ghost
  is the problem's own state; the reference's reflecting ghost uses the
  radial normal of the supersonic vortex's arcs, so a flat wall is not a
  test of it); unmatched edges and the quirk mode are refused.
"""
import os

import numpy as np
import pytest

import dgcuda_amd as dg
import oracle as orc
from helpers import GOLD_DIR, channel_lines, generated, perturbed_state, rel_max, write_elements_msh, write_msh

GM = os.path.join(GOLD_DIR, "gmsh")


def lf_text(m):
    V, es, sd, si = m.arrays()
    out = ["%d" % m.num_elem]
    for i in range(m.num_elem):
        out.append("%f %f %f %f %f %f %d %d %d" % (tuple(V[:, i]) + tuple(es[:, i])))
    out.append("%d" % m.num_sides)
    for i in range(m.num_sides):
        out.append("%f %f %f %f %d %d %d %d" % (tuple(sd[:, i]) + tuple(si[:, i])))
    return "\n".join(out) + "\n"


@pytest.mark.parametrize("quirks", [True, False])
@pytest.mark.parametrize("name", ["sv1", "sv1refined", "cylinder", "cylinder1"])
def test_shipped_pairs_convert_exactly(name, quirks):
    m = dg.Mesh.read_gmsh(os.path.join(GM, name + ".msh"), genmesh_quirks=quirks)
    assert lf_text(m) == open(os.path.join(GM, name + ".pmsh")).read()


# unit square split along a diagonal; the second triangle is clockwise
NODES = [(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)]
WALLS = [(10000, 0, 1), (20000, 1, 2), (30000, 2, 3), (10000, 3, 0)]


def jacobians(m):
    V = m.arrays()[0]
    return (V[2] - V[0]) * (V[5] - V[1]) - (V[4] - V[0]) * (V[3] - V[1])


def test_clockwise_element(tmp_path):
    path = tmp_path / "cw.msh"
    write_msh(path, NODES, WALLS, [(0, 1, 2), (0, 3, 2)])
    clean = dg.Mesh.read_gmsh(path)
    assert (jacobians(clean) > 0).all()
    V, es, sd, si = clean.arrays()
    assert sorted(si[1][si[1] < 0].tolist()) == [-3, -2, -1, -1]
    quirk = dg.Mesh.read_gmsh(path, genmesh_quirks=True)
    Vq = quirk.arrays()[0]
    # nodes 0 and 3 swapped their coordinates: the first triangle's first
    # vertex, shared with the clockwise one, moved to (0, 1)
    assert tuple(Vq[0:2, 0]) == (0.0, 1.0)
    assert tuple(V[0:2, 0]) == (0.0, 0.0)


def test_untagged_open_side(tmp_path):
    path = tmp_path / "open.msh"
    write_msh(path, NODES, WALLS[:3], [(0, 1, 2), (0, 2, 3)])
    with pytest.raises(dg.DGError) as e:
        dg.Mesh.read_gmsh(path)
    assert e.value.code == dg.STATUS_CODES["DG_ERR_PARSE"]
    m = dg.Mesh.read_gmsh(path, genmesh_quirks=True)
    si = m.arrays()[3]
    assert (si[1] == 0).sum() >= 1  # genmesh.py's default right element


def test_gmsh_1_format_refused(tmp_path):
    path = tmp_path / "old.msh"
    path.write_text("$NOD\n3\n1 0 0 0\n2 1 0 0\n3 0 1 0\n$ENDNOD\n$ELM\n1\n1 2 100 6 3 1 2 3\n$ENDELM\n")
    with pytest.raises(dg.DGError) as e:
        dg.Mesh.read_gmsh(path)
    assert e.value.code == dg.STATUS_CODES["DG_ERR_PARSE"]


@pytest.mark.parametrize("p", [1, 3])
@pytest.mark.parametrize("physics", ["advection", "euler"])
def test_periodic_pairing_matches_generated_mesh(tmp_path, physics, p):
    g = generated()
    V = g.arrays()[0]
    path = tmp_path / "box.msh"
    write_elements_msh(path, V)
    m = dg.Mesh.read_gmsh(path, periodic=True)
    assert m.period == (3.0, 2.0) == g.period
    assert m.num_elem == g.num_elem and m.num_sides == 3 * m.num_elem // 2 and m.num_boundary_sides == 0
    assert (m.arrays()[0] == V).all()
    if physics == "advection":
        prob = dg.advection_gaussian(a=(1.0, 0.5), center=(1.5, 1.0), k=4.0)
    else:
        prob = dg.euler_vortex(vortices=((1.5, 1.0, 2.0),))
    og = orc.Oracle.from_mesh(g, p, prob)
    om = orc.Oracle.from_mesh(m, p, prob)
    u = perturbed_state(og, prob, g.period, seed=p, amp=1e-2)
    assert rel_max(om.project_ic(orc.ic_array(prob, m.period)), og.project_ic(orc.ic_array(prob, g.period))) == 0
    err = rel_max(om.rhs(u, dt=0.01), og.rhs(u, dt=0.01))
    print("periodic pairing %s p=%d rel max %.3e" % (physics, p, err))
    assert err <= 1e-13
    # planted: two wrapped sides exchange partners; the RHS moves at O(1)
    V2, es2, sd2, si2 = m.arrays()
    wrap = np.nonzero((sd2[0] == sd2[2]) & np.isin(sd2[0], (0.0, 3.0)))[0]
    a, b = wrap[0], wrap[len(wrap) // 2]
    ra, fa, rb, fb = si2[1][a], si2[3][a], si2[1][b], si2[3][b]
    si2[1][a], si2[3][a], si2[1][b], si2[3][b] = rb, fb, ra, fa
    es2[fa][ra], es2[fb][rb] = b, a
    bad = dg.Mesh.from_arrays(V2, es2, sd2, si2, m.period)
    assert rel_max(orc.Oracle.from_mesh(bad, p, prob).rhs(u, dt=0.01), og.rhs(u, dt=0.01)) > 1e-3


def test_periodic_channel(tmp_path):
    nx, ny = 6, 5
    V = generated(nx, ny).arrays()[0]
    path = tmp_path / "channel.msh"
    write_elements_msh(path, V, channel_lines(V, 0.0, 2.0))
    m = dg.Mesh.read_gmsh(path, periodic=True)
    assert m.period == (3.0, 0.0)
    assert m.num_boundary_sides == 2 * nx and m.num_sides == (3 * m.num_elem + 2 * nx) // 2
    si = m.arrays()[3]
    assert (si[1][:nx] == -2).all() and (si[1][nx:2 * nx] == -3).all() and (si[1][2 * nx:] >= 0).all()
    prob = dg.euler_constant((1.0, 0.5, 0.3, 2.5))
    o = orc.Oracle.from_mesh(m, 2, prob)
    u = o.project_ic(orc.ic_array(prob, m.period))
    # rounding level: |U| ~ 2.5 times M^-1 ~ 1/J ~ 30 on these elements
    assert np.abs(o.rhs(u, dt=1.0)).max() <= 1e-12


def test_periodic_pairing_refusals(tmp_path):
    V = generated().arrays()[0].copy()
    path = tmp_path / "box.msh"
    write_elements_msh(path, V)
    with pytest.raises(dg.DGError) as e:
        dg.Mesh.read_gmsh(path, genmesh_quirks=True, periodic=True)
    assert e.value.code == dg.STATUS_CODES["DG_ERR_ARG"]
    import ctypes
    h = ctypes.c_void_p()
    assert dg.lib().dg_mesh_read_gmsh(str(path).encode(), 4, ctypes.byref(h)) == dg.STATUS_CODES["DG_ERR_ARG"]
    # one node of the right edge moved along it: its two sides lose their partners
    right = V[0::2] == 3.0
    ys = np.unique(V[1::2][right])
    y = ys[len(ys) // 2]
    V[1::2][right & (V[1::2] == y)] = y + 0.01
    write_elements_msh(path, V)
    with pytest.raises(dg.DGError) as e:
        dg.Mesh.read_gmsh(path, periodic=True)
    assert e.value.code == dg.STATUS_CODES["DG_ERR_PARSE"]
    assert "partner" in str(e.value)
