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
  the reference left empty) is refused with a parse error.
"""
import os

import numpy as np
import pytest

import dgcuda_amd as dg
from helpers import GOLD_DIR

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


def write_msh(path, nodes, lines, tris):
    with open(path, "w") as f:
        f.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%d\n" % len(nodes))
        for i, (x, y) in enumerate(nodes):
            f.write("%d %r %r 0\n" % (i + 1, x, y))
        f.write("$EndNodes\n$Elements\n%d\n" % (len(lines) + len(tris)))
        k = 1
        for tag, a, b in lines:
            f.write("%d 1 2 %d 1 %d %d\n" % (k, tag, a + 1, b + 1))
            k += 1
        for a, b, c in tris:
            f.write("%d 2 2 100 6 %d %d %d\n" % (k, a + 1, b + 1, c + 1))
            k += 1
        f.write("$EndElements\n")


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
