"""GPU parity on meshes read from gmsh files: a periodic box paired by the
converter (every kernel, including the matrix-core ones at p >= 3) and a
channel periodic in x with inflow / outflow sides (lane kernel up to p = 3,
the matrix-core kernels' boundary variant up to p = 4), through the C ABI
against the FP64 oracle.  Bars as for the generated
meshes: FP64 within 1e-13 relative max-norm per RHS, FP32 within 1e-5."""
import pytest

import dgcuda_amd as dg
import oracle as orc
from helpers import channel_lines, generated, perturbed_state, rel_max, write_elements_msh

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [dg.DG_FP64, dg.DG_FP32])
@pytest.mark.parametrize("p", [1, 4])
def test_periodic_gmsh_rhs_parity(tmp_path, p, prec):
    print("periodic gmsh rhs p=%d prec=%d" % (p, prec), flush=True)
    path = tmp_path / "box.msh"
    write_elements_msh(path, generated(24, 16).arrays()[0])
    m = dg.Mesh.read_gmsh(path, periodic=True)
    prob = dg.euler_vortex(vortices=((1.5, 1.0, 2.0),))
    o = orc.Oracle.from_mesh(m, p, prob)
    u = perturbed_state(o, prob, m.period, seed=p, amp=1e-2)
    ref = o.rhs(u, dt=0.01)
    s = dg.Solver(m, prob, p, prec)
    for kind in ([1, 2] if p >= 3 else [1]):
        s.set_kernel(kind)
        err = rel_max(s.rhs(u, dt=0.01), ref)
        print("  kernel %d rel max err %.3e" % (kind, err))
        assert err <= (1e-13 if prec == dg.DG_FP64 else 1e-5)


@pytest.mark.parametrize("p", [1, 3, 4])
def test_channel_gmsh_rhs_parity(tmp_path, p):
    print("channel gmsh rhs p=%d" % p, flush=True)
    V = generated(12, 8).arrays()[0]
    path = tmp_path / "channel.msh"
    write_elements_msh(path, V, channel_lines(V, 0.0, 2.0))
    m = dg.Mesh.read_gmsh(path, periodic=True)
    assert m.num_boundary_sides == 24
    prob = dg.euler_vortex(vortices=((1.5, 1.0, 2.0),))
    o = orc.Oracle.from_mesh(m, p, prob)
    u = perturbed_state(o, prob, m.period, seed=p, amp=1e-2)
    ref = o.rhs(u, dt=0.01)
    s = dg.Solver(m, prob, p, dg.DG_FP64)
    for kind in ([1, 2] if p <= 3 else [2]):
        s.set_kernel(kind)
        err = rel_max(s.rhs(u, dt=0.01), ref)
        print("  kernel %d rel max err %.3e" % (kind, err))
        assert err <= 1e-13
