"""The C ABI (include/dgmi.h) and the host-side parts behind it, on the CPU:
the library loads and exports every declared entry point; the mesh
generator, the .pmsh reader/writer and the geometry precompute; argument and
state errors that are reported before any device work.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import dgcuda_amd as dg
import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def header_functions():
    text = open(os.path.join(ROOT, "include", "dgmi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set(re.findall(r"\b(dg_[a-z0-9_]+)\s*\(", text))
    return names


def test_header_entry_points_exported():
    names = header_functions()
    assert len(names) >= 45
    lib = dg.lib()
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(shutil.which("nm") is None, reason="nm not available")
def test_every_exported_entry_point_is_declared():
    out = subprocess.run(["nm", "-D", "--defined-only", dg.LIB_PATH], capture_output=True, text=True, check=True)
    exported = {ln.split()[-1] for ln in out.stdout.splitlines() if re.search(r" T dg_", ln)}
    assert exported == header_functions()


def test_version_and_last_error():
    assert dg.lib().dg_version().decode()
    with pytest.raises(dg.DGError) as e:
        dg.Mesh.periodic_jittered(0, 4, 0.0, 1.0, 0.0, 1.0)
    assert e.value.code == dg.STATUS_CODES["DG_ERR_ARG"]
    assert dg.lib().dg_last_error().decode()


# ---------------------------------------------------------------- mesh generator


def cell_of(V, e):
    return V[:, e].reshape(3, 2)


def test_generator_counts_orientation_and_sides():
    nx, ny, L = 5, 3, (2.0, 1.5)
    m = dg.Mesh.periodic_jittered(nx, ny, 0.0, L[0], 0.0, L[1], 0.0, 1)
    V, es, sd, si = m.arrays()
    K, S = V.shape[1], sd.shape[1]
    assert (K, S) == (2 * nx * ny, 3 * nx * ny) and m.num_boundary_sides == 0
    J = (V[2] - V[0]) * (V[5] - V[1]) - (V[4] - V[0]) * (V[3] - V[1])
    hx, hy = L[0] / nx, L[1] / ny
    np.testing.assert_allclose(J, hx * hy, rtol=1e-13)  # CCW, half a cell each
    # lattice vertices (no jitter)
    np.testing.assert_allclose(np.round(V[0::2] / hx) * hx, V[0::2], atol=1e-13)
    np.testing.assert_allclose(np.round(V[1::2] / hy) * hy, V[1::2], atol=1e-13)
    le, re_, ls, rs = si
    assert np.all(le >= 0) and np.all(re_ >= 0)
    for s in range(S):
        for e, f, rev in ((le[s], ls[s], False), (re_[s], rs[s], True)):
            assert es[f, e] == s
            v = cell_of(V, e)
            a, b = v[f], v[(f + 1) % 3]
            p1, p2 = sd[0:2, s], sd[2:4, s]
            if rev:
                p1, p2 = p2, p1
            # the same segment up to a periodic image
            d1, d2 = a - p1, b - p2
            assert np.allclose(d1, d2, atol=1e-12)
            assert np.allclose(np.round(d1 / L) * L, d1, atol=1e-12)
        if not np.allclose(cell_of(V, le[s])[ls[s]], sd[0:2, s], atol=1e-12):
            pytest.fail("side %d: left element's face does not start at (sx1, sy1)" % s)


def test_generator_jitter_deterministic_and_bounded():
    n, L, jit = 8, 10.0, 0.1
    a = dg.Mesh.periodic_jittered(n, n, 0.0, L, 0.0, L, jit, 5).arrays()
    b = dg.Mesh.periodic_jittered(n, n, 0.0, L, 0.0, L, jit, 5).arrays()
    c = dg.Mesh.periodic_jittered(n, n, 0.0, L, 0.0, L, jit, 6).arrays()
    z = dg.Mesh.periodic_jittered(n, n, 0.0, L, 0.0, L, 0.0, 5).arrays()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert not np.array_equal(a[0], c[0])
    h = L / n
    d = a[0] - z[0]
    assert np.abs(d).max() <= jit * h * (1 + 1e-12)
    assert np.abs(d).max() > 0.5 * jit * h
    # vertices on x = 0 or y = 0 (the periodic boundary) stay on the lattice
    on_b = (np.abs(z[0][0::2]) < 1e-12) | (np.abs(z[0][1::2]) < 1e-12)
    assert np.all(d[0::2][on_b] == 0) and np.all(d[1::2][on_b] == 0)
    # connectivity does not depend on the jitter
    assert np.array_equal(a[1], z[1]) and np.array_equal(a[3], z[3])


def test_generator_block_order():
    """Consecutive groups of 64 elements are 8 x 4 cell blocks when the grid
    divides evenly (the matrix-core tile of the solver)."""
    n, L = 64, 64.0
    V = dg.Mesh.periodic_jittered(n, n, 0.0, L, 0.0, L, 0.0, 1).arrays()[0]
    cx = (V[0] + V[2] + V[4]) / 3
    cy = (V[1] + V[3] + V[5]) / 3
    for t in range(V.shape[1] // 64):
        sl = slice(64 * t, 64 * t + 64)
        ix, iy = np.floor(cx[sl]), np.floor(cy[sl])
        assert len(set(zip(ix, iy))) == 32
        assert ix.max() - ix.min() == 7 and iy.max() - iy.min() == 3


def test_pmsh_roundtrip(tmp_path):
    m = dg.Mesh.periodic_jittered(6, 5, -1.0, 1.0, 0.0, 3.0, 0.2, 9)
    path = tmp_path / "m.pmsh"
    m.write_pmsh(path)
    r = dg.Mesh.read_pmsh(path)
    for x, y in zip(m.arrays(), r.arrays()):
        assert np.array_equal(x, y)


def test_read_reference_boundary_mesh():
    """The reference's supersonic-vortex mesh (conservation dialect, with
    boundary sides): read and cross-referenced; the solver accepts it and,
    with no GPU here, fails with DG_ERR_HIP (there is no CPU fallback)."""
    m = dg.Mesh.read_pmsh(os.path.join(GOLD, "sv1.pmsh"))
    assert m.num_elem > 0 and m.num_boundary_sides > 0
    V, es, sd, si = m.arrays()
    assert np.sum(si[1] < 0) == m.num_boundary_sides
    J = (V[2] - V[0]) * (V[5] - V[1]) - (V[4] - V[0]) * (V[3] - V[1])
    assert np.all(J != 0)
    with pytest.raises(dg.DGError) as e:
        dg.Solver(m, dg.supersonic_vortex(), 1)
    assert e.value.code == dg.STATUS_CODES["DG_ERR_HIP"]
    # boundary sides: p <= 3 on the lane kernel, p <= 4 on the matrix-core
    # kernels (Euler and advection); p = 5 refused before any device work
    for prob in (dg.supersonic_vortex(), dg.advection_gaussian()):
        with pytest.raises(dg.DGError) as e:
            dg.Solver(m, prob, 4)
        assert e.value.code == dg.STATUS_CODES["DG_ERR_HIP"]
        with pytest.raises(dg.DGError) as e:
            dg.Solver(m, prob, 5)
        assert e.value.code == dg.STATUS_CODES["DG_ERR_UNSUPPORTED"]


def test_read_errors(tmp_path):
    with pytest.raises(dg.DGError) as e:
        dg.Mesh.read_pmsh(tmp_path / "missing.pmsh")
    assert e.value.code == dg.STATUS_CODES["DG_ERR_IO"]
    bad = tmp_path / "bad.pmsh"
    bad.write_text("3\n0 0 1 0\n")
    with pytest.raises(dg.DGError) as e:
        dg.Mesh.read_pmsh(bad)
    assert e.value.code == dg.STATUS_CODES["DG_ERR_PARSE"]


def test_from_arrays_rejects_inconsistent_sides():
    V, es, sd, si = dg.Mesh.periodic_jittered(3, 3, 0.0, 1.0, 0.0, 1.0, 0.0, 1).arrays()
    si2 = si.copy()
    si2[0, 0] = V.shape[1] + 5  # left element out of range
    with pytest.raises(dg.DGError):
        dg.Mesh.from_arrays(V, es, sd, si2, period=(1.0, 1.0))
    es2 = es.copy()
    es2[0, 0] = es[1, 0]  # element faces no longer match the side list
    with pytest.raises(dg.DGError):
        dg.Mesh.from_arrays(V, es2, sd, si, period=(1.0, 1.0))


# ---------------------------------------------------------------- geometry


@pytest.mark.parametrize("quirks", [0, 1, 2, 3])
def test_geometry_bitwise_vs_oracle(quirks):
    m = dg.Mesh.periodic_jittered(7, 6, 0.0, 3.0, -1.0, 2.0, 0.15, 4)
    ge, gs, r = m.geometry(quirks)
    V, es, sd, si = m.arrays()
    o = orc.Oracle(V, es, sd, si, 2, 4, quirks=quirks)
    oge, ogs, orr = o.geometry()
    assert np.array_equal(ge, oge) and np.array_equal(gs, ogs) and r == orr
    # unit normals, outward from the left element
    assert np.allclose(np.hypot(gs[1], gs[2]), 1.0, atol=1e-6 if quirks & 2 else 1e-14)


# ---------------------------------------------------------------- solver errors before device work


@pytest.mark.parametrize("kw,code", [
    (dict(p=7), "DG_ERR_ARG"),
    (dict(precision=5), "DG_ERR_ARG"),
    (dict(prob=dg.problem(3, dg.DG_IC_CONSTANT)), "DG_ERR_ARG"),
    (dict(prob=dg.problem(dg.DG_PHYS_ADVECTION, dg.DG_IC_VORTEX, vortices=((1.0, 0.0, 0.0),))), "DG_ERR_ARG"),
    (dict(prob=dg.problem(dg.DG_PHYS_EULER, dg.DG_IC_VORTEX)), "DG_ERR_ARG"),  # no vortex given
    (dict(prob=dg.euler_vortex(gamma=1.0)), "DG_ERR_ARG"),
    (dict(prob=dg.euler_vortex(quirks=16)), "DG_ERR_ARG"),
    (dict(prob=dg.supersonic_vortex(mach=0.0)), "DG_ERR_ARG"),
    (dict(prob=dg.problem(dg.DG_PHYS_ADVECTION, dg.DG_IC_SUPERSONIC)), "DG_ERR_ARG"),
])
def test_solver_argument_errors(kw, code):
    m = dg.Mesh.periodic_jittered(4, 4, 0.0, 1.0, 0.0, 1.0, 0.0, 1)
    args = dict(prob=dg.euler_vortex(), p=2, precision=dg.DG_FP64)
    args.update(kw)
    with pytest.raises(dg.DGError) as e:
        dg.Solver(m, args["prob"], args["p"], args["precision"])
    assert e.value.code == dg.STATUS_CODES[code]


def test_without_a_gpu_the_solver_fails_loudly():
    if dg.device_count() > 0:
        pytest.skip("a GPU is present")
    m = dg.Mesh.periodic_jittered(4, 4, 0.0, 1.0, 0.0, 1.0, 0.0, 1)
    with pytest.raises(dg.DGError) as e:
        dg.Solver(m, dg.euler_vortex(), 2)
    assert e.value.code == dg.STATUS_CODES["DG_ERR_HIP"]
