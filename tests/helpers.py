"""Shared helpers of the test suite (no GPU needed to import)."""
import os

import numpy as np

import dgcuda_amd as dg
import oracle as orc

VORTEX_C2 = ((5.0, 5.0, 5.0),)


def mesh(n, lo=0.0, hi=10.0, jitter=0.1, seed=1):
    return dg.Mesh.periodic_jittered(n, n, lo, hi, lo, hi, jitter, seed)


def problem_for(N, quirks=0):
    if N == 1:
        return dg.advection_gaussian(a=(1.0, 1.0), center=(5.0, 5.0), k=0.5, quirks=quirks)
    return dg.euler_vortex(VORTEX_C2, quirks=quirks)


def perturbed_state(o, prob, period, seed=0, amp=1e-2):
    """Projected initial condition plus a seeded perturbation of every mode,
    its amplitude shrinking with the number of modes above p = 2 so point
    values stay within a few percent (keeps density and pressure positive
    for the Euler vortex at every order)."""
    u = o.project_ic(orc.ic_array(prob, period))
    rng = np.random.default_rng(seed)
    scale = np.abs(u).max(axis=(1, 2), keepdims=True) + 1e-3
    amp = amp * min(1.0, 6.0 / u.shape[1])
    return u + amp * scale * rng.standard_normal(u.shape)


def shuffled_mesh(n, seed=7, lo=0.0, hi=10.0):
    """A jittered periodic mesh with elements and sides renumbered at random
    (so the solver's Morton reordering and irregular tile halos are used)."""
    m = mesh(n, lo, hi)
    V, es, sd, si = m.arrays()
    K, S = V.shape[1], sd.shape[1]
    rng = np.random.default_rng(seed)
    pe, ps = rng.permutation(K), rng.permutation(S)   # new -> old
    ie, is_ = np.argsort(pe), np.argsort(ps)          # old -> new
    V2 = V[:, pe]
    es2 = is_[es[:, pe]].astype(np.int32)
    sd2 = sd[:, ps]
    si2 = si[:, ps].copy()
    si2[0] = ie[si2[0]]
    si2[1] = ie[si2[1]]
    return dg.Mesh.from_arrays(V2, es2, sd2, si2, period=(hi - lo, hi - lo))


def rel_max(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


GOLD_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_mesh(name):
    """A shipped reference mesh kept under tests/golden."""
    return dg.Mesh.read_pmsh(os.path.join(GOLD_DIR, name + ".pmsh"))


def relabelled(m):
    """The same mesh with every element's vertices relabelled (V2, V3, V1):
    local face f becomes face (f + 2) % 3 (boundary sides on face 2 move to
    face 1)."""
    V, es, sd, si = m.arrays()
    si = si.copy()
    Vn = np.vstack([V[2], V[3], V[4], V[5], V[0], V[1]])
    esn = np.vstack([es[1], es[2], es[0]])
    si[2] = (si[2] + 2) % 3
    inner = si[1] >= 0
    si[3][inner] = (si[3][inner] + 2) % 3
    return dg.Mesh.from_arrays(Vn, esn, sd, si, m.period)


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


def write_elements_msh(path, V, lines=()):
    """The elements of V[6, K] as gmsh triangles in order, vertex order kept,
    nodes shared by exact coordinates; lines are (tag, (x, y), (x, y))."""
    ids, nodes = {}, []

    def node(x, y):
        x, y = float(x), float(y)
        if (x, y) not in ids:
            ids[(x, y)] = len(nodes)
            nodes.append((x, y))
        return ids[(x, y)]

    tris = [tuple(node(V[2 * c, e], V[2 * c + 1, e]) for c in range(3)) for e in range(V.shape[1])]
    segs = [(tag, node(*a), node(*b)) for tag, a, b in lines]
    write_msh(path, nodes, segs, tris)


def generated(nx=6, ny=5):
    return dg.Mesh.periodic_jittered(nx, ny, 0.0, 3.0, 0.0, 2.0, jitter=0.1, seed=7)


def channel_lines(V, ylo, yhi):
    """Inflow (30000) lines on the element faces on y = ylo, outflow (20000)
    on y = yhi."""
    out = []
    for e in range(V.shape[1]):
        pts = [(V[2 * c, e], V[2 * c + 1, e]) for c in range(3)]
        for k in range(3):
            a, b = pts[k], pts[(k + 1) % 3]
            if a[1] == b[1] and a[1] in (ylo, yhi):
                out.append((30000 if a[1] == ylo else 20000, a, b))
    return out
