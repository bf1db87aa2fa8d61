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
