"""Shared helpers of the test suite (no GPU needed to import)."""
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
    """Projected initial condition plus a seeded perturbation of every mode
    (keeps density and pressure positive for the Euler vortex)."""
    u = o.project_ic(orc.ic_array(prob, period))
    rng = np.random.default_rng(seed)
    scale = np.abs(u).max(axis=(1, 2), keepdims=True) + 1e-3
    return u + amp * scale * rng.standard_normal(u.shape)


def rel_max(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
