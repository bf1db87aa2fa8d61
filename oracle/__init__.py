"""FP64 serial CPU oracle of the reference DG path (oracle/dg_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product (dgcuda_amd) never
imports it.

Parity status of each oracle function (what pins it; see tests/test_oracle*.py):
  tables (set_quadrature, phi, preval_basis)  transcription parsed from the reference
      source; quadrature exactness on monomials; orthonormality bounds; exact
      Gram-Schmidt regeneration of the basis
  geometry                                    closed forms (areas, unit normals,
      outward orientation, inscribed diameter of right triangles)
  project_ic                                  polynomial reproduction, and the
      reference's own golden vertex dumps src/2d/tests/t0/*.out
  rhs (surface + volume + gather)             independent physical-space weak-form
      evaluation on tiny meshes; free-stream preservation; discrete conservation;
      LLF consistency and symmetry; exactness for advection
  rk_step / integrate                         Butcher coefficients (closed form on a
      linear ODE per mode), temporal order 4 / 2
  cfl_dt                                      closed form
  end to end                                  p+1 convergence on the isentropic vortex
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        D, I, VP = C.c_double, C.c_int, C.c_void_p
        dp, ip = C.POINTER(D), C.POINTER(I)
        sig = {
            "or_create": (VP, [I, I, dp, ip, dp, ip, I, I, dp, I, I]),
            "or_destroy": (None, [VP]), "or_set_hoist": (None, [VP, I]),
            "or_phi": (D, [D, D, I]), "or_phi_x": (D, [D, D, I]), "or_phi_y": (D, [D, D, I]),
            "or_sizes": (None, [VP, ip]),
            "or_tables": (None, [VP] + [dp] * 10),
            "or_geometry": (D, [VP, dp, dp]),
            "or_project_ic": (None, [VP, dp, dp]),
            "or_set_bc_ic": (None, [VP, dp, I]), "or_set_wall": (None, [VP, D, D]),
            "or_set_flux": (None, [VP, I]),
            "or_ic_point": (None, [dp, I, D, D, D, dp]),
            "or_rhs": (I, [VP, dp, dp, D]),
            "or_rhs_parts": (I, [VP, dp, dp, dp, dp]),
            "or_cell_lambda": (None, [VP, dp, dp]),
            "or_cfl_dt": (D, [VP, dp, D]),
            "or_rk_step": (I, [VP, dp, D, I]),
            "or_integrate": (I, [VP, dp, D, D, I, dp]),
            "or_run_steps": (D, [VP, dp, I, D, I]),
            "or_eval_vertices": (None, [VP, dp, I, dp]),
            "or_mass": (None, [VP, dp, dp]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype, f.argtypes = r, a
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def phi(r, s, i):
    return lib().or_phi(r, s, i)


def phi_x(r, s, i):
    return lib().or_phi_x(r, s, i)


def phi_y(r, s, i):
    return lib().or_phi_y(r, s, i)


def ic_poly(alpha):
    """(x - y)^alpha, the reference's projection test function (test_all.cu)."""
    return np.array([3.0, float(alpha)])


def ic_point(ic, N, gamma, x, y):
    """The oracle's U0(x, y) for an IC parameter array."""
    ic = np.ascontiguousarray(ic, np.float64)
    U = np.zeros(4)
    lib().or_ic_point(_dp(ic), N, gamma, x, y, _dp(U))
    return U[:N]


def ic_array(prob, period=(0.0, 0.0)):
    """Flatten a dgcuda_amd.ProblemDesc into the oracle's IC parameter array."""
    if prob.ic == 0:
        return np.array([0.0] + list(prob.constant))
    if prob.ic == 1:
        return np.array([1.0, prob.gauss_x, prob.gauss_y, prob.gauss_k, period[0], period[1]])
    if prob.ic == 3:  # supersonic vortex: {4, mach, float_pow}
        return np.array([4.0, prob.mach, 1.0 if prob.quirks & 4 else 0.0])
    if prob.ic == 4:  # (x - y)^alpha: {3, alpha}
        return ic_poly(prob.poly_alpha)
    a = [2.0, float(prob.n_vortex), prob.rho_inf, prob.u_inf, prob.v_inf, prob.p_inf, period[0], period[1]]
    for k in range(prob.n_vortex):
        a += [prob.vortex_beta[k], prob.vortex_x[k], prob.vortex_y[k]]
    return np.array(a)


class Oracle:
    """FP64 restatement of one solve: mesh arrays in read_mesh order."""

    def __init__(self, V, elem_s, sides, side_int, p, physics, gamma=1.4, a=(1.0, 1.0), quirks=0, hoist=1):
        self.V = np.ascontiguousarray(V, np.float64).reshape(-1)
        self.es = np.ascontiguousarray(elem_s, np.int32).reshape(-1)
        self.sd = np.ascontiguousarray(sides, np.float64).reshape(-1)
        self.si = np.ascontiguousarray(side_int, np.int32).reshape(-1)
        self.K, self.S = self.V.size // 6, self.sd.size // 4
        self.phys = np.array([gamma, a[0], a[1]])
        self._h = lib().or_create(self.K, self.S, _dp(self.V), _ip(self.es), _dp(self.sd), _ip(self.si), p, physics,
                                  _dp(self.phys), quirks, hoist)
        if not self._h:
            raise ValueError("oracle rejects this mesh / order / physics")
        sz = np.zeros(4, np.int32)
        lib().or_sizes(self._h, _ip(sz))
        self.n_p, self.n_q, self.n_q1, self.N = (int(x) for x in sz)
        self.p = p

    @classmethod
    def from_mesh(cls, mesh, p, prob, hoist=1):
        V, es, sd, si = mesh.arrays()
        o = cls(V, es, sd, si, p, prob.physics, prob.gamma, (prob.ax, prob.ay), prob.quirks, hoist)
        if prob.flux:
            o.set_flux(prob.flux)
        if (np.asarray(si).reshape(4, -1)[1] < 0).any():  # boundary sides: inflow / outflow states, wall centre
            o.set_bc_ic(ic_array(prob, mesh.period))
            o.set_wall(prob.wall_x, prob.wall_y)
        return o

    def set_flux(self, flux):
        """Advection flux: 0 local Lax-Friedrichs, 1 central (lambda = 0)."""
        lib().or_set_flux(self._h, flux)

    def set_wall(self, x, y):
        """Centre of the reflecting sides' radial wall normal."""
        lib().or_set_wall(self._h, x, y)

    def set_bc_ic(self, ic):
        """IC parameter array (as for project_ic) of the inflow / outflow states."""
        self._bc_ic = np.ascontiguousarray(ic, np.float64)
        lib().or_set_bc_ic(self._h, _dp(self._bc_ic), self._bc_ic.size)

    @property
    def shape(self):
        return (self.N, self.n_p, self.K)

    def set_hoist(self, h):
        lib().or_set_hoist(self._h, h)

    def tables(self):
        np_, nq, n1 = self.n_p, self.n_q, self.n_q1
        t = dict(r1=np.zeros(nq), r2=np.zeros(nq), w=np.zeros(nq), s_r=np.zeros(n1), w1=np.zeros(n1),
                 basis=np.zeros(np_ * nq), gx=np.zeros(np_ * nq), gy=np.zeros(np_ * nq),
                 side=np.zeros(3 * np_ * n1), vertex=np.zeros(3 * np_))
        keys = ["r1", "r2", "w", "s_r", "w1", "basis", "gx", "gy", "side", "vertex"]
        lib().or_tables(self._h, *[_dp(t[k]) for k in keys])
        return t

    def geometry(self):
        ge, gs = np.zeros((5, self.K)), np.zeros((3, self.S))
        r = lib().or_geometry(self._h, _dp(ge), _dp(gs))
        return ge, gs, r

    def project_ic(self, ic):
        u = np.zeros(self.shape)
        ic = np.ascontiguousarray(ic, np.float64)
        lib().or_project_ic(self._h, _dp(ic), _dp(u))
        return u

    def rhs(self, u, dt=1.0):
        u = np.ascontiguousarray(u, np.float64)
        k = np.zeros(self.shape)
        if lib().or_rhs(self._h, _dp(u), _dp(k), dt) != 0:
            raise ValueError("oracle: mesh has boundary sides and no inflow / outflow state (set_bc_ic)")
        return k

    def rhs_parts(self, u):
        u = np.ascontiguousarray(u, np.float64)
        q = np.zeros(self.shape)
        l, r = np.zeros((self.N, self.n_p, self.S)), np.zeros((self.N, self.n_p, self.S))
        if lib().or_rhs_parts(self._h, _dp(u), _dp(q), _dp(l), _dp(r)) != 0:
            raise ValueError("oracle: mesh has boundary sides and no inflow / outflow state (set_bc_ic)")
        return q, l, r

    def cell_lambda(self, u):
        u = np.ascontiguousarray(u, np.float64)
        lam = np.zeros(self.K)
        lib().or_cell_lambda(self._h, _dp(u), _dp(lam))
        return lam

    def cfl_dt(self, u, cfl):
        u = np.ascontiguousarray(u, np.float64)
        return lib().or_cfl_dt(self._h, _dp(u), cfl)

    def rk_step(self, u, dt, scheme=1):
        u = np.array(u, np.float64, copy=True)
        if lib().or_rk_step(self._h, _dp(u), dt, scheme) != 0:
            raise ValueError("oracle: mesh has boundary sides and no inflow / outflow state (set_bc_ic)")
        return u

    def integrate(self, u, t_end, cfl, scheme=1):
        u = np.array(u, np.float64, copy=True)
        t = C.c_double()
        n = lib().or_integrate(self._h, _dp(u), t_end, cfl, scheme, C.byref(t))
        if n < 0:
            raise ValueError("oracle: mesh has boundary sides and no inflow / outflow state (set_bc_ic)")
        return u, n, t.value

    def run_steps(self, u, nsteps, cfl, scheme=1):
        u = np.array(u, np.float64, copy=True)
        t = lib().or_run_steps(self._h, _dp(u), nsteps, cfl, scheme)
        return u, t

    def eval_vertices(self, u, var):
        u = np.ascontiguousarray(u, np.float64)
        uv = np.zeros((self.K, 3))
        lib().or_eval_vertices(self._h, _dp(u), var, _dp(uv))
        return uv

    def mass(self, u):
        u = np.ascontiguousarray(u, np.float64)
        m = np.zeros(self.N)
        lib().or_mass(self._h, _dp(u), _dp(m))
        return m

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.or_destroy(self._h)
            self._h = None


def rel_l2(a, b, J=None):
    """Relative L2 distance of two modal states (Parseval on the orthonormal
    basis): sqrt(sum_K J_K sum |a-b|^2) / sqrt(sum_K J_K sum |b|^2)."""
    a, b = np.asarray(a), np.asarray(b)
    w = np.ones(a.shape[-1]) if J is None else np.abs(np.asarray(J))
    num = np.sqrt(np.sum(w * np.sum((a - b) ** 2, axis=tuple(range(a.ndim - 1)))))
    den = np.sqrt(np.sum(w * np.sum(b ** 2, axis=tuple(range(b.ndim - 1)))))
    return num / den
