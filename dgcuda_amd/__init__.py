"""dgcuda_amd -- MI355X-native DG right-hand side and explicit RK step for 2D
advection and Euler on unstructured triangle meshes.

Python view of the C ABI in include/dgmi.h (libdgmi.so, built in-tree by
dgcuda_amd.build).  The classes mirror the reference's host entry points
(src/2d/conservation/main.cu:8-281): read_mesh -> Mesh.read_pmsh,
init_gpu + set_quadrature + preval_basis -> Solver(...), init_conditions ->
Solver.project_ic, eval_surface/eval_volume/eval_rhs -> Solver.rhs,
time_integrate_rk4/rk2 -> Solver.integrate, eval_u -> Solver.eval_vertices.
State arrays are numpy float64 in the reference layout
c[var][mode][elem] (shape (N, n_p, num_elem)).

There is no CPU fallback: importing works without a GPU (meshes and tables
are host code), but every solver call needs the HIP library and a GPU and
raises DGError otherwise.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# DGMI_LIB selects another build of the library (kernel-variant comparisons)
LIB_PATH = os.environ.get("DGMI_LIB") or os.path.join(_HERE, "libdgmi.so")

DG_OK = 0
DG_PHYS_ADVECTION, DG_PHYS_EULER = 1, 4
DG_FP64, DG_FP32, DG_GEOM_FP32 = 0, 1, 0x100
DG_RK4, DG_RK2 = 1, 2
DG_IC_CONSTANT, DG_IC_GAUSSIAN, DG_IC_VORTEX, DG_IC_SUPERSONIC, DG_IC_POLY = 0, 1, 2, 3, 4
DG_FLUX_LLF, DG_FLUX_CENTRAL = 0, 1
DG_QUIRK_FLOAT_SQRT, DG_QUIRK_FLOAT_GEOM, DG_QUIRK_FLOAT_POW, DG_QUIRK_REFLECT_MIRROR = 1, 2, 4, 8
DG_GMSH_QUIRKS, DG_GMSH_PERIODIC = 1, 2
DG_BATCH_ONE_LAUNCH, DG_BATCH_PER_CLASS = 0, 1
# the launch form a new batch starts with, per precision (the library's default)
BATCH_DEFAULT_LAUNCH_MODE = {DG_FP32: DG_BATCH_ONE_LAUNCH, DG_FP64: DG_BATCH_PER_CLASS}
STATUS_NAMES = {0: "DG_OK", 1: "DG_ERR_ARG", 2: "DG_ERR_IO", 3: "DG_ERR_PARSE", 4: "DG_ERR_HIP",
                5: "DG_ERR_UNSUPPORTED", 6: "DG_ERR_NUMERIC", 7: "DG_ERR_BUDGET", 8: "DG_ERR_STATE"}
STATUS_CODES = {v: k for k, v in STATUS_NAMES.items()}


class DGError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (STATUS_NAMES.get(code, code), msg))
        self.code = code


class ProblemDesc(C.Structure):
    _fields_ = [
        ("physics", C.c_int), ("ic", C.c_int), ("quirks", C.c_int), ("n_vortex", C.c_int),
        ("gamma", C.c_double), ("ax", C.c_double), ("ay", C.c_double),
        ("gauss_x", C.c_double), ("gauss_y", C.c_double), ("gauss_k", C.c_double),
        ("rho_inf", C.c_double), ("u_inf", C.c_double), ("v_inf", C.c_double), ("p_inf", C.c_double),
        ("vortex_beta", C.c_double * 4), ("vortex_x", C.c_double * 4), ("vortex_y", C.c_double * 4),
        ("constant", C.c_double * 4), ("mach", C.c_double), ("wall_x", C.c_double), ("wall_y", C.c_double),
        ("flux", C.c_int), ("poly_alpha", C.c_int),
    ]


class AuditOut(C.Structure):
    _fields_ = [("mass", C.c_double * 4), ("min_density", C.c_double), ("min_pressure", C.c_double),
                ("n_bad", C.c_int64), ("t", C.c_double), ("steps", C.c_int64)]


class ServerConfig(C.Structure):
    _fields_ = [("precision", C.c_int), ("max_slots", C.c_int), ("hbm_budget", C.c_int64),
                ("steps_per_round", C.c_int), ("cfl", C.c_double), ("use_graph", C.c_int),
                ("solo_min_elems", C.c_int), ("prep_threads", C.c_int), ("fifo_min_elems", C.c_int)]


class Request(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("x0", C.c_double), ("x1", C.c_double), ("y0", C.c_double),
                ("y1", C.c_double), ("jitter", C.c_double), ("mesh_seed", C.c_uint64), ("p", C.c_int),
                ("t_end", C.c_double), ("prob", ProblemDesc), ("arrival_s", C.c_double)]


class Result(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int), ("steps", C.c_int64), ("t", C.c_double),
                ("arrival_s", C.c_double), ("admit_s", C.c_double), ("finish_s", C.c_double),
                ("mass", C.c_double * 4), ("mass0", C.c_double * 4), ("n_bad", C.c_int64)]


_lib = None


def lib():
    """Load libdgmi.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libdgmi.so not built at %s; run python -m dgcuda_amd.build" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        P, D, I, I64, VP = C.POINTER, C.c_double, C.c_int, C.c_int64, C.c_void_p
        dp, ip = P(D), P(I)
        sig = {
            "dg_last_error": (C.c_char_p, []), "dg_version": (C.c_char_p, []), "dg_device_count": (I, []),
            "dg_mesh_read_pmsh": (I, [C.c_char_p, P(VP)]),
            "dg_mesh_read_gmsh": (I, [C.c_char_p, I, P(VP)]),
            "dg_mesh_read_vertex_pmsh": (I, [C.c_char_p, I, P(VP)]),
            "dg_mesh_period": (I, [VP, P(D), P(D)]),
            "dg_mesh_from_arrays": (I, [I, I, dp, ip, dp, ip, D, D, P(VP)]),
            "dg_mesh_periodic_jittered": (I, [I, I, D, D, D, D, D, C.c_uint64, P(VP)]),
            "dg_mesh_write_pmsh": (I, [VP, C.c_char_p]),
            "dg_mesh_info": (I, [VP, ip, ip, ip]),
            "dg_mesh_get_arrays": (I, [VP, dp, ip, dp, ip]),
            "dg_mesh_geometry": (I, [VP, I, dp, dp, dp]),
            "dg_mesh_destroy": (None, [VP]),
            "dg_tables": (I, [I, ip, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp]),
            "dg_solver_create": (I, [VP, P(ProblemDesc), I, I, P(VP)]),
            "dg_solver_destroy": (None, [VP]),
            "dg_solver_info": (I, [VP, ip, dp]),
            "dg_project_ic": (I, [VP]), "dg_set_state": (I, [VP, dp]), "dg_get_state": (I, [VP, dp]),
            "dg_rhs": (I, [VP, dp, dp, D, D]), "dg_rk_step": (I, [VP, D, I]), "dg_cfl_dt": (I, [VP, D, dp]),
            "dg_rhs_device_noalloc": (I, [VP, VP, VP, D, VP]),
            "dg_integrate": (I, [VP, D, D, I, P(I64), dp]), "dg_run_steps": (I, [VP, I, D, I, I]),
            "dg_audit": (I, [VP, P(AuditOut)]), "dg_eval_vertices": (I, [VP, I, dp]),
            "dg_solver_sync": (I, [VP]), "dg_solver_set_kernel": (I, [VP, I]), "dg_solver_get_kernel": (I, [VP, ip]),
            "dg_solver_set_max_groups": (I, [VP, I]), "dg_solver_launch_info": (I, [VP, ip]),
            "dg_batch_set_max_groups": (I, [VP, I]), "dg_batch_launch_info": (I, [VP, ip]),
            "dg_batch_set_launch_mode": (I, [VP, I]),
            "dg_time_stage_kernels": (I, [VP, I, D, dp]), "dg_time_step_graph": (I, [VP, I, D, dp]),
            "dg_solver_device_bytes": (I, [VP, P(I64)]),
            "dg_bandwidth_probe": (I, [I64, I, dp]), "dg_bandwidth_probe_ex": (I, [I64, I, dp, I]),
            "dg_selftest_mfma": (I, [dp]),
            "dg_write_msh": (I, [VP, I, dp, C.c_char_p]),
            "dg_write_msh_view": (I, [VP, C.c_char_p, I, dp, C.c_char_p]),
            "dg_eval_vertices_velocity": (I, [VP, dp]), "dg_eval_vertices_pressure": (I, [VP, dp]),
            "dg_l2_error": (I, [VP, P(ProblemDesc), I, dp, dp]),
            "dg_batch_create": (I, [I, P(VP)]), "dg_batch_destroy": (None, [VP]),
            "dg_batch_add": (I, [VP, VP, P(ProblemDesc), I, D, ip]), "dg_batch_project_ic": (I, [VP]),
            "dg_batch_run_steps": (I, [VP, I, D, I]), "dg_batch_integrate": (I, [VP, D, I, P(I64)]),
            "dg_batch_get_state": (I, [VP, I, dp]), "dg_batch_audit": (I, [VP, I, P(AuditOut)]),
            "dg_batch_info": (I, [VP, ip, P(I64), P(I64)]),
            "dg_server_create": (I, [P(ServerConfig), P(VP)]),
            "dg_server_submit": (I, [VP, P(Request), P(C.c_uint64)]),
            "dg_server_poll": (I, [VP, P(Result), I, ip]),
            "dg_server_get_state": (I, [VP, C.c_uint64, dp, I64]),
            "dg_server_keep_states": (I, [VP, I]), "dg_server_drain": (I, [VP]),
            "dg_server_now": (D, [VP]), "dg_server_destroy": (None, [VP]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


EXPORTED = None  # filled lazily: the names declared in include/dgmi.h


def _check(rc):
    if rc != DG_OK:
        raise DGError(rc, lib().dg_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def device_count():
    return lib().dg_device_count()


def bandwidth_probe(nbytes=1 << 30, iters=20, with_variant=False):
    """Streaming device copy, the fastest of 4 copy kernels: (ms per copy, GB/s
    counting read + write[, the kernel's number])."""
    out = np.zeros(3)
    _check(lib().dg_bandwidth_probe_ex(nbytes, iters, _dp(out), 3))
    return (out[0], out[1], int(out[2])) if with_variant else (out[0], out[1])


def n_modes(p):
    return (p + 1) * (p + 2) // 2


def tables(p):
    """FP64 quadrature and basis tables of order p in the reference's layouts."""
    sizes = np.zeros(3, np.int32)
    _check(lib().dg_tables(p, _ip(sizes), *([None] * 10)))
    np_, nq, n1 = (int(x) for x in sizes)
    out = dict(r1=np.zeros(nq), r2=np.zeros(nq), w=np.zeros(nq), s_r=np.zeros(n1), w1=np.zeros(n1),
               basis=np.zeros(np_ * nq), gx=np.zeros(np_ * nq), gy=np.zeros(np_ * nq),
               side=np.zeros(3 * np_ * n1), vertex=np.zeros(3 * np_))
    keys = ["r1", "r2", "w", "s_r", "w1", "basis", "gx", "gy", "side", "vertex"]
    _check(lib().dg_tables(p, _ip(sizes), *[_dp(out[k]) for k in keys]))
    out.update(n_p=np_, n_q=nq, n_q1=n1)
    return out


class Mesh:
    """Triangle mesh in the reference's read_mesh layout (host only)."""

    def __init__(self, handle, period=(0.0, 0.0)):
        self._h = handle
        self.period = period
        K, S, B = C.c_int(), C.c_int(), C.c_int()
        _check(lib().dg_mesh_info(self._h, C.byref(K), C.byref(S), C.byref(B)))
        self.num_elem, self.num_sides, self.num_boundary_sides = K.value, S.value, B.value

    @classmethod
    def read_pmsh(cls, path):
        h = C.c_void_p()
        _check(lib().dg_mesh_read_pmsh(str(path).encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def read_gmsh(cls, path, genmesh_quirks=False, periodic=False):
        """gmsh 2.2 .msh converted as the reference's genmesh.py does (with
        genmesh_quirks, exactly, including its latent bugs).  periodic pairs
        the open untagged sides on opposite edges of the bounding box."""
        h = C.c_void_p()
        flags = (DG_GMSH_QUIRKS if genmesh_quirks else 0) | (DG_GMSH_PERIODIC if periodic else 0)
        _check(lib().dg_mesh_read_gmsh(str(path).encode(), flags, C.byref(h)))
        px, py = C.c_double(), C.c_double()
        _check(lib().dg_mesh_period(h, C.byref(px), C.byref(py)))
        return cls(h, (px.value, py.value))

    @classmethod
    def read_vertex_pmsh(cls, path, open_kind=-3):
        """The advection prototype's 6-field .pmsh with its sides built as the
        prototype's read_mesh does; sides met once get right element
        open_kind (-3: the problem's exact state, the prototype's boundary)."""
        h = C.c_void_p()
        _check(lib().dg_mesh_read_vertex_pmsh(str(path).encode(), open_kind, C.byref(h)))
        return cls(h)

    @classmethod
    def periodic_jittered(cls, nx, ny, x0, x1, y0, y1, jitter=0.1, seed=0):
        h = C.c_void_p()
        _check(lib().dg_mesh_periodic_jittered(nx, ny, x0, x1, y0, y1, jitter, seed, C.byref(h)))
        return cls(h, (x1 - x0, y1 - y0))

    @classmethod
    def from_arrays(cls, V, elem_s, sides, side_int, period=(0.0, 0.0)):
        V = np.ascontiguousarray(V, np.float64).reshape(-1)
        elem_s = np.ascontiguousarray(elem_s, np.int32).reshape(-1)
        sides = np.ascontiguousarray(sides, np.float64).reshape(-1)
        side_int = np.ascontiguousarray(side_int, np.int32).reshape(-1)
        K, S = V.size // 6, sides.size // 4
        h = C.c_void_p()
        _check(lib().dg_mesh_from_arrays(K, S, _dp(V), _ip(elem_s), _dp(sides), _ip(side_int),
                                         period[0], period[1], C.byref(h)))
        return cls(h, period)

    def arrays(self):
        """(V[6,K], elem_s[3,K], sides[4,S], side_int[4,S]) in read_mesh order."""
        K, S = self.num_elem, self.num_sides
        V, es = np.zeros((6, K)), np.zeros((3, K), np.int32)
        sd, si = np.zeros((4, S)), np.zeros((4, S), np.int32)
        _check(lib().dg_mesh_get_arrays(self._h, _dp(V), _ip(es), _dp(sd), _ip(si)))
        return V, es, sd, si

    def geometry(self, quirks=0):
        """(geo_e[5,K] = J,xr,yr,xs,ys ; geo_s[3,S] = len,Nx,Ny ; min_r)"""
        ge, gs, r = np.zeros((5, self.num_elem)), np.zeros((3, self.num_sides)), C.c_double()
        _check(lib().dg_mesh_geometry(self._h, quirks, _dp(ge), _dp(gs), C.byref(r)))
        return ge, gs, r.value

    def write_pmsh(self, path):
        _check(lib().dg_mesh_write_pmsh(self._h, str(path).encode()))

    def write_msh(self, var, uv, path):
        """gmsh view of vertex values uv[K, 3] as the reference writes output/U<var>.msh."""
        uv = np.ascontiguousarray(uv, np.float64).reshape(-1)
        _check(lib().dg_write_msh(self._h, var, _dp(uv), str(path).encode()))

    def write_msh_view(self, name, vals, path):
        """A named gmsh view (the reference CPU driver's output files):
        vals[K, 3] scalars (ST lines) or vals[K, 3, 2] vectors (VT lines)."""
        vals = np.ascontiguousarray(vals, np.float64)
        ncomp = 2 if vals.ndim == 3 else 1
        _check(lib().dg_write_msh_view(self._h, name.encode(), ncomp, _dp(vals.reshape(-1)), str(path).encode()))

    def l2_error(self, prob, p, c):
        """Per-variable L2 error of a state c[N, n_p, K] against the problem's U0."""
        c = np.ascontiguousarray(c, np.float64)
        err = np.zeros(prob.physics)
        _check(lib().dg_l2_error(self._h, C.byref(prob), p, _dp(c), _dp(err)))
        return err

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.dg_mesh_destroy(self._h)
            self._h = None


def problem(physics, ic, gamma=1.4, a=(1.0, 1.0), gauss=(0.0, 0.0, 40.0), background=(1.0, 1.0, 1.0, 1.0),
            vortices=(), constant=(0.0, 0.0, 0.0, 0.0), quirks=0, wall=(0.0, 0.0), flux=DG_FLUX_LLF, alpha=0):
    """Build a ProblemDesc.  vortices = [(beta, xc, yc), ...]; wall = centre of
    the reflecting sides' radial normal; flux = DG_FLUX_LLF / DG_FLUX_CENTRAL
    (advection); alpha = exponent of DG_IC_POLY."""
    d = ProblemDesc()
    d.wall_x, d.wall_y = wall
    d.flux, d.poly_alpha = flux, alpha
    d.physics, d.ic, d.quirks, d.n_vortex = physics, ic, quirks, len(vortices)
    d.gamma, d.ax, d.ay = gamma, a[0], a[1]
    d.gauss_x, d.gauss_y, d.gauss_k = gauss
    d.rho_inf, d.u_inf, d.v_inf, d.p_inf = background
    for k, (b, x, y) in enumerate(vortices):
        d.vortex_beta[k], d.vortex_x[k], d.vortex_y[k] = b, x, y
    for k in range(4):
        d.constant[k] = constant[k]
    return d


def advection_gaussian(a=(1.0, 1.0), center=(0.0, 0.0), k=40.0, quirks=0):
    return problem(DG_PHYS_ADVECTION, DG_IC_GAUSSIAN, a=a, gauss=(center[0], center[1], k), quirks=quirks)


def advection_poly(alpha, a=(1.0, 1.0), flux=DG_FLUX_CENTRAL, quirks=0):
    """(x - y)^alpha advected by a, the advection prototype's test problem
    (src/2d/test_all.cu): stationary for a = (1, 1); central flux by default,
    as the prototype."""
    return problem(DG_PHYS_ADVECTION, DG_IC_POLY, a=a, flux=flux, alpha=alpha, quirks=quirks)


def euler_vortex(vortices=((5.0, 5.0, 5.0),), gamma=1.4, background=(1.0, 1.0, 1.0, 1.0), quirks=0):
    return problem(DG_PHYS_EULER, DG_IC_VORTEX, gamma=gamma, background=background, vortices=vortices, quirks=quirks)


def euler_constant(U, gamma=1.4, quirks=0):
    return problem(DG_PHYS_EULER, DG_IC_CONSTANT, gamma=gamma, constant=U, quirks=quirks)


def cylinder(mach=0.38, gamma=1.4, quirks=0):
    """Flow around a cylinder (euler/cylinder/cylinder.cu): the uniform state
    rho = gamma, u = mach, v = 0, p = 1 as initial, inflow and outflow state;
    reflecting sides about the centre (0.5, 0.5).  The reference integrates
    it with RK2 (DG_RK2)."""
    U = (gamma, gamma * mach, 0.0, 0.5 * gamma * mach * mach + 1.0 / (gamma - 1.0))
    return problem(DG_PHYS_EULER, DG_IC_CONSTANT, gamma=gamma, constant=U, quirks=quirks, wall=(0.5, 0.5))


def supersonic_vortex(mach=2.25, gamma=1.4, quirks=0):
    """Supersonic vortex between the arcs r = 1 and r = 1.384
    (euler/supersonic/supersonic.cu): the exact steady state is the initial
    state, which the inflow and outflow sides impose; walls reflect."""
    d = problem(DG_PHYS_EULER, DG_IC_SUPERSONIC, gamma=gamma, quirks=quirks)
    d.mach = mach
    return d


class Solver:
    """One DG solve resident on the GPU (the reference's init_gpu ... free_gpu)."""

    def __init__(self, mesh, prob, p, precision=DG_FP64):
        self._h = C.c_void_p()
        self.mesh = mesh
        _check(lib().dg_solver_create(mesh._h, C.byref(prob), p, precision, C.byref(self._h)))
        sizes, r = np.zeros(6, np.int32), C.c_double()
        _check(lib().dg_solver_info(self._h, _ip(sizes), C.byref(r)))
        self.num_elem, self.num_sides, self.n_p, self.n_q, self.n_q1, self.N = (int(x) for x in sizes)
        self.min_r = r.value
        self.p, self.precision = p, precision & ~DG_GEOM_FP32

    @property
    def shape(self):
        return (self.N, self.n_p, self.num_elem)

    def project_ic(self):
        _check(lib().dg_project_ic(self._h))

    def set_state(self, c):
        c = np.ascontiguousarray(c, np.float64)
        assert c.size == self.N * self.n_p * self.num_elem
        _check(lib().dg_set_state(self._h, _dp(c)))

    def get_state(self):
        c = np.zeros(self.shape)
        _check(lib().dg_get_state(self._h, _dp(c)))
        return c

    def rhs(self, u, dt=1.0, t=0.0):
        """k = dt * M^-1 R(u) (eval_surface + eval_volume + eval_rhs)."""
        u = np.ascontiguousarray(u, np.float64)
        k = np.zeros(self.shape)
        _check(lib().dg_rhs(self._h, _dp(u), _dp(k), t, dt))
        return k

    def rhs_device_noalloc(self, u_dev, k_dev, dt=1.0, stream=None):
        """k = dt * M^-1 R(u) on device buffers (addresses as ints, FP64,
        reference layout) through the library's fixed scratch: no
        allocation, enqueued on `stream` (None: the solver's), not waited for."""
        _check(lib().dg_rhs_device_noalloc(self._h, C.c_void_p(u_dev), C.c_void_p(k_dev), dt,
                                           C.c_void_p(stream) if stream else None))

    def rk_step(self, dt, scheme=DG_RK4):
        _check(lib().dg_rk_step(self._h, dt, scheme))

    def cfl_dt(self, cfl):
        dt = C.c_double()
        _check(lib().dg_cfl_dt(self._h, cfl, C.byref(dt)))
        return dt.value

    def integrate(self, t_end, cfl, scheme=DG_RK4):
        n, t = C.c_int64(), C.c_double()
        _check(lib().dg_integrate(self._h, t_end, cfl, scheme, C.byref(n), C.byref(t)))
        return n.value, t.value

    def run_steps(self, nsteps, cfl, scheme=DG_RK4, use_graph=True):
        _check(lib().dg_run_steps(self._h, nsteps, cfl, scheme, 1 if use_graph else 0))

    def audit(self):
        a = AuditOut()
        _check(lib().dg_audit(self._h, C.byref(a)))
        return dict(mass=list(a.mass)[: self.N], min_density=a.min_density, min_pressure=a.min_pressure,
                    n_bad=a.n_bad, t=a.t, steps=a.steps)

    def eval_vertices(self, var):
        uv = np.zeros((self.num_elem, 3))
        _check(lib().dg_eval_vertices(self._h, var, _dp(uv)))
        return uv

    def eval_vertices_velocity(self):
        """(u, v) at the three vertices of each element: shape (K, 3, 2)."""
        uv = np.zeros((self.num_elem, 3, 2))
        _check(lib().dg_eval_vertices_velocity(self._h, _dp(uv)))
        return uv

    def eval_vertices_pressure(self):
        """Pressure at the three vertices of each element: shape (K, 3)."""
        pv = np.zeros((self.num_elem, 3))
        _check(lib().dg_eval_vertices_pressure(self._h, _dp(pv)))
        return pv

    def sync(self):
        _check(lib().dg_solver_sync(self._h))

    def set_kernel(self, kind):
        _check(lib().dg_solver_set_kernel(self._h, kind))

    def kernel(self):
        k = C.c_int()
        _check(lib().dg_solver_get_kernel(self._h, C.byref(k)))
        return k.value

    def set_max_groups(self, n):
        """Cap the persistent matrix-core grid at n workgroups (0 = all resident)."""
        _check(lib().dg_solver_set_max_groups(self._h, n))

    def launch_info(self):
        """dict(kind, tiles, groups, walk): walk = most tiles one workgroup processes."""
        o = np.zeros(4, np.int32)
        _check(lib().dg_solver_launch_info(self._h, _ip(o)))
        return dict(kind=int(o[0]), tiles=int(o[1]), groups=int(o[2]), walk=int(o[3]))

    def time_stage_kernels(self, nsteps, cfl):
        out = np.zeros(4)
        _check(lib().dg_time_stage_kernels(self._h, nsteps, cfl, _dp(out)))
        return dict(stage_ms=out[0], dt_ms=out[1], total_ms=out[2], launches=int(out[3]))

    def time_step_graph(self, nsteps, cfl):
        out = np.zeros(2)
        _check(lib().dg_time_step_graph(self._h, nsteps, cfl, _dp(out)))
        return dict(step_ms=out[0], total_ms=out[1])

    def device_bytes(self):
        b = C.c_int64()
        _check(lib().dg_solver_device_bytes(self._h, C.byref(b)))
        return b.value

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.dg_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class Batch:
    """Many solves advanced by one persistent launch per RK stage."""

    def __init__(self, precision=DG_FP32):
        self._h = C.c_void_p()
        _check(lib().dg_batch_create(precision, C.byref(self._h)))
        self.meshes, self.shapes = [], []

    def add(self, mesh, prob, p, t_end):
        slot = C.c_int()
        _check(lib().dg_batch_add(self._h, mesh._h, C.byref(prob), p, t_end, C.byref(slot)))
        self.meshes.append(mesh)
        self.shapes.append((prob.physics, n_modes(p), mesh.num_elem))
        return slot.value

    def project_ic(self):
        _check(lib().dg_batch_project_ic(self._h))

    def run_steps(self, nsteps, cfl, use_graph=True):
        _check(lib().dg_batch_run_steps(self._h, nsteps, cfl, 1 if use_graph else 0))

    def integrate(self, cfl, use_graph=True):
        n = C.c_int64()
        _check(lib().dg_batch_integrate(self._h, cfl, 1 if use_graph else 0, C.byref(n)))
        return n.value

    def get_state(self, slot):
        c = np.zeros(self.shapes[slot])
        _check(lib().dg_batch_get_state(self._h, slot, _dp(c)))
        return c

    def audit(self, slot):
        a = AuditOut()
        _check(lib().dg_batch_audit(self._h, slot, C.byref(a)))
        return dict(mass=list(a.mass), min_density=a.min_density, min_pressure=a.min_pressure, n_bad=a.n_bad,
                    t=a.t, steps=a.steps)

    def set_max_groups(self, n):
        """Cap the persistent grid at n workgroups (0 = all resident)."""
        _check(lib().dg_batch_set_max_groups(self._h, n))

    def set_launch_mode(self, mode):
        """DG_BATCH_ONE_LAUNCH (one persistent launch per stage; the FP32
        default) or DG_BATCH_PER_CLASS (one per kernel class and stage; the
        FP64 default); from the next run on."""
        _check(lib().dg_batch_set_launch_mode(self._h, mode))

    def launch_info(self):
        """dict(groups, tile_items, lane_items, longest_run) of the uploaded
        work list (groups: the largest launch)."""
        o = np.zeros(4, np.int32)
        _check(lib().dg_batch_launch_info(self._h, _ip(o)))
        return dict(groups=int(o[0]), tile_items=int(o[1]), lane_items=int(o[2]), longest_run=int(o[3]))

    def info(self):
        n, e, b = C.c_int(), C.c_int64(), C.c_int64()
        _check(lib().dg_batch_info(self._h, C.byref(n), C.byref(e), C.byref(b)))
        return dict(n_solves=n.value, total_elems=e.value, device_bytes=b.value)

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.dg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


# Solves of this many triangles and more run their own stage kernels inside
# the step (the single-solve kernels, a few tiles per resident workgroup);
# smaller ones share the batch launch.  On a Poisson stream of 1k-500k
# triangle solves this threshold served 5.3 solves/s against 4.7 at 65,536
# (p99 latency 5.3 against 6.1 s).
SOLO_MIN_ELEMS = 262144
# Solves of this many triangles and more advance one at a time in arrival
# order (the smaller ones beside them): each keeps the GPU busy alone, and
# advancing them together would make the first to arrive wait for the later
# ones (served p99 latency above the plain one-at-a-time path's).
FIFO_MIN_ELEMS = 262144


class Server:
    """Continuous-batching solve server."""

    def __init__(self, precision=DG_FP32, max_slots=64, hbm_budget=64 << 30, steps_per_round=8, cfl=0.3,
                 use_graph=True, solo_min_elems=SOLO_MIN_ELEMS, prep_threads=4, fifo_min_elems=FIFO_MIN_ELEMS):
        cfg = ServerConfig(precision, max_slots, hbm_budget, steps_per_round, cfl, 1 if use_graph else 0,
                           solo_min_elems, prep_threads, fifo_min_elems)
        self._h = C.c_void_p()
        _check(lib().dg_server_create(C.byref(cfg), C.byref(self._h)))

    def submit(self, req):
        i = C.c_uint64()
        _check(lib().dg_server_submit(self._h, C.byref(req), C.byref(i)))
        return i.value

    def poll(self, max_results=256):
        buf = (Result * max_results)()
        n = C.c_int()
        _check(lib().dg_server_poll(self._h, buf, max_results, C.byref(n)))
        return [buf[i] for i in range(n.value)]

    def keep_states(self, keep=True):
        _check(lib().dg_server_keep_states(self._h, 1 if keep else 0))

    def get_state(self, rid, shape):
        c = np.zeros(shape)
        _check(lib().dg_server_get_state(self._h, rid, _dp(c), c.size))
        return c

    def drain(self):
        _check(lib().dg_server_drain(self._h))

    def now(self):
        return lib().dg_server_now(self._h)

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.dg_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def request(nx, p, t_end, prob, domain=(0.0, 10.0), jitter=0.1, mesh_seed=0, arrival_s=0.0):
    r = Request()
    r.nx = r.ny = nx
    r.x0, r.x1 = domain
    r.y0, r.y1 = domain
    r.jitter, r.mesh_seed, r.p, r.t_end, r.prob, r.arrival_s = jitter, mesh_seed, p, t_end, prob, arrival_s
    return r


def selftest_mfma():
    """Max abs error of a 16x16x4 FP32 and an FP64 matrix-core product vs the host."""
    e = C.c_double()
    _check(lib().dg_selftest_mfma(C.byref(e)))
    return e.value
