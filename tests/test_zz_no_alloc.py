"""dg_rhs_device_noalloc, the allocation-free right-hand side, through the
C ABI against dg_rhs on the same solver: equal bit for bit (the same
kernels; the state conversions round the same way on the device as on the
host).  Calls in a row share the device's scratch: a 1,152-triangle FP32
p = 4 solve on the matrix-core kernel, then a smaller FP64 p = 3 one, then
a lane-kernel FP32 p = 2 one, then the first again.  A solve whose two
state arrays exceed the 64 MiB scratch is refused with DG_ERR_BUDGET before
any device work.

Device buffers come from the HIP runtime libdgmi.so itself loads
(libamdhip64.so.7, resolved to the already loaded copy).
"""
import ctypes as C

import numpy as np
import pytest

import dgcuda_amd as dg
import oracle as orc
from helpers import mesh, perturbed_state, problem_for

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    def __init__(self):
        dg.lib()  # the runtime libdgmi.so links, loaded first
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []

    def check(self, rc, what):
        assert rc == 0, "%s failed with HIP error %d" % (what, rc)

    def malloc(self, nbytes):
        p = C.c_void_p()
        self.check(self.L.hipMalloc(C.byref(p), nbytes), "hipMalloc")
        return p.value

    def free(self, p):
        self.check(self.L.hipFree(C.c_void_p(p)), "hipFree")

    def put(self, dst, a):
        self.check(self.L.hipMemcpy(C.c_void_p(dst), a.ctypes.data_as(C.c_void_p), a.nbytes, H2D), "hipMemcpy")

    def get(self, src, shape):
        a = np.empty(shape)
        self.check(self.L.hipMemcpy(a.ctypes.data_as(C.c_void_p), C.c_void_p(src), a.nbytes, D2H), "hipMemcpy")
        return a

    def sync(self):
        self.check(self.L.hipDeviceSynchronize(), "hipDeviceSynchronize")


def case(n, N, p, prec, kind, seed):
    m = mesh(n, seed=seed)
    prob = problem_for(N)
    o = orc.Oracle.from_mesh(m, p, prob)
    u = perturbed_state(o, prob, m.period, seed=seed)
    s = dg.Solver(m, prob, p, prec)
    s.set_kernel(kind)
    return s, u


def noalloc_rhs(hip, s, u, dt):
    du, dk = hip.malloc(u.nbytes), hip.malloc(u.nbytes)
    try:
        hip.put(du, np.ascontiguousarray(u))
        hip.put(dk, np.full(u.shape, np.nan))  # every output entry must be written
        s.rhs_device_noalloc(du, dk, dt)
        hip.sync()
        return hip.get(dk, u.shape)
    finally:
        hip.free(du)
        hip.free(dk)


def test_noalloc_rhs_equals_dg_rhs_calls_in_a_row():
    hip = Hip()
    cases = [case(24, 4, 4, dg.DG_FP32, 2, 1),   # 1,152 triangles, matrix-core tiles
             case(10, 4, 3, dg.DG_FP64, 2, 2),   # smaller: 200 triangles
             case(7, 1, 2, dg.DG_FP32, 1, 3)]    # smaller still, lane kernel, advection
    for k in (0, 1, 2, 0):
        s, u = cases[k]
        print("noalloc rhs case %d: %d triangles p=%d prec=%d kernel %d"
              % (k, s.num_elem, s.p, s.precision, s.kernel()), flush=True)
        got = noalloc_rhs(hip, s, u, 0.01)
        want = s.rhs(u, dt=0.01)
        assert np.isfinite(got).all()
        assert np.array_equal(got, want), (k, float(np.abs(got - want).max()))


def test_noalloc_rhs_refuses_a_solve_beyond_the_scratch():
    hip = Hip()
    m = mesh(192, seed=4)  # 73,728 triangles: FP64 p = 4 needs 2 x 37.7 MB
    prob = problem_for(4)
    s = dg.Solver(m, prob, 4, dg.DG_FP64)
    s.set_kernel(2)
    n = 4 * 15 * m.num_elem
    du, dk = hip.malloc(8 * n), hip.malloc(8 * n)
    try:
        with pytest.raises(dg.DGError) as ei:
            s.rhs_device_noalloc(du, dk, 0.01)
        assert ei.value.code == dg.STATUS_CODES["DG_ERR_BUDGET"]
        assert "scratch" in str(ei.value)
        hip.sync()
    finally:
        hip.free(du)
        hip.free(dk)
