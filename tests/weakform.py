"""Independent evaluation of the DG weak form in physical coordinates.

Written from the mathematics, not from the reference's loop nest: for each
element K with affine map x = v1 + A (r, s),

  k_i = dt / J_K * ( int_K F(u_h) . grad_x phi_i dx
                     - sum_faces int_e F*(u_in, u_out, n_out) phi_i ds )

with grad_x phi = A^-T grad_r phi, n_out the outward unit normal of K, u_out
the neighbour's trace at the same physical point (the neighbour located
through the side list, its vertices shifted by the periodic offset), and LLF
F* = 1/2 [(F(a) + F(b)) . n + lambda (a - b)].  The mass matrix is J*I, as the
reference assumes.  Used only by the tests.
"""
import numpy as np


def monomial_exps(n):
    out = []
    d = 0
    while len(out) < n:
        for b in range(d + 1):
            out.append((d - b, b))
        d += 1
    return out[:n]


class Basis:
    def __init__(self, coeff, n_p):
        self.c = np.array(coeff)[:n_p, :n_p]
        self.e = monomial_exps(n_p)
        self.n_p = n_p

    def val(self, r, s):
        m = np.array([r ** a * s ** b for a, b in self.e])
        return self.c @ m

    def grad(self, r, s):
        mr = np.array([a * r ** (a - 1) * s ** b if a > 0 else 0.0 for a, b in self.e])
        ms = np.array([b * r ** a * s ** (b - 1) if b > 0 else 0.0 for a, b in self.e])
        return self.c @ mr, self.c @ ms


def euler_flux(U, g):
    rho, ru, rv, E = U
    p = (g - 1) * (E - 0.5 * (ru * ru + rv * rv) / rho)
    fx = np.array([ru, ru * ru / rho + p, ru * rv / rho, ru * (E + p) / rho])
    fy = np.array([rv, ru * rv / rho, rv * rv / rho + p, rv * (E + p) / rho])
    return fx, fy, p


def llf(a, b, n, phys, g, adv, central=False):
    """Local Lax-Friedrichs flux; central: the same without the dissipation
    term, (F(a) + F(b)).n / 2 (advection)."""
    if phys == 1:
        lam = 0.0 if central else abs(adv[0] * n[0] + adv[1] * n[1])
        fa = np.array([adv[0] * a[0]]), np.array([adv[1] * a[0]])
        fb = np.array([adv[0] * b[0]]), np.array([adv[1] * b[0]])
    else:
        fxa, fya, pa = euler_flux(a, g)
        fxb, fyb, pb = euler_flux(b, g)
        ca, cb = np.sqrt(g * pa / a[0]), np.sqrt(g * pb / b[0])
        sa = abs((a[1] * n[0] + a[2] * n[1]) / a[0]) + ca
        sb = abs((b[1] * n[0] + b[2] * n[1]) / b[0]) + cb
        lam = max(sa, sb)
        fa, fb = (fxa, fya), (fxb, fyb)
    return 0.5 * ((fa[0] + fb[0]) * n[0] + (fa[1] + fb[1]) * n[1] + lam * (a - b))


def rhs(V, elem_s, sides, side_int, period, u, p, phys, coeff, vol_rule, face_rule, dt=1.0, g=1.4,
        adv=(1.0, 1.0), central=False):
    """vol_rule = (r, s, w) on the reference triangle (w sums to 1/2);
    face_rule = (x, w) Gauss points on [-1, 1]."""
    V = np.asarray(V).reshape(6, -1)
    K = V.shape[1]
    side_int = np.asarray(side_int).reshape(4, -1)
    elem_s = np.asarray(elem_s).reshape(3, -1)
    N, n_p = u.shape[0], u.shape[1]
    B = Basis(coeff, n_p)
    Lx, Ly = period
    verts = [np.array([[V[0, e], V[1, e]], [V[2, e], V[3, e]], [V[4, e], V[5, e]]]) for e in range(K)]

    def amap(e):
        v = verts[e]
        A = np.array([[v[1, 0] - v[0, 0], v[2, 0] - v[0, 0]], [v[1, 1] - v[0, 1], v[2, 1] - v[0, 1]]])
        return v[0], A

    def trace(e, x, shift):
        v0, A = amap(e)
        rs = np.linalg.solve(A, x - v0 - shift)
        return u[:, :, e] @ B.val(rs[0], rs[1])

    out = np.zeros_like(u)
    rq, sq, wq = vol_rule
    xf, wf = face_rule
    for e in range(K):
        v0, A = amap(e)
        J = np.linalg.det(A)
        AinvT = np.linalg.inv(A).T
        acc = np.zeros((N, n_p))
        for r, s, w in zip(rq, sq, wq):
            U = u[:, :, e] @ B.val(r, s)
            if phys == 1:
                fx, fy = np.array([adv[0] * U[0]]), np.array([adv[1] * U[0]])
            else:
                fx, fy, _ = euler_flux(U, g)
            gr, gs = B.grad(r, s)
            gx = AinvT[0, 0] * gr + AinvT[0, 1] * gs
            gy = AinvT[1, 0] * gr + AinvT[1, 1] * gs
            acc += w * J * (np.outer(fx, gx) + np.outer(fy, gy))
        v = verts[e]
        for f in range(3):
            a, b = v[f], v[(f + 1) % 3]
            d = b - a
            ln = np.hypot(d[0], d[1])
            n = np.array([d[1], -d[0]]) / ln
            s_id = elem_s[f, e]
            le, re_, ls, rs_ = side_int[:, s_id]
            nb, nbf = (re_, rs_) if le == e else (le, ls)
            vb = verts[nb]
            mid_nb = 0.5 * (vb[nbf] + vb[(nbf + 1) % 3])
            mid = 0.5 * (a + b)
            # the neighbour's copy of this face lies at mid - shift
            shift = np.zeros(2)
            if Lx > 0:
                shift[0] = Lx * np.round((mid[0] - mid_nb[0]) / Lx)
            if Ly > 0:
                shift[1] = Ly * np.round((mid[1] - mid_nb[1]) / Ly)
            for xi, wi in zip(xf, wf):
                x = mid + 0.5 * xi * d
                uin = trace(e, x, np.zeros(2))
                uout = trace(nb, x, shift)
                Fs = llf(uin, uout, n, phys, g, adv, central) if central else llf(uin, uout, n, phys, g, adv)
                rs = np.linalg.solve(A, x - v0)
                acc -= 0.5 * ln * wi * np.outer(Fs, B.val(rs[0], rs[1]))
        out[:, :, e] = dt / J * acc
    return out


def collapsed_rule(n):
    """Degree >= 2n-2 rule on the reference triangle (Duffy-collapsed Gauss)."""
    x, w = np.polynomial.legendre.leggauss(n)
    t, wt = 0.5 * (x + 1), 0.5 * w
    r, s, ww = [], [], []
    for i in range(n):
        for j in range(n):
            r.append(t[i])
            s.append(t[j] * (1 - t[i]))
            ww.append(wt[i] * wt[j] * (1 - t[i]))
    return np.array(r), np.array(s), np.array(ww)
