/*
 * dg_oracle.cpp -- FP64 serial CPU restatement of the reference's modal DG
 * conservation-law path (2D advection / Euler, LLF flux, RK4 / RK2).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is built into
 * oracle/_build/liboracle.so and loaded by tests/, by __graft_entry__.smoke()
 * and by bench.py's cpu_baseline leg.  The product library (dgcuda_amd) never
 * links, loads or calls it.
 *
 * Every function follows the reference's operation order (sums in the same
 * index order, the same association of products) so that it reproduces the
 * reference's FP64 arithmetic; the only deliberate change is loop-invariant
 * hoisting (hoist=1), which evaluates the traces, fluxes and wave speeds once
 * per quadrature point instead of once per (mode, point) pair.  It performs
 * the same operations in the same order on the same values, so it is
 * bit-for-bit neutral (checked by a test); hoist=0 keeps the reference's
 * redundant loop nest for the reference-structured CPU timing.
 *
 * Reference files followed (paths relative to the reference root):
 *   src/2d/conservation/conserv.cu:37-107            quadrature selection
 *   src/2d/basis.cu:50-125, 150-226                   phi, phi_x, phi_y, preval_basis
 *   src/2d/conservation/conserv_kernels.cu:338-514    geometry precompute
 *   src/2d/conservation/main.cu:84-113                min inscribed diameter
 *   src/2d/conservation/conserv_kernels.cu:167-187    init_conditions
 *   src/2d/conservation/euler/euler.cu:31-59          evalU0
 *   src/2d/conservation/euler/euler.cu:61-168         pressure, eval_c, eval_flux, eval_lambda
 *   src/2d/2dadvec/2dadvec_kernels.cu:150-155         advection flux (a = (ax, ay) here)
 *   src/2d/conservation/conserv_kernels.cu:628-811    eval_left_right, eval_surface, eval_volume
 *   src/2d/conservation/time_integrator.cu:130-188     eval_rhs (gather + 1/J * dt)
 *   src/2d/conservation/time_integrator.cu:194-434     rk_tempstorage, rk4, time_integrate_rk4
 *   src/2d/conservation/time_integrator.cu:445-614     rk2, time_integrate_rk2
 *   src/2d/conservation/euler/euler.cu:252-278        eval_global_lambda
 *   src/2d/conservation/euler/euler.cu:177-241        inflow_boundary, outflow_boundary
 *   src/2d/conservation/euler/supersonic/supersonic.cu:28-108  U0, U_inflow, U_outflow,
 *                                                     reflecting_boundary (supersonic vortex)
 *   src/2d/conservation/conserv_kernels.cu:818-841    eval_u (vertex values)
 *
 * Built with -O2 -ffp-contract=off and no fast-math (no FMA contraction, so each
 * product and sum is rounded as the reference source states it).
 */
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../dgcuda_amd/csrc/dg_ref_tables.inc"

namespace {

enum { PHYS_ADVECTION = 1, PHYS_EULER = 4 };
enum { QUIRK_FLOAT_SQRT = 1, QUIRK_FLOAT_GEOM = 2, QUIRK_FLOAT_POW = 4, QUIRK_REFLECT_MIRROR = 8 };
enum { IC_CONSTANT = 0, IC_GAUSSIAN = 1, IC_VORTEX = 2, IC_POLY = 3, IC_SUPERSONIC = 4 };
/* boundary side codes in right_elem (conserv_kernels.cu:647-671) */
enum { BC_REFLECTING = -1, BC_OUTFLOW = -2, BC_INFLOW = -3 };

const double PI = 3.14159265358979323846;

/* ---------------- basis (basis.cu:50-125) ---------------- */
double phi(double r, double s, int n) {
    double R[64];
    int k = 0, power = 0;
    while (k <= n) {
        for (int j = 0; j <= power; j++) { R[k] = pow(r, power - j) * pow(s, j); k++; }
        power++;
    }
    double sum = 0.;
    for (int i = 0; i <= n; i++) sum += DG_BASIS_COEFF[n][i] * R[i];
    return sum;
}

double phi_x(double r, double s, int n) {
    double R[64];
    R[0] = 0;
    int k = 1, power = 1;
    while (k <= n) {
        for (int j = 0; j <= power; j++) {
            R[k] = (power - j) * pow(r, power - j - 1) * pow(s, j);
            if (j == power) R[k] = 0;
            k++;
        }
        power++;
    }
    double sum = 0.;
    for (int i = 0; i <= n; i++) sum += DG_BASIS_COEFF[n][i] * R[i];
    return sum;
}

double phi_y(double r, double s, int n) {
    double R[64];
    R[0] = 0;
    int k = 1, power = 1;
    while (k <= n) {
        for (int j = 0; j <= power; j++) {
            R[k] = pow(r, power - j) * j * pow(s, j - 1);
            if (j == 0) R[k] = 0;
            k++;
        }
        power++;
    }
    double sum = 0.;
    for (int i = 0; i <= n; i++) sum += DG_BASIS_COEFF[n][i] * R[i];
    return sum;
}

/* float-rounded square root as the reference's sqrtf on a double argument */
inline double fsqrt_quirk(double x) { return (double)sqrtf((float)x); }

struct Ctx {
    int K = 0, S = 0, p = 0, n_p = 0, n_q = 0, n_q1 = 0, N = 0;
    int physics = 0, quirks = 0, hoist = 1;
    bool has_boundary = false;  // boundary sides: the ghost states need bc_ic
    std::vector<double> bc_ic;  // IC parameters of the inflow / outflow states
    double wall_x = 0, wall_y = 0;  // centre of the reflecting sides' radial normal
    int central = 0;                // advection: the prototype's central flux
    double gamma = 1.4, ax = 1.0, ay = 1.0;
    std::vector<double> V1x, V1y, V2x, V2y, V3x, V3y;
    std::vector<int> es1, es2, es3, left_elem, right_elem, lsn, rsn;
    std::vector<double> sx1, sy1, sx2, sy2;
    std::vector<double> J, xr, yr, xs, ys, len, Nx, Ny;
    double min_r = 0;
    std::vector<double> r1, r2, w, s_r, w1;
    std::vector<double> basis, gx, gy, side, vertex;
    std::vector<double> quad_rhs, left_rhs, right_rhs, kstar, k1, k2, k3, k4;
};

/* ---------------- quadrature selection (conserv.cu:37-107) ---------------- */
void set_quadrature(Ctx& c) {
    static const int nq[6] = {1, 3, 6, 12, 16, 25};
    c.n_q = nq[c.p];
    c.n_q1 = c.p + 1;
    const double* q2 = (c.p > 0) ? DG_QUAD2D[2 * c.p - 1] : DG_QUAD2D[0];
    c.r1.resize(c.n_q); c.r2.resize(c.n_q); c.w.resize(c.n_q);
    for (int i = 0; i < c.n_q; i++) {
        c.r1[i] = q2[3 * i];
        c.r2[i] = q2[3 * i + 1];
        c.w[i] = q2[3 * i + 2] / 2.;
    }
    c.s_r.resize(c.n_q1); c.w1.resize(c.n_q1);
    for (int i = 0; i < c.n_q1; i++) {
        c.s_r[i] = DG_QUAD1D[c.p][2 * i];
        c.w1[i] = DG_QUAD1D[c.p][2 * i + 1];
    }
}

/* ---------------- basis tables (basis.cu:150-226) ---------------- */
void preval_basis(Ctx& c) {
    const int np = c.n_p, nq = c.n_q, n1 = c.n_q1;
    c.basis.assign(nq * np, 0); c.gx.assign(nq * np, 0); c.gy.assign(nq * np, 0);
    c.side.assign(3 * n1 * np, 0); c.vertex.assign(3 * np, 0);
    for (int i = 0; i < np; i++) {
        c.vertex[i * 3 + 0] = phi(0., 0., i);
        c.vertex[i * 3 + 1] = phi(1., 0., i);
        c.vertex[i * 3 + 2] = phi(0., 1., i);
        for (int j = 0; j < nq; j++) {
            c.basis[i * nq + j] = phi(c.r1[j], c.r2[j], i);
            c.gx[i * nq + j] = c.w[j] * phi_x(c.r1[j], c.r2[j], i);
            c.gy[i * nq + j] = c.w[j] * phi_y(c.r1[j], c.r2[j], i);
        }
        for (int j = 0; j < n1; j++) {
            c.side[0 * (n1 * np) + i * n1 + j] = phi(0.5 + 0.5 * c.s_r[j], 0., i);
            c.side[1 * (n1 * np) + i * n1 + j] = phi((1. - c.s_r[j]) / 2., (1. + c.s_r[j]) / 2., i);
            c.side[2 * (n1 * np) + i * n1 + j] = phi(0., 0.5 + 0.5 * c.s_r[n1 - 1 - j], i);
        }
    }
}

/* ---------------- geometry (conserv_kernels.cu:338-514, main.cu:84-113) ---------------- */
double seg_len(double dx, double dy, bool fq) {
    if (fq) {
        float fx = (float)dx, fy = (float)dy;
        float s = fx * fx + fy * fy;
        return (double)sqrtf(s);
    }
    return sqrt(dx * dx + dy * dy);
}

void precompute_geometry(Ctx& c) {
    const bool fq = (c.quirks & QUIRK_FLOAT_GEOM) != 0;
    const int K = c.K, S = c.S;
    c.J.resize(K); c.xr.resize(K); c.yr.resize(K); c.xs.resize(K); c.ys.resize(K);
    double min_r = 0;
    for (int idx = 0; idx < K; idx++) {
        double a = seg_len(c.V1x[idx] - c.V2x[idx], c.V1y[idx] - c.V2y[idx], fq);
        double b = seg_len(c.V2x[idx] - c.V3x[idx], c.V2y[idx] - c.V3y[idx], fq);
        double cc = seg_len(c.V1x[idx] - c.V3x[idx], c.V1y[idx] - c.V3y[idx], fq);
        double k = 0.5 * (a + b + cc);
        double prod = k * (k - a) * (k - b) * (k - cc);
        double d = fq ? 2 * fsqrt_quirk(prod) / k : 2 * sqrt(prod) / k;
        if (idx == 0 || d < min_r) min_r = d;
    }
    c.min_r = min_r;
    for (int idx = 0; idx < K; idx++) {
        double x1 = c.V1x[idx], y1 = c.V1y[idx], x2 = c.V2x[idx], y2 = c.V2y[idx];
        double x3 = c.V3x[idx], y3 = c.V3y[idx];
        c.J[idx] = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1);
        c.xr[idx] = x2 - x1;
        c.yr[idx] = y2 - y1;
        c.xs[idx] = x3 - x1;
        c.ys[idx] = y3 - y1;
    }
    c.len.resize(S); c.Nx.resize(S); c.Ny.resize(S);
    for (int idx = 0; idx < S; idx++) {
        c.len[idx] = seg_len(c.sx1[idx] - c.sx2[idx], c.sy1[idx] - c.sy2[idx], fq);
        double x = c.sx2[idx] - c.sx1[idx];
        double y = c.sy2[idx] - c.sy1[idx];
        double length = seg_len(x, y, fq);
        double nx = -y / length, ny = x / length;
        int l = c.left_elem[idx];
        double tx, ty, ix, iy;
        switch (c.lsn[idx]) {
            case 0: tx = c.V3x[l]; ty = c.V3y[l]; ix = (c.V1x[l] + c.V2x[l]) / 2.; iy = (c.V1y[l] + c.V2y[l]) / 2.; break;
            case 1: tx = c.V1x[l]; ty = c.V1y[l]; ix = (c.V2x[l] + c.V3x[l]) / 2.; iy = (c.V2y[l] + c.V3y[l]) / 2.; break;
            default: tx = c.V2x[l]; ty = c.V2y[l]; ix = (c.V1x[l] + c.V3x[l]) / 2.; iy = (c.V1y[l] + c.V3y[l]) / 2.; break;
        }
        double dot = nx * (tx - ix) + ny * (ty - iy);
        if (dot > 0) { nx *= -1; ny *= -1; }
        c.Nx[idx] = nx;
        c.Ny[idx] = ny;
    }
}

/* ---------------- physics (euler.cu:61-168; 2dadvec_kernels.cu:150-155) ---------------- */
inline double pressure(const Ctx& c, const double* U) {
    double rho = U[0], rhou = U[1], rhov = U[2], E = U[3];
    return (c.gamma - 1.) * (E - (rhou * rhou + rhov * rhov) / 2. / rho);
}

inline double eval_c(const Ctx& c, const double* U) {
    double p = pressure(c, U);
    double rho = U[0];
    double a = c.gamma * p / rho;
    return (c.quirks & QUIRK_FLOAT_SQRT) ? fsqrt_quirk(a) : sqrt(a);
}

inline void eval_flux(const Ctx& c, const double* U, double* fx, double* fy) {
    if (c.physics == PHYS_ADVECTION) {
        fx[0] = c.ax * U[0];
        fy[0] = c.ay * U[0];
        return;
    }
    double p = pressure(c, U);
    double rho = U[0], rhou = U[1], rhov = U[2], E = U[3];
    fx[0] = rhou;
    fy[0] = rhov;
    fx[1] = rhou * rhou / rho + p;
    fy[1] = rhou * rhov / rho;
    fx[2] = rhou * rhov / rho;
    fy[2] = rhov * rhov / rho + p;
    fx[3] = rhou * (E + p) / rho;
    fy[3] = rhov * (E + p) / rho;
}

inline double eval_lambda(const Ctx& c, const double* UL, const double* UR, double nx, double ny) {
    /* advection: upwind |a.n|; the prototype's central flux
     * (2dadvec_kernels.cu:162-164, riemann = (u_L + u_R)/2) has no
     * dissipation term */
    if (c.physics == PHYS_ADVECTION) return c.central ? 0. : fabs(c.ax * nx + c.ay * ny);
    double cl = eval_c(c, UL), cr = eval_c(c, UR);
    double ul = UL[1] / UL[0], vl = UL[2] / UL[0];
    double ur = UR[1] / UR[0], vr = UR[2] / UR[0];
    double sl = nx * ul + ny * vl;
    double sr = nx * ur + ny * vr;
    double lm = (sl > 0.) ? sl + cl : -sl + cl;
    double rm = (sr > 0.) ? sr + cr : -sr + cr;
    return (fabs(lm) > fabs(rm)) ? fabs(lm) : fabs(rm);
}

/* ---------------- surface (conserv_kernels.cu:628-758) ---------------- */
void trace(const Ctx& c, const double* C, int face, int j, double* U) {
    const int np = c.n_p, n1 = c.n_q1, N = c.N;
    for (int n = 0; n < N; n++) U[n] = 0.;
    for (int i = 0; i < np; i++)
        for (int n = 0; n < N; n++) U[n] += C[n * np + i] * c.side[face * np * n1 + i * n1 + j];
}

void ic_point(const double* ic, int N, double gamma, double x, double y, double* U);

/* physical point of the left element's face point j (inflow_boundary,
 * euler.cu:186-204): face 0 (0.5 + 0.5 r_j, 0), face 1 ((1 - r_j)/2,
 * (1 + r_j)/2), face 2 (0, 0.5 + 0.5 r_{n1-1-j}); x = x2 r + x3 s + x1 (1 - r - s).
 * `mirror` takes face 1 as ((1 + r_j)/2, (1 - r_j)/2), as reflecting_boundary
 * does (supersonic.cu:86-89). */
void face_point(const Ctx& c, int l, int ls, int j, bool mirror, double* x, double* y) {
    const int n1 = c.n_q1;
    double r1, r2;
    switch (ls) {
        case 0: r1 = 0.5 + 0.5 * c.s_r[j]; r2 = 0.; break;
        case 1:
            if (mirror) { r1 = (1. + c.s_r[j]) / 2.; r2 = (1. - c.s_r[j]) / 2.; }
            else { r1 = (1. - c.s_r[j]) / 2.; r2 = (1. + c.s_r[j]) / 2.; }
            break;
        default: r1 = 0.; r2 = 0.5 + 0.5 * c.s_r[n1 - 1 - j]; break;
    }
    *x = c.V2x[l] * r1 + c.V3x[l] * r2 + c.V1x[l] * (1 - r1 - r2);
    *y = c.V2y[l] * r1 + c.V3y[l] * r2 + c.V1y[l] * (1 - r1 - r2);
}

/* ghost state of a boundary side (eval_left_right, conserv_kernels.cu:647-671):
 * reflecting keeps rho and E and the velocity's component along the wall,
 * whose normal is the radial direction at the point (supersonic.cu:64-108);
 * inflow and outflow take the problem's state there (euler.cu:177-241). */
void ghost_state(const Ctx& c, int code, int l, int ls, int j, double nx, double ny, const double* UL, double* UR) {
    double x, y;
    if (code == BC_REFLECTING) {
        face_point(c, l, ls, j, (c.quirks & QUIRK_REFLECT_MIRROR) != 0, &x, &y);
        /* the wall centre: the origin for the supersonic vortex, (0.5, 0.5)
         * for the cylinder (cylinder.cu:84-90) */
        x -= c.wall_x;
        y -= c.wall_y;
        UR[0] = UL[0];
        if (c.N == 1) return;
        UR[3] = UL[3];
        double dot = (c.quirks & QUIRK_FLOAT_SQRT) ? fsqrt_quirk(x * x + y * y) : sqrt(x * x + y * y);
        double Nx = x / dot, Ny = y / dot;
        if (Nx * nx + Ny * ny < 0) {
            Nx *= -1;
            Ny *= -1;
        }
        UR[1] = (UL[1] * Ny - UL[2] * Nx) * Ny;
        UR[2] = -(UL[1] * Ny - UL[2] * Nx) * Nx;
        return;
    }
    face_point(c, l, ls, j, false, &x, &y);
    ic_point(c.bc_ic.data(), c.N, c.gamma, x, y, UR);
}

void riemann_point(const Ctx& c, const double* CL, const double* CR, int ls, int rs, int j,
                   double nx, double ny, double* s_out, int l = 0, int code = 0) {
    const int N = c.N, n1 = c.n_q1;
    double UL[4] = {0, 0, 0, 0}, UR[4] = {0, 0, 0, 0}, fxl[4], fyl[4], fxr[4], fyr[4];
    trace(c, CL, ls, j, UL);
    if (code < 0)
        ghost_state(c, code, l, ls, j, nx, ny, UL, UR);
    else
        trace(c, CR, rs, n1 - j - 1, UR);
    eval_flux(c, UL, fxl, fyl);
    eval_flux(c, UR, fxr, fyr);
    double lambda = eval_lambda(c, UL, UR, nx, ny);
    for (int n = 0; n < N; n++)
        s_out[n] = 0.5 * ((fxl[n] + fxr[n]) * nx + (fyl[n] + fyr[n]) * ny + lambda * (UL[n] - UR[n]));
}

void eval_surface(Ctx& c, const double* u) {
    const int np = c.n_p, n1 = c.n_q1, N = c.N, K = c.K, S = c.S;
    std::vector<double> CL(N * np), CR(N * np), sj(n1 * 4);
    for (int idx = 0; idx < S; idx++) {
        int l = c.left_elem[idx], r = c.right_elem[idx];
        int ls = c.lsn[idx], rs = c.rsn[idx];
        double nx = c.Nx[idx], ny = c.Ny[idx], len = c.len[idx];
        const int code = r < 0 ? r : 0;
        for (int i = 0; i < np; i++)
            for (int n = 0; n < N; n++) {
                CL[n * np + i] = u[K * np * n + i * K + l];
                CR[n * np + i] = r >= 0 ? u[K * np * n + i * K + r] : 0.;
            }
        if (r < 0) rs = 0;
        if (c.hoist)
            for (int j = 0; j < n1; j++)
                riemann_point(c, CL.data(), CR.data(), ls, rs, j, nx, ny, &sj[j * 4], l, code);
        for (int i = 0; i < np; i++) {
            double sum_left[4] = {0, 0, 0, 0}, sum_right[4] = {0, 0, 0, 0};
            for (int j = 0; j < n1; j++) {
                double s_loc[4];
                const double* s;
                if (c.hoist) {
                    s = &sj[j * 4];
                } else {
                    riemann_point(c, CL.data(), CR.data(), ls, rs, j, nx, ny, s_loc, l, code);
                    s = s_loc;
                }
                for (int n = 0; n < N; n++) {
                    sum_left[n] += c.w1[j] * s[n] * c.side[ls * np * n1 + i * n1 + j];
                    sum_right[n] += c.w1[j] * s[n] * c.side[rs * np * n1 + i * n1 + n1 - 1 - j];
                }
            }
            for (int n = 0; n < N; n++) {
                c.left_rhs[S * np * n + i * S + idx] = -len / 2. * sum_left[n];
                c.right_rhs[S * np * n + i * S + idx] = len / 2. * sum_right[n];
            }
        }
    }
}

/* ---------------- volume (conserv_kernels.cu:766-811) ---------------- */
void volume_point(const Ctx& c, const double* C, int j, double* fx, double* fy) {
    const int np = c.n_p, nq = c.n_q, N = c.N;
    double U[4];
    for (int n = 0; n < N; n++) U[n] = 0.;
    for (int k = 0; k < np; k++)
        for (int n = 0; n < N; n++) U[n] += C[n * np + k] * c.basis[nq * k + j];
    eval_flux(c, U, fx, fy);
}

void eval_volume(Ctx& c, const double* u) {
    const int np = c.n_p, nq = c.n_q, N = c.N, K = c.K;
    std::vector<double> C(N * np), FX(nq * 4), FY(nq * 4);
    for (int idx = 0; idx < K; idx++) {
        for (int i = 0; i < np; i++)
            for (int n = 0; n < N; n++) C[n * np + i] = u[K * np * n + i * K + idx];
        double x_r = c.xr[idx], y_r = c.yr[idx], x_s = c.xs[idx], y_s = c.ys[idx];
        if (c.hoist)
            for (int j = 0; j < nq; j++) volume_point(c, C.data(), j, &FX[j * 4], &FY[j * 4]);
        for (int i = 0; i < np; i++) {
            double sum[4] = {0, 0, 0, 0};
            for (int j = 0; j < nq; j++) {
                double fxl[4], fyl[4];
                const double *fx, *fy;
                if (c.hoist) {
                    fx = &FX[j * 4]; fy = &FY[j * 4];
                } else {
                    volume_point(c, C.data(), j, fxl, fyl);
                    fx = fxl; fy = fyl;
                }
                double bgx = c.gx[nq * i + j], bgy = c.gy[nq * i + j];
                for (int n = 0; n < N; n++)
                    sum[n] += fx[n] * (bgx * y_s - bgy * y_r) + fy[n] * (-bgx * x_s + bgy * x_r);
            }
            for (int n = 0; n < N; n++) c.quad_rhs[K * np * n + i * K + idx] = sum[n];
        }
    }
}

/* ---------------- gather + inverse mass (time_integrator.cu:130-188) ---------------- */
void eval_rhs_gather(Ctx& c, double* k, double dt) {
    const int np = c.n_p, N = c.N, K = c.K, S = c.S;
    for (int idx = 0; idx < K; idx++) {
        double j = c.J[idx];
        int s_idx[3] = {c.es1[idx], c.es2[idx], c.es3[idx]};
        for (int i = 0; i < np; i++) {
            double se[3][4];
            for (int f = 0; f < 3; f++) {
                const std::vector<double>& arr = (idx == c.left_elem[s_idx[f]]) ? c.left_rhs : c.right_rhs;
                for (int n = 0; n < N; n++) se[f][n] = arr[S * np * n + i * S + s_idx[f]];
            }
            for (int n = 0; n < N; n++)
                k[K * np * n + i * K + idx] =
                    1. / j * dt * (c.quad_rhs[K * np * n + i * K + idx] + se[0][n] + se[1][n] + se[2][n]);
        }
    }
}

void rhs(Ctx& c, const double* u, double* k, double dt) {
    eval_surface(c, u);
    eval_volume(c, u);
    eval_rhs_gather(c, k, dt);
}

/* ---------------- CFL (euler.cu:252-278, time_integrator.cu:254-263) ---------------- */
double cell_lambda(const Ctx& c, const double* u, int idx) {
    const int K = c.K, np = c.n_p;
    if (c.physics == PHYS_ADVECTION) {
        double a = c.ax * c.ax + c.ay * c.ay;
        return (c.quirks & QUIRK_FLOAT_SQRT) ? fsqrt_quirk(a) : sqrt(a);
    }
    double U[4];
    for (int n = 0; n < 4; n++) U[n] = u[K * np * n + idx] * c.basis[0];
    double cs = eval_c(c, U);
    double m2 = U[1] * U[1] + U[2] * U[2];
    double s = ((c.quirks & QUIRK_FLOAT_SQRT) ? fsqrt_quirk(m2) : sqrt(m2)) / U[0];
    return (s > 0) ? s + cs : -s + cs;
}

double max_lambda(const Ctx& c, const double* u) {
    double m = cell_lambda(c, u, 0);
    for (int i = 0; i < c.K; i++) {
        double l = cell_lambda(c, u, i);
        m = (l > m) ? l : m;
    }
    return m;
}

double cfl_dt(const Ctx& c, const double* u, double cfl) {
    return cfl * c.min_r / max_lambda(c, u) / (2. * c.p + 1.);
}

/* ---------------- RK (time_integrator.cu:194-218, 276-429, 445-578) ---------------- */
void rk_step(Ctx& c, double* u, double dt, int scheme) {
    const size_t M = (size_t)c.N * c.n_p * c.K;
    if (scheme == 2) {
        rhs(c, u, c.k1.data(), dt);
        for (size_t i = 0; i < M; i++) c.kstar[i] = u[i] + 0.5 * c.k1[i];
        rhs(c, c.kstar.data(), c.k1.data(), dt);
        for (size_t i = 0; i < M; i++) u[i] += c.k1[i];
        return;
    }
    rhs(c, u, c.k1.data(), dt);
    for (size_t i = 0; i < M; i++) c.kstar[i] = u[i] + 0.5 * c.k1[i];
    rhs(c, c.kstar.data(), c.k2.data(), dt);
    for (size_t i = 0; i < M; i++) c.kstar[i] = u[i] + 0.5 * c.k2[i];
    rhs(c, c.kstar.data(), c.k3.data(), dt);
    for (size_t i = 0; i < M; i++) c.kstar[i] = u[i] + 1.0 * c.k3[i];
    rhs(c, c.kstar.data(), c.k4.data(), dt);
    for (size_t i = 0; i < M; i++) u[i] += c.k1[i] / 6. + c.k2[i] / 3. + c.k3[i] / 3. + c.k4[i] / 6.;
}

/* ---------------- initial conditions (conserv_kernels.cu:167-187, euler.cu:31-59) ---------------- */
inline double wrap(double d, double L) { return (L > 0) ? d - L * nearbyint(d / L) : d; }

void ic_point(const double* ic, int N, double gamma, double x, double y, double* U) {
    int kind = (int)ic[0];
    if (kind == IC_CONSTANT) {
        for (int n = 0; n < N; n++) U[n] = ic[1 + n];
        return;
    }
    if (kind == IC_POLY) {
        /* (x - y)^alpha in every variable: the reference's own projection test
         * (src/2d/test_all.cu:21-51, u = (x - y)^alpha) */
        for (int n = 0; n < N; n++) U[n] = pow(x - y, ic[1]);
        return;
    }
    if (kind == IC_GAUSSIAN) {
        double dx = wrap(x - ic[1], ic[4]), dy = wrap(y - ic[2], ic[5]);
        U[0] = exp(-ic[3] * (dx * dx + dy * dy));
        return;
    }
    if (kind == IC_SUPERSONIC) {
        /* supersonic vortex (supersonic.cu:28-38): ic = {4, mach, float_pow}.
         * float_pow rounds each power's arguments and result to float, as the
         * reference's powf on double arguments does; otherwise FP64 pow. */
        const double mach = ic[1];
        const bool fp = ic[2] != 0.;
        auto pw = [fp](double a, double b) {
            return fp ? (double)(float)pow((double)(float)a, (double)(float)b) : pow(a, b);
        };
        double r = sqrt(x * x + y * y);
        U[0] = pw(1 + 1.0125 * (1. - 1. / (r * r)), 2.5);
        if (N == 1) return;
        U[1] = U[0] * sin(atan(y / x)) * mach / r;
        U[2] = U[0] * -cos(atan(y / x)) * mach / r;
        double p = (1.0 / gamma) * pw(U[0], gamma);
        U[3] = 0.5 * U[0] * (mach * mach / (r * r)) + p * (1. / (gamma - 1.));
        return;
    }
    /* isentropic vortex superposition: ic = {2, nv, rho_inf, u_inf, v_inf, p_inf, Lx, Ly, (beta, xc, yc) x nv} */
    int nv = (int)ic[1];
    double rho_inf = ic[2], u_inf = ic[3], v_inf = ic[4], p_inf = ic[5], Lx = ic[6], Ly = ic[7];
    double du = 0, dv = 0, dT = 0;
    for (int k = 0; k < nv; k++) {
        double beta = ic[8 + 3 * k], xc = ic[9 + 3 * k], yc = ic[10 + 3 * k];
        double dx = wrap(x - xc, Lx), dy = wrap(y - yc, Ly);
        double r2 = dx * dx + dy * dy;
        double e = exp(0.5 * (1. - r2));
        du += -beta / (2. * PI) * dy * e;
        dv += beta / (2. * PI) * dx * e;
        dT += -(gamma - 1.) * beta * beta / (8. * gamma * PI * PI) * e * e;
    }
    double T_inf = p_inf / rho_inf;
    double T = T_inf + dT;
    double rho = rho_inf * pow(T / T_inf, 1. / (gamma - 1.));
    double p = p_inf * pow(rho / rho_inf, gamma);
    double u = u_inf + du, v = v_inf + dv;
    U[0] = rho;
    U[1] = rho * u;
    U[2] = rho * v;
    U[3] = p / (gamma - 1.) + 0.5 * rho * (u * u + v * v);
}

void project_ic(const Ctx& c, const double* ic, double* u) {
    const int np = c.n_p, nq = c.n_q, N = c.N, K = c.K;
    for (int idx = 0; idx < K; idx++) {
        double v1x = c.V1x[idx], v1y = c.V1y[idx], v2x = c.V2x[idx], v2y = c.V2y[idx];
        double v3x = c.V3x[idx], v3y = c.V3y[idx];
        for (int i = 0; i < np; i++) {
            double U[4] = {0, 0, 0, 0};
            for (int j = 0; j < nq; j++) {
                double x = c.r1[j] * v2x + c.r2[j] * v3x + (1 - c.r1[j] - c.r2[j]) * v1x;
                double y = c.r1[j] * v2y + c.r2[j] * v3y + (1 - c.r1[j] - c.r2[j]) * v1y;
                double u0[4];
                ic_point(ic, N, c.gamma, x, y, u0);
                for (int n = 0; n < N; n++) U[n] += c.w[j] * u0[n] * c.basis[i * nq + j];
            }
            for (int n = 0; n < N; n++) u[K * np * n + i * K + idx] = U[n];
        }
    }
}

/* a mesh with boundary sides needs the inflow / outflow IC */
bool bc_ready(const Ctx& c) { return !c.has_boundary || !c.bc_ic.empty(); }

}  // namespace

extern "C" {

/* Mesh arrays in the reference's read_mesh layout (conserv.cu:118-192):
 *   V[6*K]      = V1x | V1y | V2x | V2y | V3x | V3y
 *   elem_s[3*K] = elem_s1 | elem_s2 | elem_s3
 *   sides[4*S]  = sides_x1 | sides_y1 | sides_x2 | sides_y2
 *   side_int[4*S] = left_elem | right_elem | left_side_number | right_side_number
 * phys = {gamma, ax, ay}.  Returns NULL for p outside 0..5.  Boundary sides
 * (right_elem -1 reflecting, -2 outflow, -3 inflow) take ghost states; the
 * inflow / outflow states come from the IC given to or_set_bc_ic, without
 * which or_rhs / or_rk_step / or_integrate refuse such a mesh. */
void* or_create(int K, int S, const double* V, const int* elem_s, const double* sides, const int* side_int,
                int p, int physics, const double* phys, int quirks, int hoist) {
    if (p < 0 || p > 5 || (physics != PHYS_ADVECTION && physics != PHYS_EULER)) return nullptr;
    Ctx* c = new Ctx;
    for (int s = 0; s < S; s++)
        if (side_int[S + s] < 0) c->has_boundary = true;
    c->K = K; c->S = S; c->p = p; c->n_p = (p + 1) * (p + 2) / 2;
    c->N = physics; c->physics = physics; c->quirks = quirks; c->hoist = hoist;
    c->gamma = phys[0]; c->ax = phys[1]; c->ay = phys[2];
    c->V1x.assign(V, V + K); c->V1y.assign(V + K, V + 2 * K);
    c->V2x.assign(V + 2 * K, V + 3 * K); c->V2y.assign(V + 3 * K, V + 4 * K);
    c->V3x.assign(V + 4 * K, V + 5 * K); c->V3y.assign(V + 5 * K, V + 6 * K);
    c->es1.assign(elem_s, elem_s + K); c->es2.assign(elem_s + K, elem_s + 2 * K);
    c->es3.assign(elem_s + 2 * K, elem_s + 3 * K);
    c->sx1.assign(sides, sides + S); c->sy1.assign(sides + S, sides + 2 * S);
    c->sx2.assign(sides + 2 * S, sides + 3 * S); c->sy2.assign(sides + 3 * S, sides + 4 * S);
    c->left_elem.assign(side_int, side_int + S); c->right_elem.assign(side_int + S, side_int + 2 * S);
    c->lsn.assign(side_int + 2 * S, side_int + 3 * S); c->rsn.assign(side_int + 3 * S, side_int + 4 * S);
    set_quadrature(*c);
    preval_basis(*c);
    precompute_geometry(*c);
    size_t M = (size_t)c->N * c->n_p * K, MS = (size_t)c->N * c->n_p * S;
    c->quad_rhs.assign(M, 0); c->left_rhs.assign(MS, 0); c->right_rhs.assign(MS, 0);
    c->kstar.assign(M, 0); c->k1.assign(M, 0); c->k2.assign(M, 0); c->k3.assign(M, 0); c->k4.assign(M, 0);
    return c;
}

void or_destroy(void* h) { delete (Ctx*)h; }

/* IC parameters (the or_project_ic array) of the inflow / outflow states */
void or_set_bc_ic(void* h, const double* ic, int n) { ((Ctx*)h)->bc_ic.assign(ic, ic + n); }

/* advection flux: 0 local Lax-Friedrichs, 1 the prototype's central flux */
void or_set_flux(void* h, int central) { ((Ctx*)h)->central = central; }

/* centre of the reflecting sides' radial wall normal */
void or_set_wall(void* h, double x, double y) {
    ((Ctx*)h)->wall_x = x;
    ((Ctx*)h)->wall_y = y;
}

void or_set_hoist(void* h, int hoist) { ((Ctx*)h)->hoist = hoist; }

double or_phi(double r, double s, int i) { return phi(r, s, i); }
double or_phi_x(double r, double s, int i) { return phi_x(r, s, i); }
double or_phi_y(double r, double s, int i) { return phi_y(r, s, i); }

/* sizes[0..3] = n_p, n_q, n_q1, N */
void or_sizes(void* h, int* sizes) {
    Ctx* c = (Ctx*)h;
    sizes[0] = c->n_p; sizes[1] = c->n_q; sizes[2] = c->n_q1; sizes[3] = c->N;
}

/* r1,r2,w [n_q]; s_r,w1 [n_q1]; basis,gx,gy [n_p*n_q]; side [3*n_p*n_q1]; vertex [3*n_p] */
void or_tables(void* h, double* r1, double* r2, double* w, double* s_r, double* w1, double* basis, double* gx,
               double* gy, double* side, double* vertex) {
    Ctx* c = (Ctx*)h;
    memcpy(r1, c->r1.data(), c->n_q * 8); memcpy(r2, c->r2.data(), c->n_q * 8);
    memcpy(w, c->w.data(), c->n_q * 8); memcpy(s_r, c->s_r.data(), c->n_q1 * 8);
    memcpy(w1, c->w1.data(), c->n_q1 * 8);
    memcpy(basis, c->basis.data(), c->basis.size() * 8); memcpy(gx, c->gx.data(), c->gx.size() * 8);
    memcpy(gy, c->gy.data(), c->gy.size() * 8); memcpy(side, c->side.data(), c->side.size() * 8);
    memcpy(vertex, c->vertex.data(), c->vertex.size() * 8);
}

/* geo_e[5*K] = J | xr | yr | xs | ys ; geo_s[3*S] = len | Nx | Ny ; returns min_r */
double or_geometry(void* h, double* geo_e, double* geo_s) {
    Ctx* c = (Ctx*)h;
    const int K = c->K, S = c->S;
    if (geo_e) {
        memcpy(geo_e, c->J.data(), K * 8); memcpy(geo_e + K, c->xr.data(), K * 8);
        memcpy(geo_e + 2 * K, c->yr.data(), K * 8); memcpy(geo_e + 3 * K, c->xs.data(), K * 8);
        memcpy(geo_e + 4 * K, c->ys.data(), K * 8);
    }
    if (geo_s) {
        memcpy(geo_s, c->len.data(), S * 8); memcpy(geo_s + S, c->Nx.data(), S * 8);
        memcpy(geo_s + 2 * S, c->Ny.data(), S * 8);
    }
    return c->min_r;
}

void or_project_ic(void* h, const double* ic, double* u) { project_ic(*(Ctx*)h, ic, u); }

void or_ic_point(const double* ic, int N, double gamma, double x, double y, double* U) {
    ic_point(ic, N, gamma, x, y, U);
}

/* k = dt * M^-1 R(u) with M = J*I, in the reference layout [var][mode][elem] */
int or_rhs(void* h, const double* u, double* k, double dt) {
    if (!bc_ready(*(Ctx*)h)) return -1;
    rhs(*(Ctx*)h, u, k, dt);
    return 0;
}

/* reference-side split of the RHS: quad_rhs [N*n_p*K], left/right side arrays [N*n_p*S] */
int or_rhs_parts(void* h, const double* u, double* quad, double* left, double* right) {
    Ctx* c = (Ctx*)h;
    if (!bc_ready(*c)) return -1;
    eval_surface(*c, u);
    eval_volume(*c, u);
    size_t M = (size_t)c->N * c->n_p * c->K, MS = (size_t)c->N * c->n_p * c->S;
    memcpy(quad, c->quad_rhs.data(), M * 8);
    memcpy(left, c->left_rhs.data(), MS * 8);
    memcpy(right, c->right_rhs.data(), MS * 8);
    return 0;
}

void or_cell_lambda(void* h, const double* u, double* lam) {
    Ctx* c = (Ctx*)h;
    for (int i = 0; i < c->K; i++) lam[i] = cell_lambda(*c, u, i);
}

double or_cfl_dt(void* h, const double* u, double cfl) { return cfl_dt(*(Ctx*)h, u, cfl); }

int or_rk_step(void* h, double* u, double dt, int scheme) {
    if (!bc_ready(*(Ctx*)h)) return -1;
    rk_step(*(Ctx*)h, u, dt, scheme);
    return 0;
}

/* time_integrate_rk4 / rk2 loop: dt from the CFL condition each step, t
 * advanced before the stages, last step clamped to t_end.  Returns steps. */
int or_integrate(void* h, double* u, double t_end, double cfl, int scheme, double* t_out) {
    Ctx* c = (Ctx*)h;
    if (!bc_ready(*c)) return -1;
    double t = 0;
    int steps = 0;
    while (t < t_end) {
        double dt = cfl_dt(*c, u, cfl);
        if (t + dt > t_end) {
            dt = t_end - t;
            t = t_end;
        } else {
            t += dt;
        }
        rk_step(*c, u, dt, scheme);
        steps++;
    }
    if (t_out) *t_out = t;
    return steps;
}

/* nsteps CFL steps without an end time; returns the final time */
double or_run_steps(void* h, double* u, int nsteps, double cfl, int scheme) {
    Ctx* c = (Ctx*)h;
    if (!bc_ready(*c)) return -1.0;
    double t = 0;
    for (int s = 0; s < nsteps; s++) {
        double dt = cfl_dt(*c, u, cfl);
        t += dt;
        rk_step(*c, u, dt, scheme);
    }
    return t;
}

/* uv[3*K]: values at the three vertices of each element (eval_u) */
void or_eval_vertices(void* h, const double* u, int var, double* uv) {
    Ctx* c = (Ctx*)h;
    const int K = c->K, np = c->n_p;
    for (int idx = 0; idx < K; idx++) {
        double a = 0, b = 0, d = 0;
        for (int i = 0; i < np; i++) {
            double ci = u[K * np * var + i * K + idx];
            a += ci * c->vertex[i * 3 + 0];
            b += ci * c->vertex[i * 3 + 1];
            d += ci * c->vertex[i * 3 + 2];
        }
        uv[3 * idx] = a; uv[3 * idx + 1] = b; uv[3 * idx + 2] = d;
    }
}

/* mass[n] = sum_K J_K * c0 * phi_0 / 2 (integral of each conserved variable), long double sum */
void or_mass(void* h, const double* u, double* mass) {
    Ctx* c = (Ctx*)h;
    const int K = c->K, np = c->n_p;
    for (int n = 0; n < c->N; n++) {
        long double s = 0;
        for (int idx = 0; idx < K; idx++)
            s += (long double)c->J[idx] * u[K * np * n + idx] * DG_BASIS_COEFF[0][0] * 0.5L;
        mass[n] = (double)s;
    }
}

}  // extern "C"
