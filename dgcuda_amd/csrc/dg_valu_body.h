// One element per lane (VALU): the fused stage body shared by the
// single-solve VALU kernel and the batched kernel.
//
// One call does what the reference does with four launches
// (eval_surface, eval_volume, eval_rhs, rk_tempstorage;
// src/2d/conservation/time_integrator.cu:274-307):
//   volume integral    (conserv_kernels.cu:766-811)
//   LLF surface terms  (conserv_kernels.cu:628-758) evaluated for the three
//                      faces of the element, each face flux computed in the
//                      side's canonical left->right orientation so both
//                      elements of a side obtain bit-identical fluxes
//                      (race-free, no atomics, no face buffers)
//   gather + M^-1      (time_integrator.cu:130-188, M = J*I)
//   RK stage update    (time_integrator.cu:194-218, 445-464; RK4 in stage
//                      states, rk_update)
//   CFL wave speed     (euler.cu:252-278) of the new state on the last stage.
// Euler up to p = 3 uses the normal-flux forms of the matrix-core kernels
// (fewer operations and registers); otherwise the reference's operation
// order is kept.
#pragma once
#include "dg_kernels.h"

namespace dgmi {

// phi on face f at point index j where j is a runtime value in [0, NQ1)
// given as the compile-time pair (j, NQ1-1-j) selected by `rev`
template <typename R, int P>
__device__ __forceinline__ R face_phi(const TabP<R, P>& T, int f, int m, int j, bool rev) {
    constexpr int NQ1 = P + 1;
    return rev ? T.S[f][m][NQ1 - 1 - j] : T.S[f][m][j];
}

// Processes element e (clamped for loads when invalid, no stores then);
// returns its cell wave speed after the last stage of a step, else 0.
// BND: the variant for meshes with boundary sides (p <= VALU_BOUNDARY_MAX_P);
// without it the face loop carries no branch and the loads of the three
// faces can be scheduled together.
template <int N, int P, typename R, bool BND = false>
__device__ __forceinline__ double valu_process(const StageArgs<R>& a, int e, int kind) {
    static_assert(!BND || P <= VALU_BOUNDARY_MAX_P, "boundary sides are compiled for p <= VALU_BOUNDARY_MAX_P");
    constexpr int NP = n_modes(P), NQ = n_quad_vol(P), NQ1 = P + 1;
    // Euler up to p = 3 uses the normal-flux forms of the matrix-core
    // kernels (fewer operations and registers; p >= 4 keeps its allocation)
    constexpr bool NF_FORM = N == 4 && P <= 3;
    const TabP<R, P>& T = tab<R, P>();
    const bool valid = e < a.K;
    const int ee = valid ? e : a.K - 1;
    const int Kp = a.Kp;
    const R dt = (R)a.ctrl->dt;
    R u[N][NP];
#pragma unroll
    for (int v = 0; v < N; v++)
#pragma unroll
        for (int m = 0; m < NP; m++) u[v][m] = gp(a.u)[(size_t)(v * NP + m) * Kp + ee];

    R tot[N][NP];
#pragma unroll
    for (int v = 0; v < N; v++)
#pragma unroll
        for (int m = 0; m < NP; m++) tot[v][m] = R(0);

    // The neighbour words and the first face's neighbour modes are requested
    // before the volume integral, whose arithmetic covers their latency (the
    // kernel waits on its loads for about half of its time; 1-11% faster at
    // p = 1, 2).  The BND variant keeps its loads in the face loop.
    int wf[3];
    R nu[N][NP];
    if constexpr (!BND) {
#pragma unroll
        for (int f = 0; f < 3; f++) wf[f] = gp(a.nbr)[f * Kp + ee];
        const int nb0 = wf[0] & NB_ELEM_MASK;
#pragma unroll
        for (int v = 0; v < N; v++)
#pragma unroll
            for (int m = 0; m < NP; m++) nu[v][m] = gp(a.u)[(size_t)(v * NP + m) * Kp + nb0];
    }

    // ---- volume: sum_q F(U_q) . (metric * grad phi_m) ----
    {
        const R xr = gp(a.geo_e)[ee], yr = gp(a.geo_e)[Kp + ee], xs = gp(a.geo_e)[2 * Kp + ee], ys = gp(a.geo_e)[3 * Kp + ee];
#pragma unroll 2
        for (int q = 0; q < NQ; q++) {
            R U[N];
#pragma unroll
            for (int v = 0; v < N; v++) U[v] = R(0);
#pragma unroll
            for (int m = 0; m < NP; m++)
#pragma unroll
                for (int v = 0; v < N; v++) U[v] += u[v][m] * T.B[m][q];
            R fx[N], fy[N];
            if constexpr (NF_FORM) {
                // Euler: fluxes contracted with the inverse map first
                // (metric_flux), then with the weighted basis gradients
                metric_flux<R>(a.phys, U, xr, yr, xs, ys, fx, fy);
#pragma unroll
                for (int m = 0; m < NP; m++) {
                    const R bgx = T.Gx[m][q], bgy = T.Gy[m][q];
#pragma unroll
                    for (int v = 0; v < N; v++) tot[v][m] += fx[v] * bgx + fy[v] * bgy;
                }
            } else {
                flux<N, R>(a.phys, U, fx, fy);
#pragma unroll
                for (int m = 0; m < NP; m++) {
                    const R bgx = T.Gx[m][q], bgy = T.Gy[m][q];
#pragma unroll
                    for (int v = 0; v < N; v++)
                        tot[v][m] += fx[v] * (bgx * ys - bgy * yr) + fy[v] * (-bgx * xs + bgy * xr);
                }
            }
        }
    }

    // ---- the three faces, added in face order as eval_rhs adds s1, s2, s3 ----
#pragma unroll
    for (int f = 0; f < 3; f++) {
        int w;
        if constexpr (BND) {
            w = gp(a.nbr)[f * Kp + ee];
        } else {
            w = wf[f];
            if (f > 0) {
                const int nbf = w & NB_ELEM_MASK;
#pragma unroll
                for (int v = 0; v < N; v++)
#pragma unroll
                    for (int m = 0; m < NP; m++) nu[v][m] = gp(a.u)[(size_t)(v * NP + m) * Kp + nbf];
            }
        }
        // boundary side (BND variant): no neighbour, a ghost state
        bool bnd = false;
        int nb = w & NB_ELEM_MASK;
        if constexpr (BND) {
            bnd = w < 0;
            if (bnd) nb = ee;
        }
        const int fn = (w >> 28) & 3;
        const bool left = (w >> 30) & 1;
        const R nx = gp(a.geo_f)[(3 * f + 0) * Kp + ee];
        const R ny = gp(a.geo_f)[(3 * f + 1) * Kp + ee];
        const R len = gp(a.geo_f)[(3 * f + 2) * Kp + ee];
        R FS[NQ1][N];
#pragma unroll
        for (int j = 0; j < NQ1; j++) {
            // canonical point j of the side: the left element's trace at its
            // point j, the right element's at NQ1-1-j
            R Uo[N], Un[N];
#pragma unroll
            for (int v = 0; v < N; v++) {
                Uo[v] = R(0);
                Un[v] = R(0);
            }
#pragma unroll
            for (int m = 0; m < NP; m++) {
                const R so = face_phi<R, P>(T, f, m, j, !left);
                const R s0 = face_phi<R, P>(T, 0, m, j, left);
                const R s1 = face_phi<R, P>(T, 1, m, j, left);
                const R s2 = face_phi<R, P>(T, 2, m, j, left);
                const R sn = (fn == 0) ? s0 : ((fn == 1) ? s1 : s2);
#pragma unroll
                for (int v = 0; v < N; v++) {
                    Uo[v] += u[v][m] * so;
                    if constexpr (BND)
                        Un[v] += gp(a.u)[(size_t)(v * NP + m) * Kp + nb] * sn;
                    else
                        Un[v] += nu[v][m] * sn;
                }
            }
            if constexpr (BND)
                if (bnd) boundary_ghost<N, NQ1, R>(a.bnd, w, j, Uo, Un);
            if constexpr (NF_FORM) {
                // Euler: normal fluxes of both sides, the side's sign in
                // the speed (the matrix-core kernel's form; both elements of
                // the side still get bit-identical values)
                R Fo[N], Fb[N], so, sb;
                normal_flux_speed<R>(a.phys, Uo, nx, ny, Fo, so);
                normal_flux_speed<R>(a.phys, Un, nx, ny, Fb, sb);
                const R lm = fmax(so, sb);
                const R lms = left ? lm : -lm;
#pragma unroll
                for (int v = 0; v < N; v++) FS[j][v] = T.W1[j] * (R(0.5) * fma(lms, Uo[v] - Un[v], Fo[v] + Fb[v]));
            } else {
                R UL[N], UR[N];
#pragma unroll
                for (int v = 0; v < N; v++) {
                    UL[v] = left ? Uo[v] : Un[v];
                    UR[v] = left ? Un[v] : Uo[v];
                }
                R fxl[N], fyl[N], fxr[N], fyr[N];
                flux<N, R>(a.phys, UL, fxl, fyl);
                flux<N, R>(a.phys, UR, fxr, fyr);
                const R lam = llf_speed<N, R>(a.phys, UL, UR, nx, ny);
#pragma unroll
                for (int v = 0; v < N; v++)
                    FS[j][v] = T.W1[j] *
                               (R(0.5) * ((fxl[v] + fxr[v]) * nx + (fyl[v] + fyr[v]) * ny + lam * (UL[v] - UR[v])));
            }
        }
        const R coef = left ? -len / R(2) : len / R(2);
#pragma unroll
        for (int m = 0; m < NP; m++)
#pragma unroll
            for (int v = 0; v < N; v++) {
                R s = R(0);
#pragma unroll
                for (int j = 0; j < NQ1; j++) s += FS[j][v] * face_phi<R, P>(T, f, m, j, !left);
                tot[v][m] += coef * s;
            }
    }

    // ---- inverse mass (M = J*I; 1/J precomputed in FP64) and the RK stage update ----
    const R cj = gp(a.geo_e)[4 * Kp + ee] * dt;
    double lam = 0.0;
    if (kind == ST_RHS) {
        if (valid)
#pragma unroll
            for (int v = 0; v < N; v++)
#pragma unroll
                for (int m = 0; m < NP; m++) gp(a.out)[(size_t)(v * NP + m) * Kp + e] = cj * tot[v][m];
    } else {
        // RK4 in stage states (rk_update): per stage the update's operands
        // are c (the input u, or u_n read) and aa (the input u, or W read).
        // They are loaded per variable before its first store: the output
        // can be the base state itself (the last stage writes u_n+1 over
        // u_n), so a load after a store could not be moved above it and
        // each would wait for its own round trip to memory.  One round trip
        // per variable instead of one per mode (grouping all variables made
        // the batched kernel spill).
        const bool c_in = (kind == ST_RK4_1 || kind == ST_RK4_4 || kind == ST_RK2_1);
        const bool aa_in = (kind == ST_RK4_2);
        const bool load_c = (kind == ST_RK4_2 || kind == ST_RK4_3 || kind == ST_RK2_2);
        const bool load_acc = (kind == ST_RK4_4);
        R c0[N];
#pragma unroll
        for (int v = 0; v < N; v++) {
            R cv[NP], av[NP];
#pragma unroll
            for (int m = 0; m < NP; m++) {
                const size_t ix = (size_t)(v * NP + m) * Kp + ee;
                cv[m] = c_in ? u[v][m] : (load_c ? gp(a.c)[ix] : R(0));
                av[m] = aa_in ? u[v][m] : (load_acc ? gp(a.acc)[ix] : R(0));
            }
#pragma unroll
            for (int m = 0; m < NP; m++) {
                const size_t ix = (size_t)(v * NP + m) * Kp + ee;
                const R k = cj * tot[v][m];
                R o, w = R(0);
                switch (kind) {
                    case ST_RK4_1: o = rk_update<ST_RK4_1, R>(cv[m], av[m], k, &w); break;
                    case ST_RK4_2: o = rk_update<ST_RK4_2, R>(cv[m], av[m], k, &w); break;
                    case ST_RK4_3: o = rk_update<ST_RK4_3, R>(cv[m], av[m], k, &w); break;
                    case ST_RK4_4: o = rk_update<ST_RK4_4, R>(cv[m], av[m], k, &w); break;
                    case ST_RK2_1: o = rk_update<ST_RK2_1, R>(cv[m], av[m], k, &w); break;
                    default: o = rk_update<ST_RK2_2, R>(cv[m], av[m], k, &w); break;
                }
                if (valid && kind == ST_RK4_2) gp(a.acc)[ix] = w;
                if (valid) gp(a.out)[ix] = o;
                if (m == 0) c0[v] = o;
            }
        }
        if (kind == ST_RK4_4 || kind == ST_RK2_2) {
            double U[4];
            const double b0 = tab<double, P>().B[0][0];
#pragma unroll
            for (int v = 0; v < N; v++) U[v] = (double)c0[v] * b0;
            lam = valid ? cell_speed<N>(a.phys, U) : 0.0;
        }
    }
    return lam;
}

}  // namespace dgmi
