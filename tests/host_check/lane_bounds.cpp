// Host build of the lane-per-element stage body (valu_process, the body of
// k_stage_valu) with the boundary path compiled in at every order, run over
// every element of a mesh with buffers of exactly the sizes the solver
// allocates on the device (built with -fsanitize=address: any access outside
// them aborts the run).  Mirrors SolverT::init in dg_solver.hip: Morton
// order, geometry and neighbour words from elem_geometry, padded elements
// replicating element K - 1, per boundary face point 4 FP64 values.
//
//   lane_bounds <mesh.pmsh> <N: 1|4> <p: 1..5> <precision: 0 FP64 | 1 FP32>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "dg_geometry.h"
#include "dg_mesh.h"
#include "dg_valu_body.h"

using namespace dgmi;

template <typename R, int P>
static void fill_tab(TabP<R, P>& d) {
    Tables t = build_tables(P);
    constexpr int NP = TabP<R, P>::NP, NQ = TabP<R, P>::NQ, NQ1 = TabP<R, P>::NQ1;
    for (int m = 0; m < NP; m++) {
        for (int q = 0; q < NQ; q++) {
            d.B[m][q] = (R)t.basis[m * NQ + q];
            d.Gx[m][q] = (R)t.gx[m * NQ + q];
            d.Gy[m][q] = (R)t.gy[m * NQ + q];
        }
        for (int f = 0; f < 3; f++)
            for (int j = 0; j < NQ1; j++) d.S[f][m][j] = (R)t.side[(f * NP + m) * NQ1 + j];
        for (int v = 0; v < 3; v++) d.V[m][v] = (R)t.vertex[m * 3 + v];
    }
    for (int j = 0; j < NQ1; j++) d.W1[j] = (R)t.w1[j];
    for (int q = 0; q < NQ; q++) {
        d.R1[q] = (R)t.r1[q];
        d.R2[q] = (R)t.r2[q];
        d.W[q] = (R)t.w[q];
    }
}

template <int N, int P, typename R>
static int run(const Mesh& m) {
    constexpr int NP = n_modes(P), NQ1 = P + 1;
    const int K = m.K, Kp = (K + 63) / 64 * 64;
    std::vector<int> perm = morton_order(m);
    ElemGeometry g;
    std::string why = elem_geometry(m, perm, 0, g);
    if (!why.empty()) {
        printf("geometry: %s\n", why.c_str());
        return 1;
    }
    // exact-size buffers (mode-major state layout of the lane kernel)
    std::vector<R> u((size_t)N * NP * Kp), c, acc(u.size()), out(u.size());
    std::vector<R> ge(5 * (size_t)Kp), gf(9 * (size_t)Kp);
    std::vector<int> nbr(3 * (size_t)Kp);
    for (int i = 0; i < Kp; i++) {
        const int e = std::min(i, K - 1);
        ge[i] = (R)g.xr[e];
        ge[Kp + i] = (R)g.yr[e];
        ge[2 * Kp + i] = (R)g.xs[e];
        ge[3 * Kp + i] = (R)g.ys[e];
        ge[4 * Kp + i] = (R)(1.0 / g.J[e]);  // the inverse mass, as the solver stores it
        for (int f = 0; f < 3; f++) {
            gf[(3 * f) * Kp + i] = (R)g.fnx[f][e];
            gf[(3 * f + 1) * Kp + i] = (R)g.fny[f][e];
            gf[(3 * f + 2) * Kp + i] = (R)g.flen[f][e];
            nbr[f * Kp + i] = g.nbr[f][e];
        }
        // a positive state: density 1.4, energy 3.5, small modes
        for (int v = 0; v < N; v++)
            for (int k = 0; k < NP; k++)
                u[(v * NP + k) * Kp + i] = k == 0 ? R(v == 0 ? 1.4 : v == 3 ? 3.5 : 0.2) : R(1e-3 * k);
    }
    c = u;
    // boundary data: the reflecting faces' wall normal, the others a state
    std::vector<double> bnd((size_t)g.bfaces.size() * NQ1 * 4, 0.0);
    for (size_t b = 0; b < g.bfaces.size(); b++)
        for (int j = 0; j < NQ1; j++) {
            double* o = &bnd[(b * NQ1 + j) * 4];
            if (g.bfaces[b].kind == BC_REFLECTING) {
                o[0] = 0.6;
                o[1] = 0.8;
            } else {
                o[0] = 1.2;
                o[1] = 0.3;
                o[2] = -0.1;
                o[3] = 2.9;
            }
        }
    Ctrl ctrl{};
    ctrl.dt = 1e-3;
    ctrl.active = 1;
    StageArgs<R> a{};
    a.u = u.data();
    a.c = c.data();
    a.acc = acc.data();
    a.out = out.data();
    a.geo_e = ge.data();
    a.geo_f = gf.data();
    a.nbr = nbr.data();
    a.ctrl = &ctrl;
    a.lam_bits = &ctrl.lam_bits;
    a.K = K;
    a.Kp = Kp;
    a.phys.gamma = 1.4;
    a.phys.ax = 1.0;
    a.phys.ay = 1.0;
    a.bnd = g.bfaces.empty() ? nullptr : bnd.data();
    // every thread of the launch: 128-element tiles, their count rounded up to 8
    const int threads = ((Kp + 127) / 128 + 7) / 8 * 8 * 128;
    double lam = 0;
    for (int kind : {ST_RHS, ST_RK4_1, ST_RK4_2, ST_RK4_3, ST_RK4_4, ST_RK2_1, ST_RK2_2}) {
        a.kind = kind;
        for (int e = 0; e < threads; e++)
            lam = std::max(lam, a.bnd ? valu_process<N, P, R, true>(a, e, kind) : valu_process<N, P, R, false>(a, e, kind));
    }
    double sum = 0;
    for (R x : out) sum += (double)x;
    printf("N=%d p=%d %s K=%d boundary faces %zu: every access in bounds (checksum %.6e, lambda %.4f)\n", N, P,
           sizeof(R) == 8 ? "FP64" : "FP32", K, g.bfaces.size(), sum, lam);
    return std::isfinite(sum) ? 0 : 2;
}

template <int N, typename R>
static int by_p(const Mesh& m, int p) {
    switch (p) {
        case 1: return run<N, 1, R>(m);
        case 2: return run<N, 2, R>(m);
        case 3: return run<N, 3, R>(m);
        case 4: return run<N, 4, R>(m);
        case 5: return run<N, 5, R>(m);
    }
    return 3;
}

int main(int argc, char** argv) {
    if (argc != 5) return 4;
    fill_tab(c_tab64.p1);
    fill_tab(c_tab64.p2);
    fill_tab(c_tab64.p3);
    fill_tab(c_tab64.p4);
    fill_tab(c_tab64.p5);
    fill_tab(c_tab32.p1);
    fill_tab(c_tab32.p2);
    fill_tab(c_tab32.p3);
    fill_tab(c_tab32.p4);
    fill_tab(c_tab32.p5);
    Mesh m;
    std::string why = read_pmsh(argv[1], m);
    if (!why.empty()) {
        printf("mesh: %s\n", why.c_str());
        return 1;
    }
    const int N = atoi(argv[2]), p = atoi(argv[3]), prec = atoi(argv[4]);
    if (N == 4) return prec ? by_p<4, float>(m, p) : by_p<4, double>(m, p);
    return prec ? by_p<1, float>(m, p) : by_p<1, double>(m, p);
}
