#include "dg_tables.h"

#include <cmath>

#include "dg_ref_tables.inc"

namespace dgmi {

// Monomials in graded order: index k <-> (a, b) with r^a s^b, degree d = a + b,
// k = d(d+1)/2 + b.
static void monomial_exponents(int k, int* a, int* b) {
    int d = 0;
    while ((d + 1) * (d + 2) / 2 <= k) d++;
    *b = k - d * (d + 1) / 2;
    *a = d - *b;
}

static double ipow(double x, int n) { return std::pow(x, (double)n); }

double basis_value(int i, double r, double s) {
    double sum = 0.0;
    for (int k = 0; k <= i; k++) {
        int a, b;
        monomial_exponents(k, &a, &b);
        sum += DG_BASIS_COEFF[i][k] * (ipow(r, a) * ipow(s, b));
    }
    return sum;
}

void basis_grad(int i, double r, double s, double* dr, double* ds) {
    double gr = 0.0, gs = 0.0;
    for (int k = 1; k <= i; k++) {
        int a, b;
        monomial_exponents(k, &a, &b);
        if (a > 0) gr += DG_BASIS_COEFF[i][k] * (a * ipow(r, a - 1) * ipow(s, b));
        if (b > 0) gs += DG_BASIS_COEFF[i][k] * (ipow(r, a) * b * ipow(s, b - 1));
    }
    *dr = gr;
    *ds = gs;
}

Tables build_tables(int p) {
    Tables t;
    t.p = p;
    t.n_p = n_modes(p);
    t.n_q = n_quad_vol(p);
    t.n_q1 = n_quad_face(p);
    // degree-2p triangle rule (degree 1 at p = 0); weights halved to the
    // reference triangle's area 1/2
    const double* q2 = DG_QUAD2D[p == 0 ? 0 : 2 * p - 1];
    for (int j = 0; j < t.n_q; j++) {
        t.r1.push_back(q2[3 * j]);
        t.r2.push_back(q2[3 * j + 1]);
        t.w.push_back(q2[3 * j + 2] / 2.0);
    }
    const double* q1 = DG_QUAD1D[p];
    for (int j = 0; j < t.n_q1; j++) {
        t.s_r.push_back(q1[2 * j]);
        t.w1.push_back(q1[2 * j + 1]);
    }
    const int np = t.n_p, nq = t.n_q, n1 = t.n_q1;
    t.basis.resize(np * nq);
    t.gx.resize(np * nq);
    t.gy.resize(np * nq);
    t.side.resize(3 * np * n1);
    t.vertex.resize(np * 3);
    for (int i = 0; i < np; i++) {
        for (int j = 0; j < nq; j++) {
            double dr, ds;
            t.basis[i * nq + j] = basis_value(i, t.r1[j], t.r2[j]);
            basis_grad(i, t.r1[j], t.r2[j], &dr, &ds);
            t.gx[i * nq + j] = t.w[j] * dr;
            t.gy[i * nq + j] = t.w[j] * ds;
        }
        // face points run counter-clockwise: face 0 from (0,0) to (1,0),
        // face 1 from (1,0) to (0,1), face 2 from (0,1) to (0,0)
        for (int j = 0; j < n1; j++) {
            double x = t.s_r[j], xr = t.s_r[n1 - 1 - j];
            t.side[(0 * np + i) * n1 + j] = basis_value(i, 0.5 + 0.5 * x, 0.0);
            t.side[(1 * np + i) * n1 + j] = basis_value(i, (1.0 - x) / 2.0, (1.0 + x) / 2.0);
            t.side[(2 * np + i) * n1 + j] = basis_value(i, 0.0, 0.5 + 0.5 * xr);
        }
        t.vertex[i * 3 + 0] = basis_value(i, 0.0, 0.0);
        t.vertex[i * 3 + 1] = basis_value(i, 1.0, 0.0);
        t.vertex[i * 3 + 2] = basis_value(i, 0.0, 1.0);
    }
    return t;
}

}  // namespace dgmi
