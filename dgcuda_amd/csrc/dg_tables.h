// Quadrature rules and basis tables for one polynomial order p (0..5).
#pragma once
#include <vector>

namespace dgmi {

inline constexpr int n_modes(int p) { return (p + 1) * (p + 2) / 2; }
// volume rule used at order p: degree 2p (1 point at p = 0), as selected by
// the reference's set_quadrature (src/2d/conservation/conserv.cu:61-106)
inline constexpr int n_quad_vol(int p) { return p == 0 ? 1 : p == 1 ? 3 : p == 2 ? 6 : p == 3 ? 12 : p == 4 ? 16 : 25; }
inline constexpr int n_quad_face(int p) { return p + 1; }

struct Tables {
    int p = 0, n_p = 0, n_q = 0, n_q1 = 0;
    std::vector<double> r1, r2, w;   // [n_q] volume points and weights (weights sum to 1/2)
    std::vector<double> s_r, w1;     // [n_q1] Gauss-Legendre points/weights on [-1,1]
    std::vector<double> basis;       // [n_p][n_q]   phi_i(r_j, s_j)
    std::vector<double> gx, gy;      // [n_p][n_q]   w_j * dphi_i/dr, w_j * dphi_i/ds
    std::vector<double> side;        // [3][n_p][n_q1] phi_i on face f at point j (CCW)
    std::vector<double> vertex;      // [n_p][3]     phi_i at (0,0), (1,0), (0,1)
};

Tables build_tables(int p);

// phi_i and its reference-space gradient at (r, s), from the monomial
// coefficient table
double basis_value(int i, double r, double s);
void basis_grad(int i, double r, double s, double* dr, double* ds);

}  // namespace dgmi
