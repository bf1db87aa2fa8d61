#include "dg_geometry.h"

#include <cmath>

namespace dgmi {

static double seg_len(double dx, double dy, bool float_geom) {
    if (float_geom) {
        float fx = (float)dx, fy = (float)dy;
        return (double)std::sqrt(fx * fx + fy * fy);
    }
    return std::sqrt(dx * dx + dy * dy);
}

double mesh_geometry(const Mesh& m, int quirks, std::vector<double>& geo_e, std::vector<double>& geo_s) {
    const bool fg = (quirks & 2) != 0;
    const int K = m.K, S = m.S;
    geo_e.assign(5 * (size_t)K, 0.0);
    geo_s.assign(3 * (size_t)S, 0.0);
    double min_r = 0;
    for (int e = 0; e < K; e++) {
        const double x1 = m.V1x[e], y1 = m.V1y[e], x2 = m.V2x[e], y2 = m.V2y[e], x3 = m.V3x[e], y3 = m.V3y[e];
        // inscribed-circle diameter 2*Area/s (Heron), conserv_kernels.cu:355-373
        const double a = seg_len(x1 - x2, y1 - y2, fg);
        const double b = seg_len(x2 - x3, y2 - y3, fg);
        const double c = seg_len(x1 - x3, y1 - y3, fg);
        const double k = 0.5 * (a + b + c);
        const double prod = k * (k - a) * (k - b) * (k - c);
        const double d = fg ? 2 * (double)std::sqrt((float)prod) / k : 2 * std::sqrt(prod) / k;
        if (e == 0 || d < min_r) min_r = d;
        geo_e[e] = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1);
        geo_e[K + e] = x2 - x1;
        geo_e[2 * (size_t)K + e] = y2 - y1;
        geo_e[3 * (size_t)K + e] = x3 - x1;
        geo_e[4 * (size_t)K + e] = y3 - y1;
    }
    for (int s = 0; s < S; s++) {
        geo_s[s] = seg_len(m.sx1[s] - m.sx2[s], m.sy1[s] - m.sy2[s], fg);
        const double x = m.sx2[s] - m.sx1[s], y = m.sy2[s] - m.sy1[s];
        const double length = seg_len(x, y, fg);
        double nx = -y / length, ny = x / length;
        // orient away from the left element's vertex opposite this side
        const int l = m.left_elem[s];
        double tx, ty, ix, iy;
        switch (m.left_side[s]) {
            case 0: tx = m.V3x[l]; ty = m.V3y[l]; ix = (m.V1x[l] + m.V2x[l]) / 2.; iy = (m.V1y[l] + m.V2y[l]) / 2.; break;
            case 1: tx = m.V1x[l]; ty = m.V1y[l]; ix = (m.V2x[l] + m.V3x[l]) / 2.; iy = (m.V2y[l] + m.V3y[l]) / 2.; break;
            default: tx = m.V2x[l]; ty = m.V2y[l]; ix = (m.V1x[l] + m.V3x[l]) / 2.; iy = (m.V1y[l] + m.V3y[l]) / 2.; break;
        }
        const double dot = nx * (tx - ix) + ny * (ty - iy);
        if (dot > 0) {
            nx *= -1;
            ny *= -1;
        }
        geo_s[S + s] = nx;
        geo_s[2 * (size_t)S + s] = ny;
    }
    return min_r;
}

std::string elem_geometry(const Mesh& m, const std::vector<int>& perm, int quirks, ElemGeometry& g) {
    const int K = m.K;
    for (int s = 0; s < m.S; s++)
        if (m.right_elem[s] < -3)
            return "side " + std::to_string(s) + " has boundary code " + std::to_string(m.right_elem[s]) +
                   " (reflecting -1, outflow -2 and inflow -3 are known)";
    if (K >= (1 << 28)) return "mesh too large";
    g.bfaces.clear();
    std::vector<double> ge, gs;
    g.min_r = mesh_geometry(m, quirks, ge, gs);
    g.K = K;
    std::vector<int> inv(K);
    for (int n = 0; n < K; n++) inv[perm[n]] = n;
    g.xr.resize(K); g.yr.resize(K); g.xs.resize(K); g.ys.resize(K); g.J.resize(K);
    for (int f = 0; f < 3; f++) {
        g.fnx[f].resize(K); g.fny[f].resize(K); g.flen[f].resize(K); g.nbr[f].resize(K);
    }
    for (int v = 0; v < 6; v++) g.V[v].resize(K);
    const int* es[3] = {m.elem_s1.data(), m.elem_s2.data(), m.elem_s3.data()};
    for (int n = 0; n < K; n++) {
        const int o = perm[n];
        g.J[n] = ge[o];
        g.xr[n] = ge[K + o];
        g.yr[n] = ge[2 * (size_t)K + o];
        g.xs[n] = ge[3 * (size_t)K + o];
        g.ys[n] = ge[4 * (size_t)K + o];
        g.V[0][n] = m.V1x[o]; g.V[1][n] = m.V1y[o]; g.V[2][n] = m.V2x[o];
        g.V[3][n] = m.V2y[o]; g.V[4][n] = m.V3x[o]; g.V[5][n] = m.V3y[o];
        for (int f = 0; f < 3; f++) {
            const int s = es[f][o];
            const bool left = (m.left_elem[s] == o && m.left_side[s] == f);
            const int other = left ? m.right_elem[s] : m.left_elem[s];
            const int ofc = left ? m.right_side[s] : m.left_side[s];
            g.fnx[f][n] = gs[m.S + s];
            g.fny[f][n] = gs[2 * (size_t)m.S + s];
            g.flen[f][n] = gs[s];
            if (other < 0) {  // boundary side: this element is its left element
                // the kernels' boundary word: side index, kind in bits 28..29,
                // left flag (bit 30), bit 31 set
                g.nbr[f][n] = (int)((unsigned)g.bfaces.size() | ((unsigned)-other << 28) | (1u << 30) | (1u << 31));
                g.bfaces.push_back(ElemGeometry::BFace{n, f, -other});
            } else {
                g.nbr[f][n] = (inv[other]) | (ofc << 28) | ((left ? 1 : 0) << 30);
            }
        }
    }
    return std::string();
}

}  // namespace dgmi
