#include "dg_mesh.h"

#include <algorithm>
#include <array>
#include <map>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>

namespace dgmi {

uint64_t splitmix64(uint64_t& state) {
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double uniform01(uint64_t& st) { return (double)(splitmix64(st) >> 11) * (1.0 / 9007199254740992.0); }

// .pmsh, conservation dialect (conserv.cu:134-191): num_elem, then one line per
// element "V1x V1y V2x V2y V3x V3y s1 s2 s3", num_sides, then one line per side
// "x1 y1 x2 y2 left_elem right_elem left_side_number right_side_number".
std::string read_pmsh(const char* path, Mesh& m) {
    FILE* f = fopen(path, "r");
    if (!f) return std::string("cannot open mesh file ") + path;
    char line[1024];
    auto fail = [&](const std::string& why) {
        fclose(f);
        return why;
    };
    if (!fgets(line, sizeof line, f) || sscanf(line, "%d", &m.K) != 1 || m.K <= 0)
        return fail("bad element count line");
    const int K = m.K;
    m.V1x.resize(K); m.V1y.resize(K); m.V2x.resize(K); m.V2y.resize(K); m.V3x.resize(K); m.V3y.resize(K);
    m.elem_s1.resize(K); m.elem_s2.resize(K); m.elem_s3.resize(K);
    for (int i = 0; i < K; i++) {
        if (!fgets(line, sizeof line, f)) return fail("unexpected end of file in element list");
        int items = sscanf(line, "%lf %lf %lf %lf %lf %lf %d %d %d", &m.V1x[i], &m.V1y[i], &m.V2x[i], &m.V2y[i],
                           &m.V3x[i], &m.V3y[i], &m.elem_s1[i], &m.elem_s2[i], &m.elem_s3[i]);
        if (items != 9) return fail("element line " + std::to_string(i) + ": expected 9 items, got " + std::to_string(items));
    }
    if (!fgets(line, sizeof line, f) || sscanf(line, "%d", &m.S) != 1 || m.S <= 0) return fail("bad side count line");
    const int S = m.S;
    m.sx1.resize(S); m.sy1.resize(S); m.sx2.resize(S); m.sy2.resize(S);
    m.left_elem.resize(S); m.right_elem.resize(S); m.left_side.resize(S); m.right_side.resize(S);
    for (int i = 0; i < S; i++) {
        if (!fgets(line, sizeof line, f)) return fail("unexpected end of file in side list");
        int items = sscanf(line, "%lf %lf %lf %lf %d %d %d %d", &m.sx1[i], &m.sy1[i], &m.sx2[i], &m.sy2[i],
                           &m.left_elem[i], &m.right_elem[i], &m.left_side[i], &m.right_side[i]);
        if (items != 8) return fail("side line " + std::to_string(i) + ": expected 8 items, got " + std::to_string(items));
    }
    fclose(f);
    m.period_x = m.period_y = 0;
    return validate(m);
}

// .pmsh, advection-prototype dialect (src/2d/2dadvec/2dadvec.cu:114-252):
// num_elem, then "V1x V1y V2x V2y V3x V3y" per element.  Sides are built from
// the elements in file order: a clockwise element first has V1 and V2
// swapped; each face (V1V2, V2V3, V3V1) is looked up among the sides made so
// far by its two end points' exact coordinates (either orientation); found,
// this element becomes the side's right element; not found, a new side is
// made with this element on its left and the face's end points in this
// element's orientation.  Sides met once get right element `open_kind`.
std::string read_vertex_pmsh(const char* path, Mesh& m, int open_kind) {
    if (open_kind > -1 || open_kind < -3) return "open_kind must be -1, -2 or -3";
    FILE* f = fopen(path, "r");
    if (!f) return std::string("cannot open mesh file ") + path;
    char line[1024];
    int K = 0;
    if (!fgets(line, sizeof line, f) || sscanf(line, "%d", &K) != 1 || K <= 0) {
        fclose(f);
        return "bad element count line";
    }
    m = Mesh();
    m.K = K;
    m.V1x.resize(K); m.V1y.resize(K); m.V2x.resize(K); m.V2y.resize(K); m.V3x.resize(K); m.V3y.resize(K);
    m.elem_s1.resize(K); m.elem_s2.resize(K); m.elem_s3.resize(K);
    for (int i = 0; i < K; i++) {
        if (!fgets(line, sizeof line, f) ||
            sscanf(line, "%lf %lf %lf %lf %lf %lf", &m.V1x[i], &m.V1y[i], &m.V2x[i], &m.V2y[i], &m.V3x[i],
                   &m.V3y[i]) != 6) {
            fclose(f);
            return "element line " + std::to_string(i) + ": expected 6 coordinates";
        }
    }
    fclose(f);
    // a side's key: its end points, the smaller (x, y) first
    using Pt = std::pair<double, double>;
    std::map<std::pair<Pt, Pt>, int> found;
    auto key = [](Pt a, Pt b) { return a < b ? std::make_pair(a, b) : std::make_pair(b, a); };
    for (int i = 0; i < K; i++) {
        const double J = (m.V2x[i] - m.V1x[i]) * (m.V3y[i] - m.V1y[i]) - (m.V3x[i] - m.V1x[i]) * (m.V2y[i] - m.V1y[i]);
        if (J < 0) {
            std::swap(m.V1x[i], m.V2x[i]);
            std::swap(m.V1y[i], m.V2y[i]);
        }
        const Pt P[3] = {{m.V1x[i], m.V1y[i]}, {m.V2x[i], m.V2y[i]}, {m.V3x[i], m.V3y[i]}};
        int* es[3] = {&m.elem_s1[i], &m.elem_s2[i], &m.elem_s3[i]};
        for (int fc = 0; fc < 3; fc++) {
            const Pt a = P[fc], b = P[(fc + 1) % 3];
            auto it = found.find(key(a, b));
            if (it != found.end() && m.right_elem[it->second] == open_kind) {
                const int s = it->second;
                m.right_elem[s] = i;
                m.right_side[s] = fc;
                *es[fc] = s;
                continue;
            }
            if (it != found.end()) return "side met by more than two elements (element " + std::to_string(i) + ")";
            const int s = m.S++;
            found[key(a, b)] = s;
            m.sx1.push_back(a.first);
            m.sy1.push_back(a.second);
            m.sx2.push_back(b.first);
            m.sy2.push_back(b.second);
            m.left_elem.push_back(i);
            m.right_elem.push_back(open_kind);
            m.left_side.push_back(fc);
            m.right_side.push_back(0);
            *es[fc] = s;
        }
    }
    m.period_x = m.period_y = 0;
    return validate(m);
}

// Periodic pairing of the open, untagged sides on the edges of the nodes'
// bounding box: a side on x = xmin pairs with the side on x = xmax whose end
// points are its own moved by the box width, and likewise in y.  The side met
// first keeps its index and gains the partner's element as its right element
// (the partner's face, traversed the other way as an interior side's right
// face is); the partner side is removed and the rest renumbered in order.
// genmesh.py has no periodic support; this is the pairing an interior side
// would get if the box were wrapped.
static std::string pair_periodic(const std::vector<double>& vx, const std::vector<double>& vy,
                                 std::vector<std::pair<int, int>>& sidelist, std::vector<int>& left_elem,
                                 std::vector<int>& right_elem, std::vector<int>& lsn, std::vector<int>& rsn,
                                 std::vector<char>& has_right, std::vector<std::array<int, 3>>& es, int& ns,
                                 double& period_x, double& period_y, double& x0, double& y0) {
    double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
    for (int s = 0; s < ns; s++)
        for (int v : {sidelist[s].first, sidelist[s].second}) {
            xmin = std::min(xmin, vx[v]); xmax = std::max(xmax, vx[v]);
            ymin = std::min(ymin, vy[v]); ymax = std::max(ymax, vy[v]);
        }
    const double tol = 1e-9 * std::max(xmax - xmin, ymax - ymin);
    // open untagged sides by box edge: 0 x = xmin, 1 x = xmax, 2 y = ymin, 3 y = ymax
    std::vector<int> on[4];
    for (int s = 0; s < ns; s++) {
        if (has_right[s] || right_elem[s] < 0) continue;
        const int a = sidelist[s].first, b = sidelist[s].second;
        auto at = [&](const std::vector<double>& c, double v) { return std::fabs(c[a] - v) <= tol && std::fabs(c[b] - v) <= tol; };
        const int edge = at(vx, xmin) ? 0 : at(vx, xmax) ? 1 : at(vy, ymin) ? 2 : at(vy, ymax) ? 3 : -1;
        if (edge < 0) return "side " + std::to_string(s) + " has one element, no boundary tag and is not on the box";
        on[edge].push_back(s);
    }
    std::vector<char> gone(ns, 0);
    for (int axis = 0; axis < 2; axis++) {
        auto& lo = on[2 * axis];
        auto& hi = on[2 * axis + 1];
        if (lo.size() != hi.size())
            return std::string("periodic pairing: ") + std::to_string(lo.size()) + " open sides on " +
                   (axis ? "y = ymin" : "x = xmin") + " against " + std::to_string(hi.size()) + " on the opposite edge";
        if (lo.empty()) continue;
        // the coordinate along the edge: y for the x edges, x for the y edges
        const std::vector<double>& t = axis ? vx : vy;
        auto lo_end = [&](int s) { return std::min(t[sidelist[s].first], t[sidelist[s].second]); };
        auto hi_end = [&](int s) { return std::max(t[sidelist[s].first], t[sidelist[s].second]); };
        auto by_lo = [&](int a, int b) { return lo_end(a) < lo_end(b); };
        std::sort(lo.begin(), lo.end(), by_lo);
        std::sort(hi.begin(), hi.end(), by_lo);
        for (size_t i = 0; i < lo.size(); i++) {
            const int a = lo[i], b = hi[i];
            if (std::fabs(lo_end(a) - lo_end(b)) > tol || std::fabs(hi_end(a) - hi_end(b)) > tol)
                return "periodic pairing: side " + std::to_string(a) + " has no partner on the opposite edge";
            const int keep = std::min(a, b), drop = std::max(a, b);
            right_elem[keep] = left_elem[drop];
            rsn[keep] = lsn[drop];
            has_right[keep] = 1;
            es[left_elem[drop]][lsn[drop]] = keep;
            gone[drop] = 1;
        }
        (axis ? period_y : period_x) = axis ? ymax - ymin : xmax - xmin;
    }
    x0 = xmin;
    y0 = ymin;
    std::vector<int> renum(ns, -1);
    int n = 0;
    for (int s = 0; s < ns; s++) {
        if (gone[s]) continue;
        renum[s] = n;
        sidelist[n] = sidelist[s];
        left_elem[n] = left_elem[s];
        right_elem[n] = right_elem[s];
        lsn[n] = lsn[s];
        rsn[n] = rsn[s];
        has_right[n] = has_right[s];
        n++;
    }
    for (auto& e : es)
        for (int& s : e) s = renum[s];
    ns = n;
    return std::string();
}

std::string read_gmsh(const char* path, Mesh& m, bool quirks, bool periodic) {
    FILE* f = fopen(path, "r");
    if (!f) return std::string("cannot open mesh file ") + path;
    char line[4096];
    auto fail = [&](const std::string& why) {
        fclose(f);
        return why;
    };
    auto seek = [&](const char* tag) {
        while (fgets(line, sizeof line, f))
            if (strstr(line, tag)) return true;
        return false;
    };
    if (!seek("$Nodes")) return fail("no $Nodes section");
    int nv = 0;
    if (!fgets(line, sizeof line, f) || sscanf(line, "%d", &nv) != 1 || nv <= 0) return fail("bad node count");
    std::vector<double> vx(nv), vy(nv);
    for (int i = 0; i < nv; i++) {
        int id;
        if (!fgets(line, sizeof line, f) || sscanf(line, "%d %lf %lf", &id, &vx[i], &vy[i]) != 3)
            return fail("bad node line " + std::to_string(i));
    }
    if (!seek("$Elements")) return fail("no $Elements section");
    int ne = 0;
    if (!fgets(line, sizeof line, f) || sscanf(line, "%d", &ne) != 1 || ne < 0) return fail("bad element count");
    std::vector<std::array<int, 3>> tri;
    std::vector<std::pair<int, int>> bside;
    std::vector<int> btag;
    for (int i = 0; i < ne; i++) {
        if (!fgets(line, sizeof line, f)) return fail("unexpected end of file in $Elements");
        long t[16];
        int n = 0;
        for (char* tok = strtok(line, " \t\r\n"); tok && n < 16; tok = strtok(nullptr, " \t\r\n")) t[n++] = atol(tok);
        // genmesh.py:44-58: 7 fields = a tagged line, 8 fields = a triangle
        if (n == 7) {
            bside.push_back({(int)t[5] - 1, (int)t[6] - 1});
            btag.push_back((int)t[3]);
        } else if (n == 8) {
            tri.push_back({(int)t[5] - 1, (int)t[6] - 1, (int)t[7] - 1});
        }
    }
    fclose(f);
    const int K = (int)tri.size();
    if (K == 0) return "mesh has no triangles";
    for (auto& e : tri)
        for (int v : e)
            if (v < 0 || v >= nv) return "element refers to a missing node";
    // boundary code of a vertex pair, either orientation (the last tagged line wins)
    std::map<std::pair<int, int>, int> bcode;
    for (size_t j = 0; j < bside.size(); j++) {
        const int code = btag[j] == 10000 ? -1 : btag[j] == 20000 ? -2 : btag[j] == 30000 ? -3 : 0;
        if (code == 0) continue;
        bcode[bside[j]] = code;
        bcode[{bside[j].second, bside[j].first}] = code;
    }
    std::vector<int> left_elem(3 * (size_t)K, 0), right_elem(3 * (size_t)K, 0), lsn(3 * (size_t)K, 0),
        rsn(3 * (size_t)K, 0);
    std::vector<std::array<int, 3>> es(K, {0, 0, 0});
    std::vector<std::pair<int, int>> sidelist(3 * (size_t)K);
    std::vector<char> has_right(3 * (size_t)K, 0);
    std::map<std::pair<int, int>, int> sidemap;
    int ns = 0;
    for (int i = 0; i < K; i++) {
        auto& e = tri[i];
        const double J = (vx[e[1]] - vx[e[0]]) * (vy[e[2]] - vy[e[0]]) - (vx[e[2]] - vx[e[0]]) * (vy[e[1]] - vy[e[0]]);
        if (J < 0) {
            if (quirks) {  // genmesh.py:100-103 swaps the shared coordinates
                std::swap(vx[e[0]], vx[e[1]]);
                std::swap(vy[e[0]], vy[e[1]]);
            } else {
                std::swap(e[0], e[1]);
            }
        }
        const std::pair<int, int> fs[3] = {{e[0], e[1]}, {e[1], e[2]}, {e[2], e[0]}};
        bool is_new[3] = {true, true, true};
        for (int k = 0; k < 3; k++)
            for (const auto& key : {fs[k], std::make_pair(fs[k].second, fs[k].first)}) {
                auto it = sidemap.find(key);
                if (it == sidemap.end()) continue;
                const int j = it->second;
                is_new[k] = false;
                right_elem[j] = i;
                has_right[j] = 1;
                es[i][k] = j;
                rsn[j] = k;
            }
        for (int k = 0; k < 3; k++) {
            if (!is_new[k]) continue;
            sidemap[fs[k]] = ns;
            sidelist[ns] = fs[k];
            lsn[ns] = k;
            auto b = bcode.find(fs[k]);
            if (b != bcode.end()) right_elem[ns] = b->second;
            es[i][k] = ns;
            left_elem[ns] = i;
            ns++;
        }
    }
    double period_x = 0, period_y = 0, x0 = 0, y0 = 0;
    if (periodic) {
        if (quirks) return "periodic pairing is not part of genmesh.py's conversion";
        std::string why = pair_periodic(vx, vy, sidelist, left_elem, right_elem, lsn, rsn, has_right, es, ns, period_x,
                                        period_y, x0, y0);
        if (!why.empty()) return why;
    }
    if (!quirks)
        for (int s = 0; s < ns; s++)
            if (!has_right[s] && right_elem[s] >= 0)
                return "side " + std::to_string(s) + " has one element and no boundary tag";
    // boundary sides to the front, reflecting then outflow then inflow
    // (genmesh.py:236-283: the swap of sides i and j and its fix-ups)
    auto slot = [&](int elem) -> std::array<int, 3>* {
        if (elem >= 0) return &es[elem];
        if (quirks && elem >= -3 && K + elem >= 0) return &es[K + elem];  // negative index
        return nullptr;
    };
    int j = 0;
    for (int code : {-1, -2, -3}) {
        for (int i = 0; i < ns; i++) {
            if (right_elem[i] != code) continue;
            if (auto* a = slot(left_elem[j])) (*a)[lsn[j]] = i;
            if (auto* a = slot(right_elem[j])) (*a)[rsn[j]] = i;
            if (auto* a = slot(left_elem[i])) (*a)[lsn[i]] = j;
            std::swap(sidelist[i], sidelist[j]);
            std::swap(left_elem[i], left_elem[j]);
            std::swap(right_elem[i], right_elem[j]);
            std::swap(lsn[i], lsn[j]);
            std::swap(rsn[i], rsn[j]);
            j++;
        }
    }
    m = Mesh{};
    m.K = K;
    m.S = ns;
    m.period_x = period_x;
    m.period_y = period_y;
    m.x0 = x0;
    m.y0 = y0;
    for (int i = 0; i < K; i++) {
        const auto& e = tri[i];
        m.V1x.push_back(vx[e[0]]); m.V1y.push_back(vy[e[0]]);
        m.V2x.push_back(vx[e[1]]); m.V2y.push_back(vy[e[1]]);
        m.V3x.push_back(vx[e[2]]); m.V3y.push_back(vy[e[2]]);
        m.elem_s1.push_back(es[i][0]); m.elem_s2.push_back(es[i][1]); m.elem_s3.push_back(es[i][2]);
    }
    for (int s = 0; s < ns; s++) {
        m.sx1.push_back(vx[sidelist[s].first]); m.sy1.push_back(vy[sidelist[s].first]);
        m.sx2.push_back(vx[sidelist[s].second]); m.sy2.push_back(vy[sidelist[s].second]);
        m.left_elem.push_back(left_elem[s]); m.right_elem.push_back(right_elem[s]);
        m.left_side.push_back(lsn[s]); m.right_side.push_back(rsn[s]);
    }
    return quirks ? std::string() : validate(m);
}

std::string write_pmsh(const char* path, const Mesh& m) {
    FILE* f = fopen(path, "w");
    if (!f) return std::string("cannot open ") + path + " for writing";
    fprintf(f, "%d\n", m.K);
    for (int i = 0; i < m.K; i++)
        fprintf(f, "%.17g %.17g %.17g %.17g %.17g %.17g %d %d %d\n", m.V1x[i], m.V1y[i], m.V2x[i], m.V2y[i], m.V3x[i],
                m.V3y[i], m.elem_s1[i], m.elem_s2[i], m.elem_s3[i]);
    fprintf(f, "%d\n", m.S);
    for (int i = 0; i < m.S; i++)
        fprintf(f, "%.17g %.17g %.17g %.17g %d %d %d %d\n", m.sx1[i], m.sy1[i], m.sx2[i], m.sy2[i], m.left_elem[i],
                m.right_elem[i], m.left_side[i], m.right_side[i]);
    int bad = ferror(f);
    fclose(f);
    return bad ? std::string("write error on ") + path : std::string();
}

static uint32_t spread16(uint32_t v) {
    v &= 0xFFFF;
    v = (v | (v << 8)) & 0x00FF00FF;
    v = (v | (v << 4)) & 0x0F0F0F0F;
    v = (v | (v << 2)) & 0x33333333;
    v = (v | (v << 1)) & 0x55555555;
    return v;
}

static uint32_t morton2(uint32_t i, uint32_t j) { return spread16(i) | (spread16(j) << 1); }

std::string make_periodic_jittered(int nx, int ny, double x0, double x1, double y0, double y1, double jitter,
                                   uint64_t seed, Mesh& m) {
    if (nx < 1 || ny < 1 || nx > 65535 || ny > 65535) return "grid size out of range";
    if (!(x1 > x0) || !(y1 > y0)) return "empty domain";
    if (!(jitter >= 0) || jitter >= 0.25) return "jitter must lie in [0, 0.25)";
    const double hx = (x1 - x0) / nx, hy = (y1 - y0) / ny;
    // lattice (nx+1) x (ny+1); index nx / ny is the periodic copy of 0
    std::vector<double> vx((size_t)(nx + 1) * (ny + 1)), vy(vx.size());
    uint64_t st = seed;
    auto vid = [&](int i, int j) { return (size_t)j * (nx + 1) + i; };
    for (int j = 0; j <= ny; j++)
        for (int i = 0; i <= nx; i++) {
            vx[vid(i, j)] = (i == nx) ? x1 : x0 + i * hx;
            vy[vid(i, j)] = (j == ny) ? y1 : y0 + j * hy;
        }
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            double ux = uniform01(st), uy = uniform01(st);
            if (i > 0 && j > 0) {
                vx[vid(i, j)] += (2.0 * ux - 1.0) * jitter * hx;
                vy[vid(i, j)] += (2.0 * uy - 1.0) * jitter * hy;
            }
        }
    // cells in blocks of about 8 x 4 (64 triangles, one tile of the
    // matrix-core kernel, so tiles stay compact whatever nx, ny are): the
    // block columns / rows partition the grid evenly (sizes differ by at most
    // one cell), blocks in Morton order, cells row-major inside a block
    const int C = nx * ny;
    const int nbx = (nx + 7) / 8, nby = (ny + 3) / 4;
    std::vector<int> cells(C);
    std::iota(cells.begin(), cells.end(), 0);
    auto key = [&](int c) {
        const int i = c % nx, j = c / nx;
        const int bi = (int)((int64_t)i * nbx / nx), bj = (int)((int64_t)j * nby / ny);
        return ((uint64_t)morton2(bi, bj) << 32) | ((uint64_t)j << 16) | (uint64_t)i;
    };
    std::stable_sort(cells.begin(), cells.end(), [&](int a, int b) { return key(a) < key(b); });
    std::vector<int> cell_rank(C);
    for (int r = 0; r < C; r++) cell_rank[cells[r]] = r;
    const int K = 2 * C;
    m = Mesh();
    m.K = K;
    m.V1x.resize(K); m.V1y.resize(K); m.V2x.resize(K); m.V2y.resize(K); m.V3x.resize(K); m.V3y.resize(K);
    for (int r = 0; r < C; r++) {
        int i = cells[r] % nx, j = cells[r] / nx;
        size_t a = vid(i, j), b = vid(i + 1, j), c = vid(i + 1, j + 1), d = vid(i, j + 1);
        int L = 2 * r, U = 2 * r + 1;
        m.V1x[L] = vx[a]; m.V1y[L] = vy[a]; m.V2x[L] = vx[b]; m.V2y[L] = vy[b]; m.V3x[L] = vx[c]; m.V3y[L] = vy[c];
        m.V1x[U] = vx[a]; m.V1y[U] = vy[a]; m.V2x[U] = vx[c]; m.V2y[U] = vy[c]; m.V3x[U] = vx[d]; m.V3y[U] = vy[d];
    }
    // neighbour (element, face) of every (element, face)
    auto elem_of = [&](int i, int j, int upper) {
        i = ((i % nx) + nx) % nx;
        j = ((j % ny) + ny) % ny;
        return 2 * cell_rank[j * nx + i] + upper;
    };
    std::vector<int> nb_e(3 * (size_t)K), nb_f(3 * (size_t)K);
    for (int r = 0; r < C; r++) {
        int i = cells[r] % nx, j = cells[r] / nx;
        int L = 2 * r, U = 2 * r + 1;
        // lower: f0 = bottom edge (shared with the upper triangle of the cell below, its f1)
        //        f1 = right edge (upper triangle of the cell to the right, its f2)
        //        f2 = diagonal (upper triangle of the same cell, its f0)
        nb_e[3 * L + 0] = elem_of(i, j - 1, 1); nb_f[3 * L + 0] = 1;
        nb_e[3 * L + 1] = elem_of(i + 1, j, 1); nb_f[3 * L + 1] = 2;
        nb_e[3 * L + 2] = U;                    nb_f[3 * L + 2] = 0;
        nb_e[3 * U + 0] = L;                    nb_f[3 * U + 0] = 2;
        nb_e[3 * U + 1] = elem_of(i, j + 1, 0); nb_f[3 * U + 1] = 0;
        nb_e[3 * U + 2] = elem_of(i - 1, j, 0); nb_f[3 * U + 2] = 1;
    }
    std::vector<int> side_of(3 * (size_t)K, -1);
    m.elem_s1.resize(K); m.elem_s2.resize(K); m.elem_s3.resize(K);
    for (int e = 0; e < K; e++)
        for (int f = 0; f < 3; f++) {
            if (side_of[3 * e + f] >= 0) continue;
            int s = m.S++;
            int ne = nb_e[3 * e + f], nf = nb_f[3 * e + f];
            side_of[3 * e + f] = s;
            side_of[3 * ne + nf] = s;
            double ax_, ay_, bx_, by_;
            if (f == 0) { ax_ = m.V1x[e]; ay_ = m.V1y[e]; bx_ = m.V2x[e]; by_ = m.V2y[e]; }
            else if (f == 1) { ax_ = m.V2x[e]; ay_ = m.V2y[e]; bx_ = m.V3x[e]; by_ = m.V3y[e]; }
            else { ax_ = m.V3x[e]; ay_ = m.V3y[e]; bx_ = m.V1x[e]; by_ = m.V1y[e]; }
            m.sx1.push_back(ax_); m.sy1.push_back(ay_); m.sx2.push_back(bx_); m.sy2.push_back(by_);
            m.left_elem.push_back(e); m.right_elem.push_back(ne);
            m.left_side.push_back(f); m.right_side.push_back(nf);
        }
    for (int e = 0; e < K; e++) {
        m.elem_s1[e] = side_of[3 * e + 0];
        m.elem_s2[e] = side_of[3 * e + 1];
        m.elem_s3[e] = side_of[3 * e + 2];
    }
    m.period_x = x1 - x0;
    m.period_y = y1 - y0;
    m.x0 = x0;
    m.y0 = y0;
    m.ordered = true;
    return validate(m);
}

std::string validate(const Mesh& m) {
    if (m.K <= 0 || m.S <= 0) return "empty mesh";
    const int* es[3] = {m.elem_s1.data(), m.elem_s2.data(), m.elem_s3.data()};
    for (int e = 0; e < m.K; e++)
        for (int f = 0; f < 3; f++) {
            int s = es[f][e];
            if (s < 0 || s >= m.S) return "element " + std::to_string(e) + " face " + std::to_string(f) + " has side id out of range";
            bool l = m.left_elem[s] == e && m.left_side[s] == f;
            bool r = m.right_elem[s] == e && m.right_side[s] == f;
            if (!l && !r) return "element " + std::to_string(e) + " face " + std::to_string(f) + " is not referenced by its side";
        }
    for (int s = 0; s < m.S; s++) {
        int l = m.left_elem[s], r = m.right_elem[s];
        if (l < 0 || l >= m.K) return "side " + std::to_string(s) + " has left element out of range";
        if (r >= m.K || r < -3) return "side " + std::to_string(s) + " has right element out of range";
        if (m.left_side[s] < 0 || m.left_side[s] > 2) return "side " + std::to_string(s) + " bad left side number";
        if (r >= 0 && (m.right_side[s] < 0 || m.right_side[s] > 2)) return "side " + std::to_string(s) + " bad right side number";
    }
    return std::string();
}

std::vector<int> morton_order(const Mesh& m) {
    double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
    std::vector<double> cx(m.K), cy(m.K);
    for (int e = 0; e < m.K; e++) {
        cx[e] = (m.V1x[e] + m.V2x[e] + m.V3x[e]) / 3.0;
        cy[e] = (m.V1y[e] + m.V2y[e] + m.V3y[e]) / 3.0;
        xmin = std::min(xmin, cx[e]); xmax = std::max(xmax, cx[e]);
        ymin = std::min(ymin, cy[e]); ymax = std::max(ymax, cy[e]);
    }
    double sx = (xmax > xmin) ? 65535.0 / (xmax - xmin) : 0, sy = (ymax > ymin) ? 65535.0 / (ymax - ymin) : 0;
    std::vector<uint32_t> key(m.K);
    for (int e = 0; e < m.K; e++)
        key[e] = morton2((uint32_t)((cx[e] - xmin) * sx), (uint32_t)((cy[e] - ymin) * sy));
    std::vector<int> perm(m.K);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return key[a] < key[b]; });
    return perm;
}

}  // namespace dgmi
