// extern "C" entry points of libdgmi.so (declared in include/dgmi.h).
// Every call converts internal errors into a dg_status and a thread-local
// message; nothing here exits or aborts.
#include <hip/hip_runtime.h>

#include <cstring>
#include <exception>
#include <new>
#include <string>

#include "../../include/dgmi.h"
#include "dg_batch.h"
#include "dg_geometry.h"
#include "dg_mesh.h"
#include "dg_server.h"
#include "dg_solver.h"
#include "dg_tables.h"

using namespace dgmi;

struct dg_mesh {
    Mesh m;
};
struct dg_solver {
    Solver* s = nullptr;
};
struct dg_batch {
    Batch* b = nullptr;
};
struct dg_server {
    Server* s = nullptr;
};

static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

static int ret(const Status& st) {
    if (st.code != DG_OK) g_err = st.msg;
    return st.code;
}

#define GUARD_BEGIN try {
#define GUARD_END                                                    \
    }                                                                \
    catch (const std::bad_alloc&) {                                  \
        return fail(DG_ERR_BUDGET, "host allocation failed");        \
    }                                                                \
    catch (const std::exception& ex) {                               \
        return fail(DG_ERR_STATE, std::string("exception: ") + ex.what()); \
    }

extern "C" {

const char* dg_last_error(void) { return g_err.c_str(); }

const char* dg_version(void) { return "dgmi 0.1 (gfx950)"; }

int dg_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int dg_mesh_read_pmsh(const char* path, dg_mesh** out) {
    GUARD_BEGIN
    if (!path || !out) return fail(DG_ERR_ARG, "null argument");
    auto* h = new dg_mesh();
    std::string why = read_pmsh(path, h->m);
    if (!why.empty()) {
        delete h;
        return fail(why.rfind("cannot open", 0) == 0 ? DG_ERR_IO : DG_ERR_PARSE, why);
    }
    *out = h;
    return DG_OK;
    GUARD_END
}

int dg_mesh_read_gmsh(const char* path, int flags, dg_mesh** out) {
    GUARD_BEGIN
    if (!path || !out) return fail(DG_ERR_ARG, "null argument");
    if (flags & ~(DG_GMSH_QUIRKS | DG_GMSH_PERIODIC)) return fail(DG_ERR_ARG, "unknown gmsh flags");
    if ((flags & DG_GMSH_QUIRKS) && (flags & DG_GMSH_PERIODIC))
        return fail(DG_ERR_ARG, "periodic pairing is not part of genmesh.py's conversion (no quirk mode)");
    auto* h = new dg_mesh();
    std::string why = read_gmsh(path, h->m, (flags & DG_GMSH_QUIRKS) != 0, (flags & DG_GMSH_PERIODIC) != 0);
    if (!why.empty()) {
        delete h;
        return fail(why.rfind("cannot open", 0) == 0 ? DG_ERR_IO : DG_ERR_PARSE, why);
    }
    *out = h;
    return DG_OK;
    GUARD_END
}

int dg_mesh_read_vertex_pmsh(const char* path, int open_kind, dg_mesh** out) {
    GUARD_BEGIN
    if (!path || !out) return fail(DG_ERR_ARG, "null argument");
    auto* h = new dg_mesh();
    std::string why = read_vertex_pmsh(path, h->m, open_kind);
    if (!why.empty()) {
        delete h;
        const bool io = why.rfind("cannot open", 0) == 0;
        const bool arg = why.rfind("open_kind", 0) == 0;
        return fail(io ? DG_ERR_IO : arg ? DG_ERR_ARG : DG_ERR_PARSE, why);
    }
    *out = h;
    return DG_OK;
    GUARD_END
}

int dg_mesh_from_arrays(int K, int S, const double* V, const int* elem_s, const double* sides, const int* side_int,
                        double period_x, double period_y, dg_mesh** out) {
    GUARD_BEGIN
    if (K <= 0 || S <= 0 || !V || !elem_s || !sides || !side_int || !out) return fail(DG_ERR_ARG, "invalid argument");
    auto* h = new dg_mesh();
    Mesh& m = h->m;
    m.K = K;
    m.S = S;
    m.V1x.assign(V, V + K); m.V1y.assign(V + K, V + 2 * (size_t)K);
    m.V2x.assign(V + 2 * (size_t)K, V + 3 * (size_t)K); m.V2y.assign(V + 3 * (size_t)K, V + 4 * (size_t)K);
    m.V3x.assign(V + 4 * (size_t)K, V + 5 * (size_t)K); m.V3y.assign(V + 5 * (size_t)K, V + 6 * (size_t)K);
    m.elem_s1.assign(elem_s, elem_s + K); m.elem_s2.assign(elem_s + K, elem_s + 2 * (size_t)K);
    m.elem_s3.assign(elem_s + 2 * (size_t)K, elem_s + 3 * (size_t)K);
    m.sx1.assign(sides, sides + S); m.sy1.assign(sides + S, sides + 2 * (size_t)S);
    m.sx2.assign(sides + 2 * (size_t)S, sides + 3 * (size_t)S); m.sy2.assign(sides + 3 * (size_t)S, sides + 4 * (size_t)S);
    m.left_elem.assign(side_int, side_int + S); m.right_elem.assign(side_int + S, side_int + 2 * (size_t)S);
    m.left_side.assign(side_int + 2 * (size_t)S, side_int + 3 * (size_t)S);
    m.right_side.assign(side_int + 3 * (size_t)S, side_int + 4 * (size_t)S);
    m.period_x = period_x;
    m.period_y = period_y;
    std::string why = validate(m);
    if (!why.empty()) {
        delete h;
        return fail(DG_ERR_ARG, why);
    }
    *out = h;
    return DG_OK;
    GUARD_END
}

int dg_mesh_periodic_jittered(int nx, int ny, double x0, double x1, double y0, double y1, double jitter, uint64_t seed,
                              dg_mesh** out) {
    GUARD_BEGIN
    if (!out) return fail(DG_ERR_ARG, "null argument");
    auto* h = new dg_mesh();
    std::string why = make_periodic_jittered(nx, ny, x0, x1, y0, y1, jitter, seed, h->m);
    if (!why.empty()) {
        delete h;
        return fail(DG_ERR_ARG, why);
    }
    *out = h;
    return DG_OK;
    GUARD_END
}

int dg_mesh_write_pmsh(const dg_mesh* m, const char* path) {
    GUARD_BEGIN
    if (!m || !path) return fail(DG_ERR_ARG, "null argument");
    std::string why = write_pmsh(path, m->m);
    return why.empty() ? DG_OK : fail(DG_ERR_IO, why);
    GUARD_END
}

int dg_mesh_info(const dg_mesh* m, int* num_elem, int* num_sides, int* num_boundary_sides) {
    if (!m) return fail(DG_ERR_ARG, "null mesh");
    if (num_elem) *num_elem = m->m.K;
    if (num_sides) *num_sides = m->m.S;
    if (num_boundary_sides) {
        int nb = 0;
        for (int r : m->m.right_elem) nb += (r < 0);
        *num_boundary_sides = nb;
    }
    return DG_OK;
}

int dg_mesh_period(const dg_mesh* m, double* period_x, double* period_y) {
    if (!m || !period_x || !period_y) return fail(DG_ERR_ARG, "null argument");
    *period_x = m->m.period_x;
    *period_y = m->m.period_y;
    return DG_OK;
}

int dg_mesh_get_arrays(const dg_mesh* mh, double* V, int* elem_s, double* sides, int* side_int) {
    if (!mh) return fail(DG_ERR_ARG, "null mesh");
    const Mesh& m = mh->m;
    const size_t K = m.K, S = m.S;
    if (V) {
        const std::vector<double>* a[6] = {&m.V1x, &m.V1y, &m.V2x, &m.V2y, &m.V3x, &m.V3y};
        for (int i = 0; i < 6; i++) memcpy(V + i * K, a[i]->data(), K * sizeof(double));
    }
    if (elem_s) {
        const std::vector<int>* a[3] = {&m.elem_s1, &m.elem_s2, &m.elem_s3};
        for (int i = 0; i < 3; i++) memcpy(elem_s + i * K, a[i]->data(), K * sizeof(int));
    }
    if (sides) {
        const std::vector<double>* a[4] = {&m.sx1, &m.sy1, &m.sx2, &m.sy2};
        for (int i = 0; i < 4; i++) memcpy(sides + i * S, a[i]->data(), S * sizeof(double));
    }
    if (side_int) {
        const std::vector<int>* a[4] = {&m.left_elem, &m.right_elem, &m.left_side, &m.right_side};
        for (int i = 0; i < 4; i++) memcpy(side_int + i * S, a[i]->data(), S * sizeof(int));
    }
    return DG_OK;
}

int dg_mesh_geometry(const dg_mesh* mh, int quirks, double* geo_e, double* geo_s, double* min_r) {
    GUARD_BEGIN
    if (!mh) return fail(DG_ERR_ARG, "null mesh");
    std::vector<double> ge, gs;
    double r = mesh_geometry(mh->m, quirks, ge, gs);
    if (geo_e) memcpy(geo_e, ge.data(), ge.size() * sizeof(double));
    if (geo_s) memcpy(geo_s, gs.data(), gs.size() * sizeof(double));
    if (min_r) *min_r = r;
    return DG_OK;
    GUARD_END
}

void dg_mesh_destroy(dg_mesh* m) { delete m; }

int dg_write_msh(const dg_mesh* mh, int var, const double* uv, const char* path) {
    GUARD_BEGIN
    if (!mh || !uv || !path) return fail(DG_ERR_ARG, "null argument");
    const Mesh& m = mh->m;
    FILE* f = fopen(path, "w");
    if (!f) return fail(DG_ERR_IO, std::string("cannot open ") + path + " for writing");
    fprintf(f, "View \"U%i \" {\n", var);
    for (int i = 0; i < m.K; i++)
        fprintf(f, "ST (%lf,%lf,0,%lf,%lf,0,%lf,%lf,0) {%lf,%lf,%lf};\n", m.V1x[i], m.V1y[i], m.V2x[i], m.V2y[i],
                m.V3x[i], m.V3y[i], uv[3 * (size_t)i], uv[3 * (size_t)i + 1], uv[3 * (size_t)i + 2]);
    fprintf(f, "};");
    const bool bad = ferror(f) != 0;
    fclose(f);
    if (bad) return fail(DG_ERR_IO, std::string("write error on ") + path);
    return DG_OK;
    GUARD_END
}

int dg_write_msh_view(const dg_mesh* mh, const char* name, int ncomp, const double* vals, const char* path) {
    GUARD_BEGIN
    if (!mh || !name || !vals || !path) return fail(DG_ERR_ARG, "null argument");
    if (ncomp != 1 && ncomp != 2) return fail(DG_ERR_ARG, "ncomp must be 1 (scalar) or 2 (vector)");
    const Mesh& m = mh->m;
    FILE* f = fopen(path, "w");
    if (!f) return fail(DG_ERR_IO, std::string("cannot open ") + path + " for writing");
    fprintf(f, "View \"%s\" {\n", name);
    for (int i = 0; i < m.K; i++) {
        if (ncomp == 1) {
            const double* v = vals + 3 * (size_t)i;
            fprintf(f, "ST (%lf,%lf,0,%lf,%lf,0,%lf,%lf,0) {%lf,%lf,%lf};\n", m.V1x[i], m.V1y[i], m.V2x[i],
                    m.V2y[i], m.V3x[i], m.V3y[i], v[0], v[1], v[2]);
        } else {
            const double* v = vals + 6 * (size_t)i;
            fprintf(f, "VT (%lf,%lf,0,%lf,%lf,0,%lf,%lf,0) {%lf,%lf,0,%lf,%lf,0,%lf,%lf,0};\n", m.V1x[i], m.V1y[i],
                    m.V2x[i], m.V2y[i], m.V3x[i], m.V3y[i], v[0], v[1], v[2], v[3], v[4], v[5]);
        }
    }
    fprintf(f, "};");
    const bool bad = ferror(f) != 0;
    fclose(f);
    if (bad) return fail(DG_ERR_IO, std::string("write error on ") + path);
    return DG_OK;
    GUARD_END
}

int dg_l2_error(const dg_mesh* mh, const dg_problem_desc* prob, int p, const double* c, double* err) {
    GUARD_BEGIN
    if (!mh || !prob || !c || !err) return fail(DG_ERR_ARG, "null argument");
    Status st = check_problem(*prob, p, DG_FP64);
    if (!st) return fail(st.code, st.msg);
    const Mesh& m = mh->m;
    // a rule of higher degree than the projection's: with as many points as
    // modes (p <= 2) the projection interpolates at its own rule's points.
    // The graded basis is hierarchical, so the first n_p rows of the higher
    // order's table are this order's basis.
    const Tables T = build_tables(std::min(p + 2, 5));
    const IcP ic = make_ic(*prob, m.period_x, m.period_y);
    const int K = m.K, N = prob->physics, np = n_modes(p), nq = T.n_q;
    std::vector<double> ge, gs;
    mesh_geometry(m, prob->quirks, ge, gs);
    std::vector<long double> acc(N, 0.0L);
    for (int e = 0; e < K; e++) {
        const double J = ge[e];
        for (int q = 0; q < nq; q++) {
            const double r1 = T.r1[q], r2 = T.r2[q];
            const double x = r1 * m.V2x[e] + r2 * m.V3x[e] + (1 - r1 - r2) * m.V1x[e];
            const double y = r1 * m.V2y[e] + r2 * m.V3y[e] + (1 - r1 - r2) * m.V1y[e];
            double U[4];
            ic_point(ic, x, y, U);
            for (int v = 0; v < N; v++) {
                double uh = 0;
                for (int i = 0; i < np; i++) uh += c[((size_t)v * np + i) * K + e] * T.basis[(size_t)i * nq + q];
                const double d = uh - U[v];
                acc[v] += (long double)(J * T.w[q] * d * d);
            }
        }
    }
    for (int v = 0; v < N; v++) err[v] = std::sqrt((double)acc[v]);
    return DG_OK;
    GUARD_END
}

int dg_tables(int p, int* sizes, double* r1, double* r2, double* w, double* s_r, double* w1, double* basis, double* gx,
              double* gy, double* side, double* vertex) {
    GUARD_BEGIN
    if (p < 0 || p > 5) return fail(DG_ERR_ARG, "order p must be in 0..5");
    Tables t = build_tables(p);
    if (sizes) {
        sizes[0] = t.n_p;
        sizes[1] = t.n_q;
        sizes[2] = t.n_q1;
    }
    auto cp = [](double* d, const std::vector<double>& s) {
        if (d) memcpy(d, s.data(), s.size() * sizeof(double));
    };
    cp(r1, t.r1); cp(r2, t.r2); cp(w, t.w); cp(s_r, t.s_r); cp(w1, t.w1);
    cp(basis, t.basis); cp(gx, t.gx); cp(gy, t.gy); cp(side, t.side); cp(vertex, t.vertex);
    return DG_OK;
    GUARD_END
}

int dg_solver_create(const dg_mesh* m, const dg_problem_desc* prob, int p, int precision, dg_solver** out) {
    GUARD_BEGIN
    if (!m || !prob || !out) return fail(DG_ERR_ARG, "null argument");
    Solver* s = nullptr;
    Status st = Solver::create(m->m, *prob, p, precision, &s);
    if (!st) return ret(st);
    auto* h = new dg_solver();
    h->s = s;
    *out = h;
    return DG_OK;
    GUARD_END
}

void dg_solver_destroy(dg_solver* s) {
    if (!s) return;
    delete s->s;
    delete s;
}

#define SOLVER_CHECK(h_) \
    if (!(h_) || !(h_)->s) return fail(DG_ERR_ARG, "null solver");

int dg_solver_info(const dg_solver* s, int* sizes, double* min_r) {
    SOLVER_CHECK(s);
    const Solver* v = s->s;
    if (sizes) {
        sizes[0] = v->K; sizes[1] = v->S; sizes[2] = v->NP; sizes[3] = v->NQ; sizes[4] = v->NQ1; sizes[5] = v->N;
    }
    if (min_r) *min_r = v->min_r;
    return DG_OK;
}

int dg_project_ic(dg_solver* s) {
    SOLVER_CHECK(s);
    GUARD_BEGIN return ret(s->s->project_ic()); GUARD_END
}

int dg_set_state(dg_solver* s, const double* c) {
    SOLVER_CHECK(s);
    if (!c) return fail(DG_ERR_ARG, "null state");
    GUARD_BEGIN return ret(s->s->set_state(c)); GUARD_END
}

int dg_get_state(dg_solver* s, double* c) {
    SOLVER_CHECK(s);
    if (!c) return fail(DG_ERR_ARG, "null state");
    GUARD_BEGIN return ret(s->s->get_state(c)); GUARD_END
}

int dg_rhs(dg_solver* s, const double* u, double* k, double t, double dt) {
    (void)t;
    SOLVER_CHECK(s);
    if (!u || !k) return fail(DG_ERR_ARG, "null buffer");
    GUARD_BEGIN return ret(s->s->rhs(u, k, dt)); GUARD_END
}

int dg_rhs_device_noalloc(dg_solver* s, const double* u_dev, double* k_dev, double dt, void* stream) {
    SOLVER_CHECK(s);
    if (!u_dev || !k_dev) return fail(DG_ERR_ARG, "null buffer");
    GUARD_BEGIN return ret(s->s->rhs_device_noalloc(u_dev, k_dev, dt, static_cast<hipStream_t>(stream))); GUARD_END
}

int dg_rk_step(dg_solver* s, double dt, int scheme) {
    SOLVER_CHECK(s);
    GUARD_BEGIN return ret(s->s->rk_step(dt, scheme)); GUARD_END
}

int dg_cfl_dt(dg_solver* s, double cfl, double* dt) {
    SOLVER_CHECK(s);
    if (!dt) return fail(DG_ERR_ARG, "null dt");
    GUARD_BEGIN return ret(s->s->cfl_dt(cfl, dt)); GUARD_END
}

int dg_integrate(dg_solver* s, double t_end, double cfl, int scheme, int64_t* steps, double* t_out) {
    SOLVER_CHECK(s);
    GUARD_BEGIN
    int64_t n = 0;
    int rc = ret(s->s->integrate(t_end, cfl, scheme, &n));
    if (steps) *steps = n;
    if (t_out) *t_out = s->s->t;
    return rc;
    GUARD_END
}

int dg_run_steps(dg_solver* s, int nsteps, double cfl, int scheme, int use_graph) {
    SOLVER_CHECK(s);
    if (nsteps < 0) return fail(DG_ERR_ARG, "nsteps must be >= 0");
    GUARD_BEGIN return ret(s->s->run_steps(nsteps, cfl, scheme, use_graph != 0)); GUARD_END
}

int dg_audit(dg_solver* s, dg_audit_out* out) {
    SOLVER_CHECK(s);
    if (!out) return fail(DG_ERR_ARG, "null output");
    GUARD_BEGIN return ret(s->s->audit(out)); GUARD_END
}

int dg_eval_vertices(dg_solver* s, int var, double* uv) {
    SOLVER_CHECK(s);
    if (!uv) return fail(DG_ERR_ARG, "null output");
    GUARD_BEGIN return ret(s->s->eval_vertices(var, uv)); GUARD_END
}

// vertex values of every variable: [var][K][3]
static int vertex_values(dg_solver* s, std::vector<double>& w) {
    const int K = s->s->K;
    w.assign(4 * 3 * (size_t)K, 0.0);
    for (int v = 0; v < 4; v++) {
        const int rc = ret(s->s->eval_vertices(v, w.data() + 3 * (size_t)K * v));
        if (rc != DG_OK) return rc;
    }
    return DG_OK;
}

int dg_eval_vertices_velocity(dg_solver* s, double* uv) {
    SOLVER_CHECK(s);
    if (!uv) return fail(DG_ERR_ARG, "null output");
    if (s->s->N != 4) return fail(DG_ERR_ARG, "velocity views are for Euler");
    GUARD_BEGIN
    std::vector<double> w;
    const int rc = vertex_values(s, w);
    if (rc != DG_OK) return rc;
    const size_t K3 = 3 * (size_t)s->s->K;
    for (size_t k = 0; k < K3; k++) {
        uv[2 * k] = w[K3 + k] / w[k];
        uv[2 * k + 1] = w[2 * K3 + k] / w[k];
    }
    return DG_OK;
    GUARD_END
}

int dg_eval_vertices_pressure(dg_solver* s, double* pv) {
    SOLVER_CHECK(s);
    if (!pv) return fail(DG_ERR_ARG, "null output");
    if (s->s->N != 4) return fail(DG_ERR_ARG, "pressure views are for Euler");
    GUARD_BEGIN
    std::vector<double> w;
    const int rc = vertex_values(s, w);
    if (rc != DG_OK) return rc;
    const size_t K3 = 3 * (size_t)s->s->K;
    const double g = s->s->prob.gamma;
    for (size_t k = 0; k < K3; k++) {
        const double rho = w[k], u = w[K3 + k] / rho, v = w[2 * K3 + k] / rho, E = w[3 * K3 + k];
        pv[k] = (g - 1.) * (E - (u * u + v * v) / 2. * rho);
    }
    return DG_OK;
    GUARD_END
}

int dg_solver_sync(dg_solver* s) {
    SOLVER_CHECK(s);
    return ret(s->s->sync());
}

int dg_solver_set_kernel(dg_solver* s, int kind) {
    SOLVER_CHECK(s);
    GUARD_BEGIN return ret(s->s->set_kernel(kind)); GUARD_END
}

int dg_solver_get_kernel(const dg_solver* s, int* kind) {
    SOLVER_CHECK(s);
    if (kind) *kind = s->s->kernel_kind();
    return DG_OK;
}

int dg_solver_set_max_groups(dg_solver* s, int max_groups) {
    SOLVER_CHECK(s);
    GUARD_BEGIN return ret(s->s->set_max_groups(max_groups)); GUARD_END
}

int dg_solver_launch_info(const dg_solver* s, int* out4) {
    SOLVER_CHECK(s);
    if (!out4) return fail(DG_ERR_ARG, "null output");
    s->s->launch_info(out4);
    return DG_OK;
}

int dg_time_stage_kernels(dg_solver* s, int nsteps, double cfl, double* out) {
    SOLVER_CHECK(s);
    if (nsteps <= 0 || !out) return fail(DG_ERR_ARG, "invalid argument");
    GUARD_BEGIN return ret(s->s->time_stage_kernels(nsteps, cfl, out)); GUARD_END
}

int dg_time_step_graph(dg_solver* s, int nsteps, double cfl, double* out) {
    SOLVER_CHECK(s);
    if (nsteps <= 0 || !out) return fail(DG_ERR_ARG, "invalid argument");
    GUARD_BEGIN return ret(s->s->time_step_graph(nsteps, cfl, out)); GUARD_END
}

int dg_solver_device_bytes(const dg_solver* s, int64_t* bytes) {
    SOLVER_CHECK(s);
    if (bytes) *bytes = s->s->device_bytes();
    return DG_OK;
}

int dg_bandwidth_probe(int64_t bytes, int iters, double* out) { return dg_bandwidth_probe_ex(bytes, iters, out, 2); }

int dg_bandwidth_probe_ex(int64_t bytes, int iters, double* out, int nout) {
    GUARD_BEGIN
    if (bytes <= 0 || bytes % 16 || iters <= 0 || !out) return fail(DG_ERR_ARG, "invalid argument");
    Status st = ensure_device_ready();
    if (!st) return ret(st);
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        if (a) hipFree(a);
        if (b) hipFree(b);
        if (e0) hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
        if (s) hipStreamDestroy(s);
    };
    hipError_t e = hipMalloc(&a, bytes);
    if (e == hipSuccess) e = hipMalloc(&b, bytes);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemsetAsync(a, 0, bytes, s);
    float best = 0;
    int best_v = 0;
    for (int v = 0; v < STREAM_COPY_VARIANTS && e == hipSuccess; v++) {
        e = launch_stream_copy(a, b, bytes, s, v);  // warm-up
        if (e == hipSuccess) e = hipEventRecord(e0, s);
        for (int i = 0; i < iters && e == hipSuccess; i++) e = launch_stream_copy(a, b, bytes, s, v);
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e == hipSuccess && (v == 0 || ms < best)) {
            best = ms;
            best_v = v;
        }
    }
    cleanup();
    if (e != hipSuccess) return ret(hip_status(e, "bandwidth probe"));
    out[0] = best / iters;
    out[1] = 2.0 * (double)bytes / (out[0] * 1e-3) / 1e9;
    if (nout >= 3) out[2] = best_v;
    return DG_OK;
    GUARD_END
}

int dg_selftest_mfma(double* max_err) {
    GUARD_BEGIN
    if (!max_err) return fail(DG_ERR_ARG, "null output");
    Status st = ensure_device_ready();
    if (!st) return ret(st);
    float d[256];
    double d64[256];
    hipError_t e = mfma_layout_probe(d);
    if (e == hipSuccess) e = mfma_layout_probe_f64(d64);
    if (e != hipSuccess) return ret(hip_status(e, "mfma probe"));
    double err = 0;  // both products, as laid out by the kernels' lane maps
    for (int i = 0; i < 16; i++)
        for (int j = 0; j < 16; j++) {
            double want = 0;
            for (int k = 0; k < 4; k++) want += ((i + 1) + 0.5 * k) * ((2 * k + 1) - 0.25 * j);
            err = std::max(err, std::fabs(want - (double)d[i * 16 + j]));
            err = std::max(err, std::fabs(want - d64[i * 16 + j]));
        }
    *max_err = err;
    return DG_OK;
    GUARD_END
}

/* ---------------- batch ---------------- */
int dg_batch_create(int precision, dg_batch** out) {
    GUARD_BEGIN
    if (!out) return fail(DG_ERR_ARG, "null argument");
    Batch* b = nullptr;
    Status st = Batch::create(precision, &b);
    if (!st) return ret(st);
    auto* h = new dg_batch();
    h->b = b;
    *out = h;
    return DG_OK;
    GUARD_END
}

void dg_batch_destroy(dg_batch* b) {
    if (!b) return;
    delete b->b;
    delete b;
}

#define BATCH_CHECK(b) \
    if (!(b) || !(b)->b) return fail(DG_ERR_ARG, "null batch");

int dg_batch_add(dg_batch* b, const dg_mesh* m, const dg_problem_desc* prob, int p, double t_end, int* slot) {
    BATCH_CHECK(b);
    if (!m || !prob) return fail(DG_ERR_ARG, "null argument");
    GUARD_BEGIN
    int sl = -1;
    int rc = ret(b->b->add(m->m, *prob, p, t_end, &sl));
    if (slot) *slot = sl;
    return rc;
    GUARD_END
}

int dg_batch_project_ic(dg_batch* b) {
    BATCH_CHECK(b);
    GUARD_BEGIN return ret(b->b->project_ic()); GUARD_END
}

int dg_batch_run_steps(dg_batch* b, int nsteps, double cfl, int use_graph) {
    BATCH_CHECK(b);
    GUARD_BEGIN return ret(b->b->run_steps(nsteps, cfl, use_graph != 0)); GUARD_END
}

int dg_batch_integrate(dg_batch* b, double cfl, int use_graph, int64_t* steps) {
    BATCH_CHECK(b);
    GUARD_BEGIN
    int64_t n = 0;
    int rc = ret(b->b->integrate(cfl, use_graph != 0, &n));
    if (steps) *steps = n;
    return rc;
    GUARD_END
}

int dg_batch_get_state(dg_batch* b, int slot, double* c) {
    BATCH_CHECK(b);
    if (!c) return fail(DG_ERR_ARG, "null state");
    GUARD_BEGIN return ret(b->b->get_state(slot, c)); GUARD_END
}

int dg_batch_audit(dg_batch* b, int slot, dg_audit_out* out) {
    BATCH_CHECK(b);
    if (!out) return fail(DG_ERR_ARG, "null output");
    GUARD_BEGIN return ret(b->b->audit(slot, out)); GUARD_END
}

int dg_batch_info(const dg_batch* b, int* n_solves, int64_t* total_elems, int64_t* device_bytes) {
    BATCH_CHECK(b);
    b->b->info(n_solves, total_elems, device_bytes);
    return DG_OK;
}

int dg_batch_set_max_groups(dg_batch* b, int max_groups) {
    BATCH_CHECK(b);
    GUARD_BEGIN return ret(b->b->set_max_groups(max_groups)); GUARD_END
}

int dg_batch_set_launch_mode(dg_batch* b, int mode) {
    BATCH_CHECK(b);
    if (mode != DG_BATCH_ONE_LAUNCH && mode != DG_BATCH_PER_CLASS)
        return fail(DG_ERR_ARG, "mode must be DG_BATCH_ONE_LAUNCH or DG_BATCH_PER_CLASS");
    GUARD_BEGIN b->b->set_launch_form(mode);
    return DG_OK;
    GUARD_END
}

int dg_batch_launch_info(const dg_batch* b, int* out4) {
    BATCH_CHECK(b);
    if (!out4) return fail(DG_ERR_ARG, "null output");
    b->b->launch_info(out4);
    return DG_OK;
}

/* ---------------- server ---------------- */
int dg_server_create(const dg_server_config* cfg, dg_server** out) {
    GUARD_BEGIN
    if (!cfg || !out) return fail(DG_ERR_ARG, "null argument");
    Server* s = nullptr;
    Status st = Server::create(*cfg, &s);
    if (!st) return ret(st);
    auto* h = new dg_server();
    h->s = s;
    *out = h;
    return DG_OK;
    GUARD_END
}

#define SERVER_CHECK(h_) \
    if (!(h_) || !(h_)->s) return fail(DG_ERR_ARG, "null server");

int dg_server_submit(dg_server* sv, const dg_request* req, uint64_t* id) {
    SERVER_CHECK(sv);
    if (!req) return fail(DG_ERR_ARG, "null request");
    GUARD_BEGIN return ret(sv->s->submit(*req, id)); GUARD_END
}

int dg_server_poll(dg_server* sv, dg_result* out, int max, int* n) {
    SERVER_CHECK(sv);
    GUARD_BEGIN return ret(sv->s->poll(out, max, n)); GUARD_END
}

int dg_server_get_state(dg_server* sv, uint64_t id, double* c, int64_t n) {
    SERVER_CHECK(sv);
    GUARD_BEGIN return ret(sv->s->get_state(id, c, n)); GUARD_END
}

int dg_server_keep_states(dg_server* sv, int keep) {
    SERVER_CHECK(sv);
    sv->s->keep_states(keep != 0);
    return DG_OK;
}

int dg_server_drain(dg_server* sv) {
    SERVER_CHECK(sv);
    GUARD_BEGIN return ret(sv->s->drain()); GUARD_END
}

double dg_server_now(dg_server* sv) {
    if (!sv || !sv->s) return 0.0;
    return sv->s->now();
}

void dg_server_destroy(dg_server* sv) {
    if (!sv) return;
    delete sv->s;
    delete sv;
}

}  // extern "C"
