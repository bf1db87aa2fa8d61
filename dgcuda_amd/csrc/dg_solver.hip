// Host driver of one solve: device buffers, stage sequencing, device-side
// CFL time step, hipGraph replay of a step, state I/O and audits.
//
// The reference's time loop (src/2d/conservation/time_integrator.cu:221-434)
// launches four kernels per stage and copies every element's wave speed to
// the host each step to pick dt.  Here a step is one small dt kernel plus one
// fused launch per stage; dt never leaves the device, and the host only
// synchronises once per chunk of steps to test for t >= t_end.
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstdio>
#include <tuple>
#include <map>
#include <cmath>
#include <cstring>
#include <mutex>

#include "dg_solver.h"

namespace dgmi {

Status hip_status(hipError_t e, const char* what) {
    if (e == hipSuccess) return Status::ok();
    return Status::err(DG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPST(expr)                                          \
    do {                                                     \
        hipError_t _e = (expr);                              \
        if (_e != hipSuccess) return hip_status(_e, #expr);  \
    } while (0)

// Scratch of the allocation-free right-hand side (rhs_device_noalloc): the
// stage input and output in the solver's state layout.  A file-scope device
// array of fixed size; each device that the library sets up holds its own
// copy until the process ends.  In host code the array's name is not a
// device address: its address on each device is taken once, at the
// device's set-up below, and only that address is used.
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
static std::atomic<void*> g_scratch[64];

void* noalloc_scratch(int dev) { return (dev >= 0 && dev < 64) ? g_scratch[dev].load() : nullptr; }

Status ensure_device_ready() {
    static std::mutex mu;
    static bool done[64] = {false};
    std::lock_guard<std::mutex> lk(mu);
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return Status::err(DG_ERR_HIP, "no GPU device visible");
    int dev = 0;
    HIPST(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return Status::err(DG_ERR_HIP, "device id out of range");
    if (!done[dev]) {
        HIPST(upload_tables());
        void* ptr = nullptr;
        HIPST(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch[dev].store(ptr);
        done[dev] = true;
    }
    return Status::ok();
}

namespace {
struct Pool {
    std::mutex mu;
    std::map<void*, std::pair<int, size_t>> live;           // block -> (device, bytes)
    std::multimap<std::pair<int, size_t>, void*> idle;      // (device, bytes) -> block
    std::map<int, size_t> idle_bytes;
    std::map<int, std::vector<hipStream_t>> streams;
};
Pool& pool() {
    static Pool* p = new Pool();  // never destroyed: blocks live until the process ends
    return *p;
}
constexpr size_t POOL_IDLE_CAP = size_t(16) << 30;
size_t pool_round(size_t n) {
    const size_t g = n <= (size_t(1) << 20) ? 4096 : (size_t(2) << 20);
    return (n + g - 1) / g * g;
}
}  // namespace

Status pool_alloc(void** out, size_t bytes) {
    int dev = 0;
    HIPST(hipGetDevice(&dev));
    const size_t need = pool_round(std::max<size_t>(bytes, 1));
    Pool& P = pool();
    {
        std::lock_guard<std::mutex> lk(P.mu);
        auto it = P.idle.lower_bound(std::make_pair(dev, need));
        if (it != P.idle.end() && it->first.first == dev && it->first.second <= need + need / 2) {
            *out = it->second;
            P.live[it->second] = it->first;
            P.idle_bytes[dev] -= it->first.second;
            P.idle.erase(it);
            return Status::ok();
        }
    }
    HIPST(hipMalloc(out, need));
    std::lock_guard<std::mutex> lk(P.mu);
    P.live[*out] = std::make_pair(dev, need);
    return Status::ok();
}

void pool_free(void* p) {
    if (!p) return;
    Pool& P = pool();
    std::vector<void*> release;
    {
        std::lock_guard<std::mutex> lk(P.mu);
        auto it = P.live.find(p);
        if (it == P.live.end()) return;
        const auto key = it->second;
        P.live.erase(it);
        P.idle.emplace(key, p);
        size_t& ib = P.idle_bytes[key.first];
        ib += key.second;
        // over the cap: give back the largest idle blocks of this device
        while (ib > POOL_IDLE_CAP) {
            auto last = P.idle.lower_bound(std::make_pair(key.first + 1, size_t(0)));
            if (last == P.idle.begin()) break;
            --last;
            if (last->first.first != key.first) break;
            ib -= last->first.second;
            release.push_back(last->second);
            P.idle.erase(last);
        }
    }
    for (void* q : release) (void)hipFree(q);
}

Status pool_stream(hipStream_t* s) {
    int dev = 0;
    HIPST(hipGetDevice(&dev));
    Pool& P = pool();
    {
        std::lock_guard<std::mutex> lk(P.mu);
        auto& v = P.streams[dev];
        if (!v.empty()) {
            *s = v.back();
            v.pop_back();
            return Status::ok();
        }
    }
    HIPST(hipStreamCreateWithFlags(s, hipStreamNonBlocking));
    return Status::ok();
}

void pool_release_stream(hipStream_t s) {
    if (!s) return;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    Pool& P = pool();
    std::lock_guard<std::mutex> lk(P.mu);
    P.streams[dev].push_back(s);
}

PhysP make_phys(const dg_problem_desc& d) {
    PhysP ph{};
    ph.gamma = d.gamma;
    ph.ax = d.ax;
    ph.ay = d.ay;
    ph.quirks = d.quirks;
    ph.flux = d.flux;
    return ph;
}

IcP make_ic(const dg_problem_desc& d, double Lx, double Ly) {
    IcP ic{};
    ic.kind = d.ic;
    ic.nv = d.n_vortex;
    ic.gamma = d.gamma;
    ic.gx = d.gauss_x;
    ic.gy = d.gauss_y;
    ic.gk = d.gauss_k;
    ic.Lx = Lx;
    ic.Ly = Ly;
    ic.rho_inf = d.rho_inf;
    ic.u_inf = d.u_inf;
    ic.v_inf = d.v_inf;
    ic.p_inf = d.p_inf;
    for (int k = 0; k < 4; k++) {
        ic.beta[k] = d.vortex_beta[k];
        ic.vx[k] = d.vortex_x[k];
        ic.vy[k] = d.vortex_y[k];
        ic.constant[k] = d.constant[k];
    }
    ic.mach = d.mach;
    ic.fpow = (d.quirks & DG_QUIRK_FLOAT_POW) ? 1 : 0;
    ic.alpha = d.poly_alpha;
    return ic;
}

// Per boundary side and face point, 4 FP64 values for the lane kernel
// (boundary_ghost): inflow and outflow (euler.cu:177-241) take the problem's
// state U0 at the point; reflecting (supersonic.cu:64-108,
// cylinder.cu:58-101) the unit direction from the problem's wall centre to
// the point, turned to the side's normal, as the wall normal.  The
// point is the left element's face point j mapped by the reference's
// formula x = x2 r + x3 s + x1 (1 - r - s): face 0 (0.5 + 0.5 r_j, 0), face
// 1 ((1 - r_j)/2, (1 + r_j)/2), face 2 (0, 0.5 + 0.5 r_{n-1-j}); the
// reflecting side maps face 1 as ((1 + r_j)/2, (1 - r_j)/2) under
// DG_QUIRK_REFLECT_MIRROR, as the reference does, and takes the radial
// length through a float square root under DG_QUIRK_FLOAT_SQRT.
std::vector<double> boundary_data(const ElemGeometry& g, const Tables& T, const IcP& ic, int quirks, double wall_x,
                                  double wall_y) {
    const int n1 = T.n_q1;
    std::vector<double> out((size_t)g.bfaces.size() * n1 * 4, 0.0);
    for (size_t b = 0; b < g.bfaces.size(); b++) {
        const int e = g.bfaces[b].elem, f = g.bfaces[b].face, kind = g.bfaces[b].kind;
        const double v1x = g.V[0][e], v1y = g.V[1][e], v2x = g.V[2][e], v2y = g.V[3][e];
        const double v3x = g.V[4][e], v3y = g.V[5][e];
        const bool mirror = kind == BC_REFLECTING && (quirks & DG_QUIRK_REFLECT_MIRROR);
        for (int j = 0; j < n1; j++) {
            double r1, r2;
            switch (f) {
                case 0: r1 = 0.5 + 0.5 * T.s_r[j]; r2 = 0.; break;
                case 1:
                    if (mirror) { r1 = (1. + T.s_r[j]) / 2.; r2 = (1. - T.s_r[j]) / 2.; }
                    else { r1 = (1. - T.s_r[j]) / 2.; r2 = (1. + T.s_r[j]) / 2.; }
                    break;
                default: r1 = 0.; r2 = 0.5 + 0.5 * T.s_r[n1 - 1 - j]; break;
            }
            const double x = v2x * r1 + v3x * r2 + v1x * (1 - r1 - r2);
            const double y = v2y * r1 + v3y * r2 + v1y * (1 - r1 - r2);
            double* o = &out[(b * n1 + j) * 4];
            if (kind == BC_REFLECTING) {
                const double xw = x - wall_x, yw = y - wall_y;
                const double d2 = xw * xw + yw * yw;
                const double dot = (quirks & DG_QUIRK_FLOAT_SQRT) ? (double)std::sqrt((float)d2) : std::sqrt(d2);
                double Nx = xw / dot, Ny = yw / dot;
                if (Nx * g.fnx[f][e] + Ny * g.fny[f][e] < 0) {
                    Nx *= -1;
                    Ny *= -1;
                }
                o[0] = Nx;
                o[1] = Ny;
            } else {
                ic_point(ic, x, y, o);
            }
        }
    }
    return out;
}

Status check_problem(const dg_problem_desc& d, int p, int precision) {
    if (p < 0 || p > 5) return Status::err(DG_ERR_ARG, "order p must be in 0..5");
    if (precision != DG_FP64 && precision != DG_FP32) return Status::err(DG_ERR_ARG, "precision must be DG_FP64 or DG_FP32");
    if (d.physics != DG_PHYS_ADVECTION && d.physics != DG_PHYS_EULER)
        return Status::err(DG_ERR_ARG, "physics must be DG_PHYS_ADVECTION or DG_PHYS_EULER");
    if (d.ic < 0 || d.ic > 4) return Status::err(DG_ERR_ARG, "unknown initial condition");
    if (d.ic == DG_IC_POLY && (d.poly_alpha < 0 || d.poly_alpha > 8))
        return Status::err(DG_ERR_ARG, "poly_alpha must be in 0..8");
    if (d.flux != DG_FLUX_LLF && d.flux != DG_FLUX_CENTRAL) return Status::err(DG_ERR_ARG, "unknown flux");
    if (d.flux == DG_FLUX_CENTRAL && d.physics != DG_PHYS_ADVECTION)
        return Status::err(DG_ERR_ARG, "the central flux is for advection");
    if (d.ic == DG_IC_SUPERSONIC && (d.physics != DG_PHYS_EULER || !(d.mach > 0.0)))
        return Status::err(DG_ERR_ARG, "supersonic vortex needs Euler physics and a positive Mach number");
    if (d.physics == DG_PHYS_ADVECTION && d.ic == DG_IC_VORTEX)
        return Status::err(DG_ERR_ARG, "vortex initial condition needs Euler physics");
    if (d.physics == DG_PHYS_EULER && d.ic == DG_IC_GAUSSIAN)
        return Status::err(DG_ERR_ARG, "Gaussian pulse initial condition is for advection");
    if (d.ic == DG_IC_VORTEX && (d.n_vortex < 1 || d.n_vortex > DG_MAX_VORTEX))
        return Status::err(DG_ERR_ARG, "n_vortex must be in 1..4");
    if (d.physics == DG_PHYS_EULER && !(d.gamma > 1.0)) return Status::err(DG_ERR_ARG, "gamma must exceed 1");
    if (d.quirks & ~15) return Status::err(DG_ERR_ARG, "unknown quirk bits");
    if (!std::isfinite(d.wall_x) || !std::isfinite(d.wall_y)) return Status::err(DG_ERR_ARG, "wall centre not finite");
    return Status::ok();
}

// process-wide cache of automatic kernel choices
static std::mutex& tune_mu() {
    static std::mutex m;
    return m;
}
static std::map<std::tuple<int, int, int, int, int>, int>& tune_cache() {
    static std::map<std::tuple<int, int, int, int, int>, int> c;
    return c;
}

template <typename R>
class SolverT : public Solver {
   public:
    std::vector<int> perm, inv;   // perm[internal] = caller's element id
    PhysP phys{};
    IcP ic{};
    hipStream_t stream = nullptr;
    R *c = nullptr, *acc = nullptr, *ua = nullptr, *ub = nullptr;
    R *geo_e = nullptr, *geo_f = nullptr;
    int* nbr = nullptr;
    int* perm_d = nullptr;  // [Kp]: caller's element of internal element i (padding: element K-1's)
    int device = -1;
    std::vector<int> h_nbr[3];
    uint32_t* fword = nullptr;
    int* halo = nullptr;
    void *face4 = nullptr, *metric4 = nullptr;
    int64_t tile_bytes = 0;  // fword, halo and packed geometry
    MfmaArgs mx{};
    double *V = nullptr, *Jd = nullptr, *audit_part = nullptr;
    double* bnd = nullptr;      // boundary sides' ghost data (boundary_data)
    bool has_boundary = false;  // the mesh has boundary sides (bnd holds their ghost data)
    Ctrl* ctrl = nullptr;
    Ctrl hctrl{};
    LaunchShape shape;
    int kernel_req = 0;
    int max_groups = 0;  // cap of the persistent matrix-core grid (0: every resident workgroup)
    bool lam_valid = false;
    hipGraphExec_t gexec = nullptr;
    int graph_scheme = 0;
    int64_t bytes = 0;
    int audit_blocks = 0;
    // layout of c, acc, ua, ub: mode-major for the lane kernel, the tile
    // layout of the precision for the matrix-core kernels; switched with the
    // kernel
    StateIx six{};
    bool state_set = false;  // c holds a state (projected, set or stepped)
    double b0 = 0;  // phi_0 at the first volume point (cell mean = c_0 * b0)
    bool geom_fp32 = false;  // FP64 solve with the geometry rounded through FP32 (DG_GEOM_FP32)

    ~SolverT() override {
        if (stream) (void)hipStreamSynchronize(stream);  // nothing queued may use the buffers
        if (gexec) (void)hipGraphExecDestroy(gexec);
        void* bufs[] = {c,  acc,        ua,   ub,    geo_e, geo_f, nbr,     V,      Jd,
                        audit_part, ctrl, fword, halo,  face4, metric4, bnd, perm_d};
        for (void* b : bufs) pool_free(b);
        pool_release_stream(stream);
    }

    // StateIx::tiled of this precision's tile layout
    static constexpr int tile_kind() { return sizeof(R) == 4 ? 1 : 2; }

    // entries allocated per state array: room for either layout
    size_t state_alloc() const {
        StateIx t{N, NP, Kp, mfma_supported(N, p) ? tile_kind() : 0};
        return t.elems();
    }

    Status alloc(void** p, size_t n) {
        Status st = pool_alloc(p, n);
        if (!st) return st;
        bytes += (int64_t)n;
        return Status::ok();
    }

    Status init(const Mesh& m, const dg_problem_desc& pr, int p_, int kernel) {
        prob = pr;
        p = p_;
        N = pr.physics;
        NP = n_modes(p);
        NQ = n_quad_vol(p);
        NQ1 = n_quad_face(p);
        K = m.K;
        S = m.S;
        Kp = (K + 63) / 64 * 64;
        precision = sizeof(R) == 8 ? DG_FP64 : DG_FP32;
        phys = make_phys(pr);
        ic = make_ic(pr, m.period_x, m.period_y);
        b0 = build_tables(p).basis[0];
        if (m.ordered) {
            perm.resize(K);
            for (int i = 0; i < K; i++) perm[i] = i;
        } else {
            perm = morton_order(m);
        }
        inv.resize(K);
        for (int i = 0; i < K; i++) inv[perm[i]] = i;
        ElemGeometry g;
        std::string why = elem_geometry(m, perm, pr.quirks, g);
        if (!why.empty()) return Status::err(DG_ERR_UNSUPPORTED, why);
        has_boundary = !g.bfaces.empty();
        if (has_boundary && p > VALU_BOUNDARY_MAX_P && !mfma_supported(N, p))
            return Status::err(DG_ERR_UNSUPPORTED,
                               "boundary sides are supported up to p = 4 (lane kernel up to p = 3, matrix-core "
                               "kernels up to p = 4)");
        min_r = g.min_r;
        Status st = ensure_device_ready();
        if (!st) return st;
        HIPST(hipGetDevice(&device));
        if (!(st = pool_stream(&stream))) return st;
        six = StateIx{N, NP, Kp, 0};
        const size_t M = state_alloc();
        if (!(st = alloc((void**)&c, M * sizeof(R)))) return st;
        if (!(st = alloc((void**)&acc, M * sizeof(R)))) return st;
        if (!(st = alloc((void**)&ua, M * sizeof(R)))) return st;
        if (!(st = alloc((void**)&ub, M * sizeof(R)))) return st;
        if (!(st = alloc((void**)&geo_e, 5 * (size_t)Kp * sizeof(R)))) return st;
        if (!(st = alloc((void**)&geo_f, 9 * (size_t)Kp * sizeof(R)))) return st;
        if (!(st = alloc((void**)&nbr, 3 * (size_t)Kp * sizeof(int)))) return st;
        if (!(st = alloc((void**)&V, 6 * (size_t)Kp * sizeof(double)))) return st;
        if (!(st = alloc((void**)&Jd, (size_t)Kp * sizeof(double)))) return st;
        audit_blocks = (K + 255) / 256;
        if (!(st = alloc((void**)&audit_part, (size_t)audit_blocks * 8 * sizeof(double)))) return st;
        if (!(st = alloc((void**)&ctrl, sizeof(Ctrl)))) return st;
        HIPST(hipMemsetAsync(c, 0, M * sizeof(R), stream));
        HIPST(hipMemsetAsync(acc, 0, M * sizeof(R), stream));
        HIPST(hipMemsetAsync(ua, 0, M * sizeof(R), stream));
        HIPST(hipMemsetAsync(ub, 0, M * sizeof(R), stream));
        // geometry, padded entries replicate the last element; the 5th
        // element value is the inverse mass 1/J (M = J*I), computed in FP64
        std::vector<R> he(5 * (size_t)Kp), hf(9 * (size_t)Kp);
        std::vector<int> hn(3 * (size_t)Kp);
        std::vector<double> hv(6 * (size_t)Kp), hj(Kp);
        const auto gr = [this](double x) { return (R)(geom_fp32 ? (double)(float)x : x); };
        for (int i = 0; i < Kp; i++) {
            const int e = std::min(i, K - 1);
            he[i] = gr(g.xr[e]);
            he[Kp + i] = gr(g.yr[e]);
            he[2 * (size_t)Kp + i] = gr(g.xs[e]);
            he[3 * (size_t)Kp + i] = gr(g.ys[e]);
            he[4 * (size_t)Kp + i] = gr(1.0 / g.J[e]);
            hj[i] = g.J[e];
            for (int f = 0; f < 3; f++) {
                hf[(3 * f + 0) * (size_t)Kp + i] = gr(g.fnx[f][e]);
                hf[(3 * f + 1) * (size_t)Kp + i] = gr(g.fny[f][e]);
                hf[(3 * f + 2) * (size_t)Kp + i] = gr(g.flen[f][e]);
                hn[f * (size_t)Kp + i] = g.nbr[f][e];
            }
            for (int v = 0; v < 6; v++) hv[v * (size_t)Kp + i] = g.V[v][e];
        }
        HIPST(hipMemcpyAsync(geo_e, he.data(), he.size() * sizeof(R), hipMemcpyHostToDevice, stream));
        HIPST(hipMemcpyAsync(geo_f, hf.data(), hf.size() * sizeof(R), hipMemcpyHostToDevice, stream));
        HIPST(hipMemcpyAsync(nbr, hn.data(), hn.size() * sizeof(int), hipMemcpyHostToDevice, stream));
        std::vector<int> hp(Kp);
        for (int i = 0; i < Kp; i++) hp[i] = perm[std::min(i, K - 1)];
        if (!(st = alloc((void**)&perm_d, (size_t)Kp * sizeof(int)))) return st;
        HIPST(hipMemcpyAsync(perm_d, hp.data(), hp.size() * sizeof(int), hipMemcpyHostToDevice, stream));
        for (int f = 0; f < 3; f++) h_nbr[f] = g.nbr[f];
        // host source of an async copy: lives until the stream is synchronised below
        std::vector<double> hb;
        if (has_boundary) {
            hb = boundary_data(g, build_tables(p), ic, pr.quirks, pr.wall_x, pr.wall_y);
            if (!(st = alloc((void**)&bnd, hb.size() * sizeof(double)))) return st;
            HIPST(hipMemcpyAsync(bnd, hb.data(), hb.size() * sizeof(double), hipMemcpyHostToDevice, stream));
        }
        HIPST(hipMemcpyAsync(V, hv.data(), hv.size() * sizeof(double), hipMemcpyHostToDevice, stream));
        HIPST(hipMemcpyAsync(Jd, hj.data(), hj.size() * sizeof(double), hipMemcpyHostToDevice, stream));
        hctrl = Ctrl{};
        hctrl.min_r = min_r;
        hctrl.two_p1 = 2. * p + 1.;
        if (!(st = choose_shape(kernel))) return st;
        if (!(st = push_ctrl())) return st;
        HIPST(hipStreamSynchronize(stream));
        return Status::ok();
    }

    // Matrix-core tiling: per (element, face) the LDS slot of the neighbour
    // (inside the tile, or in the tile's halo list), the halo lists, and the
    // tile's side list: the sides the tile computes (boundary and halo sides,
    // and of each side between two of its elements the left element's
    // record), numbered in record order.
    Status build_mfma_tiles(int tile_elems) {
        const int ntiles = (K + tile_elems - 1) / tile_elems;
        std::vector<uint32_t> words(3 * (size_t)Kp, 0);
        std::vector<std::vector<int>> halos(ntiles);
        int hmax = 0;
        for (int t = 0; t < ntiles; t++) {
            const int b = t * tile_elems, end = std::min(b + tile_elems, K);
            std::vector<int>& h = halos[t];
            uint32_t ntask = 0;
            for (int e = b; e < end; e++)
                for (int f = 0; f < 3; f++) {
                    const int wv = h_nbr[f][e];
                    uint32_t wd;
                    bool owner;
                    if (wv < 0) {
                        // boundary side: own slot and face, left, flag bit 15
                        // (the kernel reads the ghost data through the
                        // neighbour word)
                        wd = (uint32_t)((e - b) | (f << 12) | (1 << 14) | (1 << 15));
                        owner = true;
                    } else {
                        const int nb = wv & NB_ELEM_MASK, fn = (wv >> 28) & 3, left = (wv >> 30) & 1;
                        int slot;
                        if (nb >= b && nb < end) {
                            slot = nb - b;
                        } else {
                            auto it = std::find(h.begin(), h.end(), nb);
                            slot = tile_elems + (int)(it - h.begin());
                            if (it == h.end()) h.push_back(nb);
                        }
                        if (slot >= 4096) return Status::err(DG_ERR_UNSUPPORTED, "tile halo too large");
                        wd = (uint32_t)(slot | (fn << 12) | (left << 14));
                        owner = slot >= tile_elems || left;
                    }
                    if (owner) wd |= (ntask++ << FW_TASK_SHIFT) | FW_OWNER;
                    words[f * (size_t)Kp + e] = wd;
                }
            hmax = std::max(hmax, (int)h.size());
        }
        const int hcap = std::max(16, (hmax + 15) / 16 * 16);
        // LDS per workgroup: traces and face records, plus the 5-slot operand table
        if (hcap > MFMA_MAX_HALO || mfma_lds_bytes(N, hcap, (int)sizeof(R)) + 5 * 64 * 4 * sizeof(R) > 160 * 1024)
            return Status::err(DG_ERR_UNSUPPORTED, "tile halo too large for the matrix-core kernel");
        std::vector<int> hl((size_t)ntiles * hcap);
        for (int t = 0; t < ntiles; t++)
            for (int i = 0; i < hcap; i++)
                hl[(size_t)t * hcap + i] = i < (int)halos[t].size() ? halos[t][i] : t * tile_elems;
        HIPST(hipStreamSynchronize(stream));  // queued launches may still read the old lists
        for (void* b : {(void*)fword, (void*)halo, (void*)face4, (void*)metric4}) pool_free(b);
        fword = nullptr;
        halo = nullptr;
        face4 = nullptr;
        metric4 = nullptr;
        tile_bytes = 0;
        const size_t bw = words.size() * sizeof(uint32_t), bh = hl.size() * sizeof(int);
        const size_t bf = (size_t)Kp * 3 * 4 * sizeof(R), bm = (size_t)Kp * 4 * sizeof(R);
        Status sa = pool_alloc((void**)&fword, bw);
        if (sa) sa = pool_alloc((void**)&halo, bh);
        if (sa) sa = pool_alloc((void**)&face4, bf);
        if (sa) sa = pool_alloc((void**)&metric4, bm);
        if (!sa) return sa;
        tile_bytes = (int64_t)(bw + bh + bf + bm);
        HIPST(hipMemcpyAsync(fword, words.data(), bw, hipMemcpyHostToDevice, stream));
        HIPST(hipMemcpyAsync(halo, hl.data(), bh, hipMemcpyHostToDevice, stream));
        HIPST(launch_mfma_pack<R>(geo_e, geo_f, fword, K, Kp, face4, metric4, stream));
        HIPST(hipStreamSynchronize(stream));
        mx.fword = fword;
        mx.halo = halo;
        mx.face4 = face4;
        mx.metric4 = metric4;
        mx.hcap = hcap;
        return Status::ok();
    }

    // Launch shape of kernel kind 1 (lane per element), 2 (persistent
    // matrix-core tiles) or 3 (one matrix-core tile per workgroup).
    Status make_shape(int kind, LaunchShape& s) {
        const bool mf = mfma_supported(N, p);
        if ((kind == 2 || kind == 3) && !mf)
            return Status::err(DG_ERR_UNSUPPORTED, "matrix-core kernel needs p = 1..4");
        if (kind == 1 && has_boundary && p > VALU_BOUNDARY_MAX_P)
            return Status::err(DG_ERR_UNSUPPORTED, "lane kernel with boundary sides needs p <= 3");
        if (kind < 1 || kind > 3) return Status::err(DG_ERR_ARG, "kernel kind must be 0, 1, 2 or 3");
        s = LaunchShape{};
        s.kind = kind;
        s.tile_elems = (kind == 1) ? VALU_TILE_ELEMS : MFMA_TILE_ELEMS;
        s.ntiles = (K + s.tile_elems - 1) / s.tile_elems;
        s.xcd_remap = 1;
        s.grid = (s.ntiles + 7) / 8 * 8;
        if (kind >= 2) {
            Status st = build_mfma_tiles(s.tile_elems);
            if (!st) return st;
            if (kind == 2) {
                // persistent: as many workgroups as fit on the device at once
                const int res = mfma_resident_groups(N, p, mx.hcap, (int)sizeof(R), has_boundary);
                if (res >= 8) s.grid = std::min(s.grid, res / 8 * 8);
                if (max_groups >= 8) s.grid = std::min(s.grid, max_groups / 8 * 8);
            }
        }
        return Status::ok();
    }

    // Automatic kernel choice: time every kind this solve supports on its
    // own buffers (3 right-hand-side launches after a warm-up, HIP events)
    // and take the fastest; cached per (physics, p, precision, log2 K).
    // The candidates agree to the stated tolerances (FP32 rounding, or 1e-12
    // in FP64), so the choice changes speed, not results beyond those.
    Status autotune(int* out_kind) {
        *out_kind = 1;
        if (!mfma_supported(N, p)) return Status::ok();
        int lk = 0;
        while (lk < 30 && (1 << (lk + 1)) <= K) lk++;
        const auto key = std::make_tuple(N, p, (int)sizeof(R), lk, (int)has_boundary);
        {
            std::lock_guard<std::mutex> lk_(tune_mu());
            auto it = tune_cache().find(key);
            if (it != tune_cache().end()) {
                *out_kind = it->second;
                return Status::ok();
            }
        }
        // timed on the scratch buffers (zeros), so a held state is untouched
        HIPST(hipMemsetAsync(ub, 0, state_alloc() * sizeof(R), stream));
        HIPST(launch_set_dt(ctrl, 0.0, 0, stream));  // activates the stage kernels
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIPST(hipEventCreate(&e0));
        HIPST(hipEventCreate(&e1));
        float best = 0.f;
        int best_kind = 0;
        const bool log = getenv("DGMI_TUNE_LOG") != nullptr;
        Status st;
        for (int kind : {1, 2, 3}) {
            LaunchShape cand;
            if (!make_shape(kind, cand)) continue;
            shape = cand;
            if (!(st = stage(ST_RHS, ub, ub, ua))) break;
            if (hipEventRecord(e0, stream) != hipSuccess) break;
            for (int r = 0; r < 3 && st; r++) st = stage(ST_RHS, ub, ub, ua);
            if (!st) break;
            float ms = 0.f;
            if (hipEventRecord(e1, stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                st = Status::err(DG_ERR_HIP, "autotune timing failed");
                break;
            }
            if (log) fprintf(stderr, "dgmi autotune N=%d p=%d K=%d kind %d: %.4f ms per launch\n", N, p, K, kind, ms / 3);
            if (best_kind == 0 || ms < best) {
                best = ms;
                best_kind = kind;
            }
        }
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        if (!st) return st;
        if (best_kind == 0) return Status::err(DG_ERR_UNSUPPORTED, "no stage kernel supports this solve");
        lam_valid = false;  // set_dt cleared the accumulated wave speed
        {
            std::lock_guard<std::mutex> lk_(tune_mu());
            tune_cache()[key] = best_kind;
        }
        *out_kind = best_kind;
        return Status::ok();
    }

    Status choose_shape(int req) {
        kernel_req = req;
        int kind = req;
        if (kind == 0) {
            Status st = autotune(&kind);
            if (!st) return st;
        }
        LaunchShape s;
        Status st = make_shape(kind, s);
        if (!st) {
            if (req != 0) return st;
            // automatic choice falls back to the lane kernel
            if (!(st = make_shape(1, s))) return st;
        }
        const StateIx nix{N, NP, Kp, s.kind >= 2 ? tile_kind() : 0};
        if (nix.tiled != six.tiled) {
            // acc, ua and ub are written in full (padding included) by a
            // stage before any stage reads them; c is carried over
            if (state_set) {
                std::vector<double> h((size_t)N * NP * K);
                if (!(st = download_ref(c, h.data()))) return st;
                six = nix;
                if (!(st = upload_ref(h.data(), c))) return st;
            } else {
                six = nix;
                HIPST(hipMemsetAsync(c, 0, state_alloc() * sizeof(R), stream));
            }
        }
        shape = s;
        lam_valid = false;
        if (gexec) {
            hipGraphExecDestroy(gexec);
            gexec = nullptr;
        }
        return Status::ok();
    }

    Status push_ctrl(hipStream_t on = nullptr) {
        HIPST(hipMemcpyAsync(ctrl, &hctrl, sizeof(Ctrl), hipMemcpyHostToDevice, on ? on : stream));
        return Status::ok();
    }

    Status pull_ctrl() {
        HIPST(hipMemcpyAsync(&hctrl, ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, stream));
        HIPST(hipStreamSynchronize(stream));
        return Status::ok();
    }

    Status stage(int kind, const R* u, const R* cc, R* out, hipStream_t on = nullptr) {
        StageArgs<R> a{};
        a.u = u;
        a.c = cc;
        a.acc = acc;
        a.out = out;
        a.geo_e = geo_e;
        a.geo_f = geo_f;
        a.nbr = nbr;
        a.ctrl = ctrl;
        a.lam_bits = &ctrl->lam_bits;
        a.K = K;
        a.Kp = Kp;
        a.kind = kind;
        a.ntiles = shape.ntiles;
        a.xcd_remap = shape.xcd_remap;
        a.phys = phys;
        a.bnd = bnd;
        hipError_t e;
        hipStream_t sx = on ? on : stream;
        if (shape.kind >= 2)
            e = launch_stage_mfma<R>(N, p, a, mx, dim3(shape.grid), sx);
        else
            e = launch_stage_valu<R>(N, p, a, dim3(shape.grid), sx);
        return hip_status(e, "stage kernel launch");
    }

    Status stage_on(int kind, hipStream_t st) override {
        switch (kind) {
            case ST_RK4_1: return stage(ST_RK4_1, c, c, ua, st);
            case ST_RK4_2: return stage(ST_RK4_2, ua, c, ub, st);
            case ST_RK4_3: return stage(ST_RK4_3, ub, c, ua, st);
            case ST_RK4_4: return stage(ST_RK4_4, ua, c, c, st);
            default: return Status::err(DG_ERR_ARG, "stage_on: RK4 stages only");
        }
    }

    Status ensure_lambda(hipStream_t on = nullptr) {
        if (lam_valid) return Status::ok();
        HIPST(launch_lambda<R>(c, six, K, b0, phys, ctrl, on ? on : stream));
        lam_valid = true;
        return Status::ok();
    }

    // the stages of one step; the step's dt is already in ctrl
    Status enqueue_stages(int scheme) {
        Status st;
        if (scheme == DG_RK2) {
            if (!(st = stage(ST_RK2_1, c, c, ua))) return st;
            return stage(ST_RK2_2, ua, c, c);
        }
        if (!(st = stage(ST_RK4_1, c, c, ua))) return st;
        if (!(st = stage(ST_RK4_2, ua, c, ub))) return st;
        if (!(st = stage(ST_RK4_3, ub, c, ua))) return st;
        return stage(ST_RK4_4, ua, c, c);
    }

    Status enqueue_cfl_step(int scheme, bool use_graph) {
        if (use_graph) {
            if (!gexec || graph_scheme != scheme) {
                if (gexec) {
                    hipGraphExecDestroy(gexec);
                    gexec = nullptr;
                }
                hipGraph_t g = nullptr;
                HIPST(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                Status st = hip_status(launch_dt(ctrl, stream), "dt kernel");
                if (st) st = enqueue_stages(scheme);
                hipError_t ee = hipStreamEndCapture(stream, &g);
                if (!st) {
                    if (g) hipGraphDestroy(g);
                    return st;
                }
                HIPST(ee);
                hipError_t ei = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
                hipGraphDestroy(g);
                HIPST(ei);
                graph_scheme = scheme;
            }
            HIPST(hipGraphLaunch(gexec, stream));
            return Status::ok();
        }
        HIPST(launch_dt(ctrl, stream));
        return enqueue_stages(scheme);
    }

    Status begin_run(double t_end, bool has_end, double cfl, hipStream_t on = nullptr) {
        hctrl.t = t;
        hctrl.steps = steps;
        hctrl.t_end = t_end;
        hctrl.has_end = has_end ? 1 : 0;
        hctrl.done = 0;
        hctrl.active = 0;
        hctrl.cfl = cfl;
        hctrl.min_r = min_r;
        hctrl.two_p1 = 2. * p + 1.;
        Status st = push_ctrl(on);
        if (!st) return st;
        return ensure_lambda(on);
    }

    Status project_ic() override {
        HIPST(launch_project_ic<R>(p, c, six, V, K, ic, stream));
        HIPST(hipStreamSynchronize(stream));
        state_set = true;
        t = 0;
        steps = 0;
        lam_valid = false;
        return Status::ok();
    }

    // c_ref[var][mode][elem] (FP64, caller's element order) -> device, in
    // the current layout; padding elements replicate the last element and
    // padding modes are zero
    Status upload_ref(const double* c_ref, R* dst) {
        std::vector<R> h(six.elems(), R(0));
        for (int v = 0; v < N; v++)
            for (int m = 0; m < NP; m++) {
                const double* src = c_ref + ((size_t)v * NP + m) * K;
                for (int i = 0; i < Kp; i++) h[six.at(v, m, i)] = (R)src[perm[std::min(i, K - 1)]];
            }
        HIPST(hipMemcpyAsync(dst, h.data(), h.size() * sizeof(R), hipMemcpyHostToDevice, stream));
        HIPST(hipStreamSynchronize(stream));
        return Status::ok();
    }

    Status download_ref(const R* src, double* c_ref) {
        std::vector<R> h(six.elems());
        HIPST(hipMemcpyAsync(h.data(), src, h.size() * sizeof(R), hipMemcpyDeviceToHost, stream));
        HIPST(hipStreamSynchronize(stream));
        for (int v = 0; v < N; v++)
            for (int m = 0; m < NP; m++) {
                double* d = c_ref + ((size_t)v * NP + m) * K;
                for (int i = 0; i < K; i++) d[perm[i]] = (double)h[six.at(v, m, i)];
            }
        return Status::ok();
    }

    Status set_state(const double* c_ref) override {
        Status st = upload_ref(c_ref, c);
        state_set = (bool)st;
        t = 0;
        steps = 0;
        lam_valid = false;
        return st;
    }

    Status get_state(double* c_ref) override { return download_ref(c, c_ref); }

    Status rhs(const double* u_ref, double* k_ref, double dt) override {
        Status st = upload_ref(u_ref, ub);
        if (!st) return st;
        HIPST(launch_set_dt(ctrl, dt, 0, stream));
        lam_valid = false;  // set_dt cleared the accumulated wave speed
        if (!(st = stage(ST_RHS, ub, ub, ua))) return st;
        return download_ref(ua, k_ref);
    }

    // dg_rhs on the caller's device buffers through the device's scratch:
    // converts u to the state layout, one RHS stage launch, converts k back;
    // enqueued on `on` (0: the solver's stream), no allocation, no wait.
    Status rhs_device_noalloc(const double* u_dev, double* k_dev, double dt, hipStream_t on) override {
        int cur = -1;
        HIPST(hipGetDevice(&cur));
        if (cur != device) return Status::err(DG_ERR_STATE, "the solver's device is not the current device");
        char* base = static_cast<char*>(noalloc_scratch(device));
        if (!base) return Status::err(DG_ERR_STATE, "the device's scratch is not set up");
        const size_t one = (six.elems() * sizeof(R) + 255) / 256 * 256;
        if (2 * one > NOALLOC_SCRATCH_BYTES)
            return Status::err(DG_ERR_BUDGET, "this solve needs " + std::to_string(2 * one) +
                                                  " bytes of scratch; the allocation-free entry point has " +
                                                  std::to_string(NOALLOC_SCRATCH_BYTES));
        R* in = reinterpret_cast<R*>(base);
        R* out = reinterpret_cast<R*>(base + one);
        hipStream_t sx = on ? on : stream;
        HIPST(launch_ref_to_state<R>(u_dev, in, perm_d, six, K, sx));
        HIPST(launch_set_dt(ctrl, dt, 0, sx));
        lam_valid = false;  // set_dt cleared the accumulated wave speed
        Status st = stage(ST_RHS, in, in, out, sx);
        if (!st) return st;
        HIPST(launch_state_to_ref<R>(out, k_dev, perm_d, six, K, sx));
        return Status::ok();
    }

    Status rk_step(double dt, int scheme) override {
        if (scheme != DG_RK4 && scheme != DG_RK2) return Status::err(DG_ERR_ARG, "scheme must be DG_RK4 or DG_RK2");
        HIPST(launch_set_dt(ctrl, dt, 1, stream));
        Status st = enqueue_stages(scheme);
        if (!st) return st;
        lam_valid = true;  // the last stage left the new state's wave speeds
        HIPST(hipStreamSynchronize(stream));
        t += dt;
        steps += 1;
        return Status::ok();
    }

    Status cfl_dt(double cfl, double* dt) override {
        Status st = ensure_lambda();
        if (!st) return st;
        unsigned long long bits = 0;
        HIPST(hipMemcpyAsync(&bits, &ctrl->lam_bits, sizeof(bits), hipMemcpyDeviceToHost, stream));
        HIPST(hipStreamSynchronize(stream));
        double mx;
        std::memcpy(&mx, &bits, sizeof(mx));
        *dt = cfl * min_r / mx / (2. * p + 1.);
        return Status::ok();
    }

    Status integrate(double t_end, double cfl, int scheme, int64_t* nsteps) override {
        if (scheme != DG_RK4 && scheme != DG_RK2) return Status::err(DG_ERR_ARG, "scheme must be DG_RK4 or DG_RK2");
        if (!(cfl > 0)) return Status::err(DG_ERR_ARG, "cfl must be positive");
        const int64_t s0 = steps;
        Status st = begin_run(t_end, true, cfl);
        if (!st) return st;
        int chunk = 8;
        while (true) {
            for (int i = 0; i < chunk; i++)
                if (!(st = enqueue_cfl_step(scheme, true))) return st;
            if (!(st = pull_ctrl())) return st;
            if (hctrl.done || !(hctrl.t < t_end)) break;
            chunk = std::min(chunk * 2, 64);
        }
        t = hctrl.t;
        steps = hctrl.steps;
        if (nsteps) *nsteps = steps - s0;
        return Status::ok();
    }

    Status run_steps(int nsteps, double cfl, int scheme, bool use_graph) override {
        if (scheme != DG_RK4 && scheme != DG_RK2) return Status::err(DG_ERR_ARG, "scheme must be DG_RK4 or DG_RK2");
        if (!(cfl > 0)) return Status::err(DG_ERR_ARG, "cfl must be positive");
        Status st = begin_run(0.0, false, cfl);
        if (!st) return st;
        for (int i = 0; i < nsteps; i++)
            if (!(st = enqueue_cfl_step(scheme, use_graph))) return st;
        if (!(st = pull_ctrl())) return st;
        t = hctrl.t;
        steps = hctrl.steps;
        return Status::ok();
    }

    Status audit(dg_audit_out* out) override {
        HIPST(launch_audit<R>(c, six, Jd, K, b0, phys.gamma, audit_part, audit_blocks, stream));
        std::vector<double> h((size_t)audit_blocks * 8);
        HIPST(hipMemcpyAsync(h.data(), audit_part, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIPST(hipStreamSynchronize(stream));
        dg_audit_out o{};
        double rho = 1e300, pr = 1e300, bad = 0;
        for (int b = 0; b < audit_blocks; b++) {
            for (int v = 0; v < 4; v++) o.mass[v] += h[b * 8 + v];
            rho = std::min(rho, h[b * 8 + 4]);
            pr = std::min(pr, h[b * 8 + 5]);
            bad += h[b * 8 + 6];
        }
        o.min_density = rho;
        o.min_pressure = (N == 4) ? pr : INFINITY;
        o.n_bad = (int64_t)bad;
        o.t = t;
        o.steps = steps;
        *out = o;
        return Status::ok();
    }

    Status eval_vertices(int var, double* uv) override {
        if (var < 0 || var >= N) return Status::err(DG_ERR_ARG, "variable index out of range");
        std::vector<double> cr((size_t)N * NP * K);
        Status st = get_state(cr.data());
        if (!st) return st;
        Tables tb = build_tables(p);
        for (int e = 0; e < K; e++) {
            double a = 0, b = 0, d = 0;
            for (int i = 0; i < NP; i++) {
                const double ci = cr[((size_t)var * NP + i) * K + e];
                a += ci * tb.vertex[i * 3 + 0];
                b += ci * tb.vertex[i * 3 + 1];
                d += ci * tb.vertex[i * 3 + 2];
            }
            uv[3 * (size_t)e] = a;
            uv[3 * (size_t)e + 1] = b;
            uv[3 * (size_t)e + 2] = d;
        }
        return Status::ok();
    }

    Status time_step_graph(int nsteps, double cfl, double* out) override {
        Status st = begin_run(0.0, false, cfl);
        if (!st) return st;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIPST(hipEventCreate(&e0));
        HIPST(hipEventCreate(&e1));
        HIPST(hipEventRecord(e0, stream));
        for (int i = 0; i < nsteps && st; i++) st = enqueue_cfl_step(DG_RK4, true);
        float ms = 0.f;
        hipError_t e = hipEventRecord(e1, stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        if (!st) return st;
        HIPST(e);
        out[0] = ms / nsteps;
        out[1] = ms;
        if (!(st = pull_ctrl())) return st;
        t = hctrl.t;
        steps = hctrl.steps;
        lam_valid = true;
        return Status::ok();
    }

    Status time_stage_kernels(int nsteps, double cfl, double* out) override {
        Status st = begin_run(0.0, false, cfl);
        if (!st) return st;
        const int nst = 4 * nsteps;
        std::vector<hipEvent_t> ev(2 * nst + 2 * nsteps + 2);
        for (auto& e : ev) HIPST(hipEventCreate(&e));
        int k = 0;
        HIPST(hipEventRecord(ev[2 * nst + 2 * nsteps], stream));
        for (int s = 0; s < nsteps; s++) {
            HIPST(hipEventRecord(ev[2 * nst + 2 * s], stream));
            HIPST(launch_dt(ctrl, stream));
            HIPST(hipEventRecord(ev[2 * nst + 2 * s + 1], stream));
            const int kinds[4] = {ST_RK4_1, ST_RK4_2, ST_RK4_3, ST_RK4_4};
            const R* ins[4] = {c, ua, ub, ua};
            R* outs[4] = {ua, ub, ua, c};
            for (int j = 0; j < 4; j++) {
                HIPST(hipEventRecord(ev[2 * k], stream));
                if (!(st = stage(kinds[j], ins[j], c, outs[j]))) return st;
                HIPST(hipEventRecord(ev[2 * k + 1], stream));
                k++;
            }
        }
        HIPST(hipEventRecord(ev[2 * nst + 2 * nsteps + 1], stream));
        HIPST(hipStreamSynchronize(stream));
        double sum = 0, sumdt = 0;
        for (int i = 0; i < nst; i++) {
            float ms = 0;
            HIPST(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
            sum += ms;
        }
        for (int s = 0; s < nsteps; s++) {
            float ms = 0;
            HIPST(hipEventElapsedTime(&ms, ev[2 * nst + 2 * s], ev[2 * nst + 2 * s + 1]));
            sumdt += ms;
        }
        float tot = 0;
        HIPST(hipEventElapsedTime(&tot, ev[2 * nst + 2 * nsteps], ev[2 * nst + 2 * nsteps + 1]));
        for (auto& e : ev) hipEventDestroy(e);
        out[0] = sum / nst;
        out[1] = sumdt / nsteps;
        out[2] = tot;
        out[3] = nst;
        if (!(st = pull_ctrl())) return st;
        t = hctrl.t;
        steps = hctrl.steps;
        return Status::ok();
    }

    void slot_info(SlotInfo* o) override {
        o->c = c;
        o->acc = acc;
        o->ua = ua;
        o->ub = ub;
        o->geo_e = geo_e;
        o->geo_f = geo_f;
        o->nbr = nbr;
        o->fword = fword;
        o->halo = halo;
        o->face4 = face4;
        o->metric4 = metric4;
        o->ctrl = ctrl;
        o->K = K;
        o->Kp = Kp;
        o->N = N;
        o->p = p;
        o->hcap = mx.hcap;
        o->kind = shape.kind;
        o->boundary = has_boundary ? 1 : 0;
        o->phys = phys;
    }

    Status begin(double t_end, bool has_end, double cfl) override {
        Status st = begin_run(t_end, has_end, cfl);
        if (!st) return st;
        HIPST(hipStreamSynchronize(stream));
        return Status::ok();
    }

    Status begin_on(double t_end, bool has_end, double cfl, hipStream_t on) override {
        return begin_run(t_end, has_end, cfl, on);
    }

    Status refresh() override {
        Status st = pull_ctrl();
        if (!st) return st;
        t = hctrl.t;
        steps = hctrl.steps;
        return Status::ok();
    }

    void set_host_ctrl(const Ctrl& cc) override {
        hctrl = cc;
        t = hctrl.t;
        steps = hctrl.steps;
    }

    void set_lambda_valid(bool v) override { lam_valid = v; }
    hipStream_t get_stream() const override { return stream; }

    Status set_kernel(int kind) override {
        Status st = choose_shape(kind);
        if (!st) return st;
        return push_ctrl();
    }
    int kernel_kind() const override { return shape.kind; }

    Status set_max_groups(int n) override {
        if (n < 0 || (n > 0 && n < 8)) return Status::err(DG_ERR_ARG, "max_groups must be 0 or at least 8");
        max_groups = n;
        if (shape.kind != 2) return Status::ok();  // taken up when kind 2 is chosen
        const int req = kernel_req;
        Status st = choose_shape(2);
        kernel_req = req;
        if (!st) return st;
        return push_ctrl();
    }

    // Kind 2 walks tiles tile0, tile0 + stride, ... inside its XCD's
    // contiguous range of ceil(ntiles / 8) tiles (stride = grid / 8); the
    // other kinds launch one workgroup per tile.
    void launch_info(int* o) const override {
        o[0] = shape.kind;
        o[1] = shape.ntiles;
        o[2] = shape.grid;
        int walk = 1;
        if (shape.grid > 0) {
            if (shape.xcd_remap) {
                const int t8 = (shape.ntiles + 7) / 8, stride = std::max(1, shape.grid / 8);
                walk = (t8 + stride - 1) / stride;
            } else {
                walk = (shape.ntiles + shape.grid - 1) / shape.grid;
            }
        }
        o[3] = walk;
    }
    int64_t device_bytes() const override { return bytes + tile_bytes; }
    Status sync() override {
        HIPST(hipStreamSynchronize(stream));
        return Status::ok();
    }
};

Status Solver::create(const Mesh& m, const dg_problem_desc& prob, int p, int precision, Solver** out, int kernel) {
    const bool geom32 = (precision & DG_GEOM_FP32) != 0;
    precision &= ~DG_GEOM_FP32;
    Status st = check_problem(prob, p, precision);
    if (!st) return st;
    if (geom32 && precision != DG_FP64) return Status::err(DG_ERR_ARG, "DG_GEOM_FP32 applies to DG_FP64 solves");
    if (precision == DG_FP64) {
        auto* s = new SolverT<double>();
        s->geom_fp32 = geom32;
        st = s->init(m, prob, p, kernel);
        if (!st) {
            delete s;
            return st;
        }
        *out = s;
    } else {
        auto* s = new SolverT<float>();
        st = s->init(m, prob, p, kernel);
        if (!st) {
            delete s;
            return st;
        }
        *out = s;
    }
    return Status::ok();
}

}  // namespace dgmi
