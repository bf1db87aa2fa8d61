// Many independent solves of uneven size and mixed order advanced together:
// one persistent launch per RK stage over a device-side work list of
// (solve, tile) items, the step (per-solve dt kernel + 4 stage launches)
// replayed from a hipGraph.
//
// Every solve keeps its own buffers, geometry and control block (its own
// CFL dt, time and end time), so a solve's arithmetic does not depend on
// which other solves share the launch: its result is the same alone and in
// any mix.  In FP32 every item runs the matrix-core tile body (64-element
// tiles; the products are padded to 16x16 at low order), in FP64 the
// one-element-per-lane body (256-element tiles).  The work list is split into 8
// contiguous chunks of equal estimated cost, one per XCD, and the
// workgroups of an XCD take the items of its chunk round-robin.
#include <algorithm>
#include <cstring>
#include <numeric>

#include "dg_batch.h"
#include "dg_mfma_body.h"
#include "dg_valu_body.h"

namespace dgmi {

DGMI_TABLE_UPLOADER(upload_tables_batch)

template <typename R>
struct SlotDesc {
    R *c, *acc, *ua, *ub;
    const R *geo_e, *geo_f;
    const int* nbr;
    const uint16_t* fword;
    const int* halo;
    Ctrl* ctrl;
    int K, Kp, N, p, hcap, path;  // path 1 = lane per element, 2 = matrix-core tiles
    PhysP phys;
};

constexpr int BATCH_VALU_TILE = 256;

template <int N, int P, typename R, int KIND>
__device__ __forceinline__ double batch_valu(const StageArgs<R>& a, int e) {
    return valu_process<N, P, R>(a, e, KIND);
}

template <int N, int P, int KIND>
__device__ __forceinline__ double batch_mfma(const StageArgs<float>& a, const MfmaArgs& mx, int tile, float4* fd,
                                             float* ts) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, j = lane & 15;
    MfmaA<N, P> A;
    A.init(g, j);
    MfmaTileState<N, P> st;
    st.load(a.u, a.K, a.Kp, mx.halo, mx.hcap, tile, w, g, j);
    return mfma_process_tile<N, P, KIND>(a, mx, A, st, tile, fd, ts);
}

template <typename R, int KIND>
__global__ void __launch_bounds__(256, 2) k_batch_stage(const SlotDesc<R>* __restrict__ slots,
                                                    const int2* __restrict__ items, const int* __restrict__ xcd_start) {
    constexpr bool LAST = (KIND == ST_RK4_4 || KIND == ST_RK2_2);
    extern __shared__ float4 lds4[];
    __shared__ double sh[8];
    const int x = blockIdx.x & 7, G8 = gridDim.x >> 3;
    const int i1 = xcd_start[x + 1];
    for (int it = xcd_start[x] + (blockIdx.x >> 3); it < i1; it += G8) {
        const int2 item = items[it];
        const SlotDesc<R>* sd = slots + item.x;
        Ctrl* ctrl = sd->ctrl;
        if (!ctrl->active) continue;
        StageArgs<R> a;
        a.ctrl = ctrl;
        a.lam_bits = &ctrl->lam_bits;
        a.geo_e = sd->geo_e;
        a.geo_f = sd->geo_f;
        a.nbr = sd->nbr;
        a.K = sd->K;
        a.Kp = sd->Kp;
        a.kind = KIND;
        a.ntiles = 0;
        a.xcd_remap = 0;
        a.pad = 0;
        a.phys = sd->phys;
        a.acc = sd->acc;
        a.c = sd->c;
        if constexpr (KIND == ST_RK4_1 || KIND == ST_RK2_1) {
            a.u = sd->c;
            a.out = sd->ua;
        } else if constexpr (KIND == ST_RK4_2) {
            a.u = sd->ua;
            a.out = sd->ub;
        } else if constexpr (KIND == ST_RK4_3) {
            a.u = sd->ub;
            a.out = sd->ua;
        } else if constexpr (KIND == ST_RK4_4 || KIND == ST_RK2_2) {
            a.u = sd->ua;
            a.out = sd->c;
        } else {
            a.u = sd->ub;
            a.out = sd->ua;
        }
        const int N = sd->N, p = sd->p;
        double lam = 0.0;
        if constexpr (sizeof(R) == 4) {
            // FP32: every order on the matrix-core tile body
            MfmaArgs mx;
            mx.fword = sd->fword;
            mx.halo = sd->halo;
            mx.hcap = sd->hcap;
            mx.pad = 0;
            float4* fd = lds4;
            float* ts = reinterpret_cast<float*>(lds4 + MFMA_TILE * 3);
            if (N == 4 && p == 4) lam = batch_mfma<4, 4, KIND>(a, mx, item.y, fd, ts);
            else if (N == 4 && p == 3) lam = batch_mfma<4, 3, KIND>(a, mx, item.y, fd, ts);
            else if (N == 4 && p == 2) lam = batch_mfma<4, 2, KIND>(a, mx, item.y, fd, ts);
            else if (N == 4 && p == 1) lam = batch_mfma<4, 1, KIND>(a, mx, item.y, fd, ts);
            else if (N == 1 && p == 4) lam = batch_mfma<1, 4, KIND>(a, mx, item.y, fd, ts);
            else if (N == 1 && p == 3) lam = batch_mfma<1, 3, KIND>(a, mx, item.y, fd, ts);
            else if (N == 1 && p == 2) lam = batch_mfma<1, 2, KIND>(a, mx, item.y, fd, ts);
            else if (N == 1 && p == 1) lam = batch_mfma<1, 1, KIND>(a, mx, item.y, fd, ts);
        } else {
            const int e = item.y * BATCH_VALU_TILE + threadIdx.x;
            if (N == 4 && p == 4) lam = batch_valu<4, 4, double, KIND>(a, e);
            else if (N == 4 && p == 3) lam = batch_valu<4, 3, double, KIND>(a, e);
            else if (N == 4 && p == 2) lam = batch_valu<4, 2, double, KIND>(a, e);
            else if (N == 4 && p == 1) lam = batch_valu<4, 1, double, KIND>(a, e);
            else if (N == 1 && p == 4) lam = batch_valu<1, 4, double, KIND>(a, e);
            else if (N == 1 && p == 3) lam = batch_valu<1, 3, double, KIND>(a, e);
            else if (N == 1 && p == 2) lam = batch_valu<1, 2, double, KIND>(a, e);
            else if (N == 1 && p == 1) lam = batch_valu<1, 1, double, KIND>(a, e);
        }
        if constexpr (LAST) {
            const double m = block_max(lam, sh);
            if (threadIdx.x == 0) lam_commit(m, a.lam_bits);
            __syncthreads();  // sh is reused by the next item
        }
    }
}

// slot count on the device, so a captured step stays valid as solves come and go
template <typename R>
__global__ void k_batch_dt(const SlotDesc<R>* __restrict__ slots, const int* __restrict__ nslots) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < *nslots) dt_update(slots[s].ctrl);
}

// done flag, steps and t of every slot into one array (one host copy)
template <typename R>
__global__ void k_batch_status(const SlotDesc<R>* __restrict__ slots, int nslots, double* out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nslots) {
        const Ctrl* c = slots[s].ctrl;
        out[3 * s] = c->done ? 1.0 : 0.0;
        out[3 * s + 1] = (double)c->steps;
        out[3 * s + 2] = c->t;
    }
}

#define HIPST(expr)                                          \
    do {                                                     \
        hipError_t _e = (expr);                              \
        if (_e != hipSuccess) return hip_status(_e, #expr);  \
    } while (0)

// Device side of a batch: slot table, work list, XCD partition, the
// persistent launch shape and the captured step.  The grid and LDS size are
// fixed at creation (LDS for the largest tile halo), and the slot count
// lives on the device, so the step graph stays valid while set_slots
// changes which solves are resident, as long as the capacities hold.
template <typename R>
class EngineT : public BatchEngine {
   public:
    hipStream_t st = nullptr;
    SlotDesc<R>* d_slots = nullptr;
    int2* d_items = nullptr;
    int* d_xcd = nullptr;   // [9] XCD chunk bounds, then [1] slot count
    double* d_status = nullptr;
    int cap_slots = 0, cap_items = 0;
    int nslots = 0, nitems = 0, grid = 0, dt_grid = 0;
    size_t lds = 0;
    hipGraphExec_t gexec = nullptr;

    ~EngineT() override {
        if (gexec) hipGraphExecDestroy(gexec);
        void* b[] = {d_slots, d_items, d_xcd, d_status};
        for (void* q : b)
            if (q) hipFree(q);
        if (st) hipStreamDestroy(st);
    }

    Status init(int max_slots, int max_items) {
        Status s0 = ensure_device_ready();
        if (!s0) return s0;
        HIPST(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIPST(hipMalloc((void**)&d_xcd, 10 * sizeof(int)));
        HIPST(hipMemsetAsync(d_xcd, 0, 10 * sizeof(int), st));
        lds = (sizeof(R) == 4) ? mfma_tile_lds_bytes(4, MFMA_MAX_HALO) : 0;
        int per_cu = 0, dev = 0, cus = 0;
        HIPST(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(&k_batch_stage<R, ST_RK4_2>), 256, lds));
        HIPST(hipGetDevice(&dev));
        HIPST(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        grid = std::max(8, per_cu * cus / 8 * 8);
        Status s1 = reserve(max_slots, max_items);
        if (!s1) return s1;
        HIPST(hipStreamSynchronize(st));
        return Status::ok();
    }

    // grows the device tables; invalidates the captured step when it does
    Status reserve(int ns, int ni) {
        if (ns > cap_slots) {
            if (d_slots) hipFree(d_slots);
            if (d_status) hipFree(d_status);
            cap_slots = std::max(ns, 2 * cap_slots);
            HIPST(hipMalloc((void**)&d_slots, cap_slots * sizeof(SlotDesc<R>)));
            HIPST(hipMalloc((void**)&d_status, cap_slots * 3 * sizeof(double)));
            dt_grid = (cap_slots + 63) / 64;
            drop_graph();
        }
        if (ni > cap_items) {
            if (d_items) hipFree(d_items);
            cap_items = std::max(ni, 2 * cap_items);
            HIPST(hipMalloc((void**)&d_items, cap_items * sizeof(int2)));
            drop_graph();
        }
        return Status::ok();
    }

    void drop_graph() {
        if (gexec) {
            hipGraphExecDestroy(gexec);
            gexec = nullptr;
        }
    }

    Status set_slots(const std::vector<Solver*>& solvers) override {
        const int ns = (int)solvers.size();
        std::vector<SlotDesc<R>> hs(ns);
        std::vector<int2> hi;
        std::vector<double> cost;
        for (int s = 0; s < ns; s++) {
            SlotInfo si;
            solvers[s]->slot_info(&si);
            if (sizeof(R) == 4 && si.kind < 2) return Status::err(DG_ERR_STATE, "FP32 batch slot without matrix-core tiling");
            if (si.hcap > MFMA_MAX_HALO) return Status::err(DG_ERR_UNSUPPORTED, "tile halo too large");
            SlotDesc<R>& d = hs[s];
            d.c = (R*)si.c;
            d.acc = (R*)si.acc;
            d.ua = (R*)si.ua;
            d.ub = (R*)si.ub;
            d.geo_e = (const R*)si.geo_e;
            d.geo_f = (const R*)si.geo_f;
            d.nbr = si.nbr;
            d.fword = si.fword;
            d.halo = si.halo;
            d.ctrl = si.ctrl;
            d.K = si.K;
            d.Kp = si.Kp;
            d.N = si.N;
            d.p = si.p;
            d.hcap = si.hcap;
            d.path = (sizeof(R) == 4) ? 2 : 1;
            d.phys = si.phys;
            const int te = (d.path == 2) ? MFMA_TILE : BATCH_VALU_TILE;
            const int nt = (si.K + te - 1) / te;
            const double per_elem = (double)si.N * n_modes(si.p) * sizeof(R);
            for (int t = 0; t < nt; t++) {
                hi.push_back(int2{s, t});
                cost.push_back(per_elem * std::min(te, si.K - t * te));
            }
        }
        const int ni = (int)hi.size();
        // 8 contiguous chunks of equal estimated cost (one per XCD), then the slot count
        std::vector<int> xs(10, ni);
        xs[0] = 0;
        const double total = std::accumulate(cost.begin(), cost.end(), 0.0);
        double run = 0;
        int x = 1;
        for (int i = 0; i < ni && x < 8; i++) {
            run += cost[i];
            while (x < 8 && run >= total * x / 8.0) xs[x++] = i + 1;
        }
        xs[9] = ns;
        Status s0 = reserve(std::max(ns, 1), std::max(ni, 1));
        if (!s0) return s0;
        if (ns) HIPST(hipMemcpyAsync(d_slots, hs.data(), ns * sizeof(SlotDesc<R>), hipMemcpyHostToDevice, st));
        if (ni) HIPST(hipMemcpyAsync(d_items, hi.data(), ni * sizeof(int2), hipMemcpyHostToDevice, st));
        HIPST(hipMemcpyAsync(d_xcd, xs.data(), 10 * sizeof(int), hipMemcpyHostToDevice, st));
        HIPST(hipStreamSynchronize(st));  // the host vectors go out of scope
        nslots = ns;
        nitems = ni;
        return Status::ok();
    }

    Status launch_stage(int kind) {
        const dim3 g(grid), b(256);
        switch (kind) {
            case ST_RK4_1: hipLaunchKernelGGL((k_batch_stage<R, ST_RK4_1>), g, b, lds, st, d_slots, d_items, d_xcd); break;
            case ST_RK4_2: hipLaunchKernelGGL((k_batch_stage<R, ST_RK4_2>), g, b, lds, st, d_slots, d_items, d_xcd); break;
            case ST_RK4_3: hipLaunchKernelGGL((k_batch_stage<R, ST_RK4_3>), g, b, lds, st, d_slots, d_items, d_xcd); break;
            case ST_RK4_4: hipLaunchKernelGGL((k_batch_stage<R, ST_RK4_4>), g, b, lds, st, d_slots, d_items, d_xcd); break;
            default: return Status::err(DG_ERR_ARG, "unsupported batch stage");
        }
        return hip_status(hipGetLastError(), "batch stage launch");
    }

    Status step_direct() {
        hipLaunchKernelGGL((k_batch_dt<R>), dim3(dt_grid), dim3(64), 0, st, d_slots, d_xcd + 9);
        HIPST(hipGetLastError());
        Status s0;
        for (int k : {ST_RK4_1, ST_RK4_2, ST_RK4_3, ST_RK4_4})
            if (!(s0 = launch_stage(k))) return s0;
        return Status::ok();
    }

    Status step(bool use_graph) override {
        if (!use_graph) return step_direct();
        if (!gexec) {
            hipGraph_t gr = nullptr;
            HIPST(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            Status s0 = step_direct();
            hipError_t ee = hipStreamEndCapture(st, &gr);
            if (!s0) {
                if (gr) hipGraphDestroy(gr);
                return s0;
            }
            HIPST(ee);
            hipError_t ei = hipGraphInstantiate(&gexec, gr, nullptr, nullptr, 0);
            hipGraphDestroy(gr);
            HIPST(ei);
        }
        HIPST(hipGraphLaunch(gexec, st));
        return Status::ok();
    }

    Status status(std::vector<double>& out) override {
        out.assign(3 * (size_t)nslots, 0.0);
        if (nslots == 0) return Status::ok();
        hipLaunchKernelGGL((k_batch_status<R>), dim3((nslots + 63) / 64), dim3(64), 0, st, d_slots, nslots, d_status);
        HIPST(hipGetLastError());
        HIPST(hipMemcpyAsync(out.data(), d_status, out.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPST(hipStreamSynchronize(st));
        return Status::ok();
    }

    Status sync() override {
        HIPST(hipStreamSynchronize(st));
        return Status::ok();
    }
    hipStream_t stream() override { return st; }
    int max_tile_halo() const override { return MFMA_MAX_HALO; }
};

Status BatchEngine::create(int precision, int max_slots, int max_items, BatchEngine** out) {
    Status s0;
    if (precision == DG_FP32) {
        auto* e = new EngineT<float>();
        if (!(s0 = e->init(max_slots, max_items))) {
            delete e;
            return s0;
        }
        *out = e;
    } else if (precision == DG_FP64) {
        auto* e = new EngineT<double>();
        if (!(s0 = e->init(max_slots, max_items))) {
            delete e;
            return s0;
        }
        *out = e;
    } else {
        return Status::err(DG_ERR_ARG, "precision must be DG_FP32 or DG_FP64");
    }
    return Status::ok();
}

Status make_batch_solver(const Mesh& m, const dg_problem_desc& prob, int p, int precision, Solver** out) {
    if (p < 1 || p > 4) return Status::err(DG_ERR_UNSUPPORTED, "batched solves support p = 1..4");
    Solver* s = nullptr;
    Status st = Solver::create(m, prob, p, precision, &s);
    if (!st) return st;
    if (precision == DG_FP32 && !(st = s->set_kernel(2))) {
        delete s;
        return st;
    }
    *out = s;
    return Status::ok();
}

// The user-facing batch: solves added up front, run together.
class BatchImpl : public Batch {
   public:
    int precision = DG_FP32;
    std::vector<Solver*> solvers;
    std::vector<double> t_end;
    BatchEngine* eng = nullptr;
    bool dirty = true;

    ~BatchImpl() override {
        delete eng;
        for (Solver* s : solvers) delete s;
    }

    Status add(const Mesh& m, const dg_problem_desc& prob, int p, double te, int* slot) override {
        if (!(te > 0)) return Status::err(DG_ERR_ARG, "end time must be positive");
        Solver* s = nullptr;
        Status st = make_batch_solver(m, prob, p, precision, &s);
        if (!st) return st;
        solvers.push_back(s);
        t_end.push_back(te);
        dirty = true;
        if (slot) *slot = (int)solvers.size() - 1;
        return Status::ok();
    }

    Status begin_all(bool has_end, double cfl) {
        if (solvers.empty()) return Status::err(DG_ERR_STATE, "empty batch");
        Status st;
        if (dirty) {
            if (!(st = eng->set_slots(solvers))) return st;
            dirty = false;
        }
        for (size_t s = 0; s < solvers.size(); s++)
            if (!(st = solvers[s]->begin(t_end[s], has_end, cfl))) return st;
        return Status::ok();
    }

    Status finish_all() {
        Status st = eng->sync();
        if (!st) return st;
        for (Solver* s : solvers) {
            if (!(st = s->refresh())) return st;
            s->set_lambda_valid(true);
        }
        return Status::ok();
    }

    Status project_ic() override {
        Status st;
        for (Solver* s : solvers)
            if (!(st = s->project_ic())) return st;
        return Status::ok();
    }

    Status run_steps(int nsteps, double cfl, bool use_graph) override {
        if (!(cfl > 0) || nsteps < 0) return Status::err(DG_ERR_ARG, "invalid arguments");
        Status st = begin_all(false, cfl);
        if (!st) return st;
        for (int i = 0; i < nsteps; i++)
            if (!(st = eng->step(use_graph))) return st;
        return finish_all();
    }

    Status integrate(double cfl, bool use_graph, int64_t* steps) override {
        if (!(cfl > 0)) return Status::err(DG_ERR_ARG, "cfl must be positive");
        Status st = begin_all(true, cfl);
        if (!st) return st;
        const int ns = (int)solvers.size();
        std::vector<double> h;
        int64_t issued = 0;
        int chunk = 8;
        while (true) {
            for (int i = 0; i < chunk; i++)
                if (!(st = eng->step(use_graph))) return st;
            issued += chunk;
            if (!(st = eng->status(h))) return st;
            bool all = true;
            for (int s = 0; s < ns; s++) all = all && (h[3 * s] != 0.0 || !(h[3 * s + 2] < t_end[s]));
            if (all) break;
            chunk = std::min(chunk * 2, 64);
        }
        if (steps) *steps = issued;
        return finish_all();
    }

    Status get_state(int slot, double* c_ref) override {
        if (slot < 0 || slot >= (int)solvers.size()) return Status::err(DG_ERR_ARG, "slot out of range");
        Status st = eng->sync();
        if (!st) return st;
        return solvers[slot]->get_state(c_ref);
    }

    Status audit(int slot, dg_audit_out* out) override {
        if (slot < 0 || slot >= (int)solvers.size()) return Status::err(DG_ERR_ARG, "slot out of range");
        Status st = eng->sync();
        if (!st) return st;
        return solvers[slot]->audit(out);
    }

    void info(int* n, int64_t* elems, int64_t* dbytes) const override {
        int64_t e = 0, b = 0;
        for (Solver* s : solvers) {
            e += s->K;
            b += s->device_bytes();
        }
        if (n) *n = (int)solvers.size();
        if (elems) *elems = e;
        if (dbytes) *dbytes = b;
    }
};

Status Batch::create(int precision, Batch** out) {
    if (precision != DG_FP32 && precision != DG_FP64) return Status::err(DG_ERR_ARG, "precision must be DG_FP32 or DG_FP64");
    auto* b = new BatchImpl();
    b->precision = precision;
    Status st = BatchEngine::create(precision, 64, 4096, &b->eng);
    if (!st) {
        delete b;
        return st;
    }
    *out = b;
    return Status::ok();
}

}  // namespace dgmi
