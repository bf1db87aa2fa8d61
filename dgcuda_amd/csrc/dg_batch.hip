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

template <typename R>
__global__ void k_batch_dt(const SlotDesc<R>* __restrict__ slots, int nslots) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nslots) dt_update(slots[s].ctrl);
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

template <typename R>
class BatchT : public Batch {
   public:
    std::vector<Solver*> solvers;
    std::vector<double> t_end;
    hipStream_t stream = nullptr;
    SlotDesc<R>* d_slots = nullptr;
    int2* d_items = nullptr;
    int* d_xcd = nullptr;
    double* d_status = nullptr;
    int cap_slots = 0, cap_items = 0;
    int nitems = 0, grid = 0;
    size_t lds = 0;
    bool dirty = true;
    hipGraphExec_t gexec = nullptr;
    int64_t bytes = 0;

    ~BatchT() override {
        if (gexec) hipGraphExecDestroy(gexec);
        for (Solver* s : solvers) delete s;
        void* b[] = {d_slots, d_items, d_xcd, d_status};
        for (void* p : b)
            if (p) hipFree(p);
        if (stream) hipStreamDestroy(stream);
    }

    Status init() {
        Status st = ensure_device_ready();
        if (!st) return st;
        HIPST(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        return Status::ok();
    }

    Status add(const Mesh& m, const dg_problem_desc& prob, int p, double te, int* slot) override {
        if (p < 1 || p > 4) return Status::err(DG_ERR_UNSUPPORTED, "batched solves support p = 1..4");
        if (!(te > 0)) return Status::err(DG_ERR_ARG, "end time must be positive");
        Solver* s = nullptr;
        Status st = Solver::create(m, prob, p, sizeof(R) == 8 ? DG_FP64 : DG_FP32, &s);
        if (!st) return st;
        if (sizeof(R) == 4) {
            st = s->set_kernel(2);
            if (!st) {
                delete s;
                return st;
            }
        }
        solvers.push_back(s);
        t_end.push_back(te);
        dirty = true;
        if (slot) *slot = (int)solvers.size() - 1;
        return Status::ok();
    }

    // slot table, work list, XCD partition and launch shape
    Status rebuild() {
        const int ns = (int)solvers.size();
        if (ns == 0) return Status::err(DG_ERR_STATE, "empty batch");
        std::vector<SlotDesc<R>> hs(ns);
        std::vector<int2> hi;
        std::vector<double> cost;
        size_t max_lds = 0;
        for (int s = 0; s < ns; s++) {
            SlotInfo si;
            solvers[s]->slot_info(&si);
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
            if (d.path == 2) max_lds = std::max(max_lds, mfma_tile_lds_bytes(si.N, si.hcap));
            const double per_elem = (double)si.N * n_modes(si.p) * sizeof(R);
            for (int t = 0; t < nt; t++) {
                hi.push_back(int2{s, t});
                cost.push_back(per_elem * std::min(te, si.K - t * te));
            }
        }
        nitems = (int)hi.size();
        // 8 contiguous chunks of equal estimated cost (one per XCD)
        std::vector<int> xs(9, nitems);
        xs[0] = 0;
        const double total = std::accumulate(cost.begin(), cost.end(), 0.0);
        double run = 0;
        int x = 1;
        for (int i = 0; i < nitems && x < 8; i++) {
            run += cost[i];
            while (x < 8 && run >= total * x / 8.0) xs[x++] = i + 1;
        }
        for (; x < 8; x++) xs[x] = nitems;
        if (ns > cap_slots) {
            if (d_slots) hipFree(d_slots);
            if (d_status) hipFree(d_status);
            cap_slots = std::max(ns, 2 * cap_slots);
            HIPST(hipMalloc((void**)&d_slots, cap_slots * sizeof(SlotDesc<R>)));
            HIPST(hipMalloc((void**)&d_status, cap_slots * 3 * sizeof(double)));
        }
        if (nitems > cap_items) {
            if (d_items) hipFree(d_items);
            cap_items = std::max(nitems, 2 * cap_items);
            HIPST(hipMalloc((void**)&d_items, cap_items * sizeof(int2)));
        }
        if (!d_xcd) HIPST(hipMalloc((void**)&d_xcd, 9 * sizeof(int)));
        HIPST(hipMemcpyAsync(d_slots, hs.data(), ns * sizeof(SlotDesc<R>), hipMemcpyHostToDevice, stream));
        HIPST(hipMemcpyAsync(d_items, hi.data(), nitems * sizeof(int2), hipMemcpyHostToDevice, stream));
        HIPST(hipMemcpyAsync(d_xcd, xs.data(), 9 * sizeof(int), hipMemcpyHostToDevice, stream));
        HIPST(hipStreamSynchronize(stream));
        lds = max_lds;
        // persistent grid: the workgroups that fit at once, a multiple of 8
        int per_cu = 0, dev = 0, cus = 0;
        HIPST(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(&k_batch_stage<R, ST_RK4_2>), 256, lds));
        HIPST(hipGetDevice(&dev));
        HIPST(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        grid = std::max(8, std::min(per_cu * cus, (nitems + 7) / 8 * 8) / 8 * 8);
        if (gexec) {
            hipGraphExecDestroy(gexec);
            gexec = nullptr;
        }
        dirty = false;
        return Status::ok();
    }

    Status stage(int kind) {
        const dim3 g(grid), b(256);
        switch (kind) {
            case ST_RK4_1: hipLaunchKernelGGL((k_batch_stage<R, ST_RK4_1>), g, b, lds, stream, d_slots, d_items, d_xcd); break;
            case ST_RK4_2: hipLaunchKernelGGL((k_batch_stage<R, ST_RK4_2>), g, b, lds, stream, d_slots, d_items, d_xcd); break;
            case ST_RK4_3: hipLaunchKernelGGL((k_batch_stage<R, ST_RK4_3>), g, b, lds, stream, d_slots, d_items, d_xcd); break;
            case ST_RK4_4: hipLaunchKernelGGL((k_batch_stage<R, ST_RK4_4>), g, b, lds, stream, d_slots, d_items, d_xcd); break;
            default: return Status::err(DG_ERR_ARG, "unsupported batch stage");
        }
        return hip_status(hipGetLastError(), "batch stage launch");
    }

    Status enqueue_step_direct() {
        const int ns = (int)solvers.size();
        hipLaunchKernelGGL((k_batch_dt<R>), dim3((ns + 63) / 64), dim3(64), 0, stream, d_slots, ns);
        HIPST(hipGetLastError());
        Status st;
        for (int k : {ST_RK4_1, ST_RK4_2, ST_RK4_3, ST_RK4_4})
            if (!(st = stage(k))) return st;
        return Status::ok();
    }

    Status enqueue_step(bool use_graph) {
        if (!use_graph) return enqueue_step_direct();
        if (!gexec) {
            hipGraph_t gr = nullptr;
            HIPST(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            Status st = enqueue_step_direct();
            hipError_t ee = hipStreamEndCapture(stream, &gr);
            if (!st) {
                if (gr) hipGraphDestroy(gr);
                return st;
            }
            HIPST(ee);
            hipError_t ei = hipGraphInstantiate(&gexec, gr, nullptr, nullptr, 0);
            hipGraphDestroy(gr);
            HIPST(ei);
        }
        HIPST(hipGraphLaunch(gexec, stream));
        return Status::ok();
    }

    Status begin_all(bool has_end, double cfl) {
        Status st;
        if (dirty && !(st = rebuild())) return st;
        for (size_t s = 0; s < solvers.size(); s++)
            if (!(st = solvers[s]->begin(t_end[s], has_end, cfl))) return st;
        return Status::ok();
    }

    Status finish_all() {
        HIPST(hipStreamSynchronize(stream));
        Status st;
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
            if (!(st = enqueue_step(use_graph))) return st;
        return finish_all();
    }

    Status integrate(double cfl, bool use_graph, int64_t* steps) override {
        if (!(cfl > 0)) return Status::err(DG_ERR_ARG, "cfl must be positive");
        Status st = begin_all(true, cfl);
        if (!st) return st;
        const int ns = (int)solvers.size();
        std::vector<double> h(3 * ns);
        int64_t issued = 0;
        int chunk = 8;
        while (true) {
            for (int i = 0; i < chunk; i++)
                if (!(st = enqueue_step(use_graph))) return st;
            issued += chunk;
            hipLaunchKernelGGL((k_batch_status<R>), dim3((ns + 63) / 64), dim3(64), 0, stream, d_slots, ns, d_status);
            HIPST(hipGetLastError());
            HIPST(hipMemcpyAsync(h.data(), d_status, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
            HIPST(hipStreamSynchronize(stream));
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
        HIPST(hipStreamSynchronize(stream));
        return solvers[slot]->get_state(c_ref);
    }

    Status audit(int slot, dg_audit_out* out) override {
        if (slot < 0 || slot >= (int)solvers.size()) return Status::err(DG_ERR_ARG, "slot out of range");
        HIPST(hipStreamSynchronize(stream));
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
    Status st;
    if (precision == DG_FP32) {
        auto* b = new BatchT<float>();
        if (!(st = b->init())) {
            delete b;
            return st;
        }
        *out = b;
    } else {
        auto* b = new BatchT<double>();
        if (!(st = b->init())) {
            delete b;
            return st;
        }
        *out = b;
    }
    return Status::ok();
}

}  // namespace dgmi
