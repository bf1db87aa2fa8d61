// Many independent solves of uneven size and mixed order advanced together:
// persistent launches per RK stage over device-side work lists of (solve,
// tile) items, the step (per-solve dt kernel + the stage launches) replayed
// from a hipGraph.  Two launch forms: one kernel per stage carrying every
// class (the FP32 default), or one launch per stage for each kernel class
// present (Euler tiles, advection tiles, Euler lane items, advection lane
// items), each at the occupancy its own bodies allow (the FP64 default: the
// all-class FP64 kernel holds one wavefront per SIMD).
//
// Every solve keeps its own buffers, geometry and control block (its own
// CFL dt, time and end time), so a solve's arithmetic does not depend on
// which other solves share the launch: its result is the same alone and in
// any mix.  Solves of p >= 3 run as 64-element tiles through the pipelined
// matrix-core body of the single-solve kernel (FP32 or FP64), solves of
// p <= 2 as 256-element lane items of the one-element-per-lane body, the
// single solve's kernel choice in both cases.  Each work list is split into
// 8 contiguous chunks of equal estimated cost, one per XCD, and the
// workgroups of an XCD take the items of its chunk round-robin.
#include <algorithm>
#include <cstring>
#include <numeric>

#include "dg_batch.h"
#include "dg_worklist.h"
#include "dg_mfma_body.h"
#include "dg_valu_body.h"

namespace dgmi {

DGMI_TABLE_UPLOADER(upload_tables_batch)

template <typename R>
struct SlotDesc {
    R *c, *acc, *ua, *ub;
    const R *geo_e, *geo_f;
    const int* nbr;
    const uint32_t* fword;
    const int* halo;
    const void* face4;
    const void* metric4;
    Ctrl* ctrl;
    int K, Kp, N, p, hcap, path;  // path 1 = lane per element, 2 = matrix-core tiles
    PhysP phys;
};

constexpr int BATCH_VALU_TILE = 256;

// Values read from the slot table or the work list are the same in every
// lane, but those reads follow the kernel's own stores, so the compiler
// cannot issue them as scalar loads and would keep pointers, sizes and
// addresses in vector registers.  readfirstlane puts them in SGPRs.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ double uni(double x) {
    const unsigned long long v = __double_as_longlong(x);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <typename T>
__device__ __forceinline__ T* uni(T* ptr) {
    const unsigned long long v = reinterpret_cast<unsigned long long>(ptr);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return reinterpret_cast<T*>(((unsigned long long)hi << 32) | lo);
}

// Tile of item `x` if it belongs to `slot`, else `fallback` (a tile of the
// same solve, so loads issued for it stay in bounds).
__device__ __forceinline__ int run_tile(const int2* __restrict__ items, int x, int i1, int slot, int fallback) {
    if (x >= i1) return fallback;
    const int2 v = items[x];
    return uni(v.x) == slot ? uni(v.y) : fallback;
}

// One workgroup's run of consecutive items of the same solve (items it,
// it+G8, ... while they belong to `slot`) through the pipelined tile body of
// the single-solve kernel (the same arithmetic, so a batched solve equals
// the solve run alone bit for bit): the state and geometry of the next tile
// and the halo ids of the one after are in flight while a tile computes.
// Returns the first item index not processed.
template <int N, int P, int KIND, typename R>
__device__ __forceinline__ int batch_mfma_run(const StageArgs<R>& a, const MfmaArgs& mx, int slot,
                                              const int2* __restrict__ items, int it, int i1, int G8,
                                              typename MfmaT<R>::V* fd, R* ts, typename MfmaT<R>::V* tabA,
                                              double& lam) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, j = lane & 15;
    const R dt = (R)uni(a.ctrl->dt);
    int tile = uni(items[it].y);
    int n1 = run_tile(items, it + G8, i1, slot, tile);
    MfmaTileState<N, P, R> st;
    int hid[MfmaTileState<N, P, R>::HPW];
    TileGeo<R> gc;
    MfmaTileState<N, P, R>::load_ids(mx.halo, mx.hcap, tile, w, j, hid);
    gc.load(a, mx, tile, w * 16 + j);
    st.load_data(a.u, a.K, tile, w, g, j, hid);
    MfmaTileState<N, P, R>::load_ids(mx.halo, mx.hcap, n1, w, j, hid);
    if (w == 0) mfma_a_to_lds<N, P, R>(tabA, lane);  // the previous run ended with a barrier
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    while (true) {
        const int nit = it + G8;
        const bool more = nit < i1 && uni(items[nit].x) == slot;
        const int n2 = run_tile(items, it + 2 * G8, i1, slot, n1);
        lam = fmax(lam, mfma_process_tile_pipe<N, P, KIND, R>(a, mx, tabA, st, hid, gc, dt, tile, n1, n2, fd, ts));
        if (!more) return nit;
        it = nit;
        tile = n1;
        n1 = n2;
    }
}

template <typename R, int KIND>
__device__ __forceinline__ StageArgs<R> slot_args(const SlotDesc<R>* sd) {
    StageArgs<R> a;
    Ctrl* ctrl = uni(sd->ctrl);
    a.ctrl = ctrl;
    a.lam_bits = &ctrl->lam_bits;
    a.geo_e = uni(sd->geo_e);
    a.geo_f = uni(sd->geo_f);
    a.nbr = uni(sd->nbr);
    a.K = uni(sd->K);
    a.Kp = uni(sd->Kp);
    a.kind = KIND;
    a.ntiles = 0;
    a.xcd_remap = 0;
    a.pad = 0;
    a.bnd = nullptr;  // batched solves have no boundary sides
    a.phys.gamma = uni(sd->phys.gamma);
    a.phys.ax = uni(sd->phys.ax);
    a.phys.ay = uni(sd->phys.ay);
    a.phys.quirks = uni(sd->phys.quirks);
    a.phys.flux = uni(sd->phys.flux);
    R* c = uni(sd->c);
    R* ua = uni(sd->ua);
    R* ub = uni(sd->ub);
    a.acc = uni(sd->acc);
    a.c = c;
    if constexpr (KIND == ST_RK4_1 || KIND == ST_RK2_1) {
        a.u = c;
        a.out = ua;
    } else if constexpr (KIND == ST_RK4_2) {
        a.u = ua;
        a.out = ub;
    } else if constexpr (KIND == ST_RK4_3) {
        a.u = ub;
        a.out = ua;
    } else if constexpr (KIND == ST_RK4_4 || KIND == ST_RK2_2) {
        a.u = ua;
        a.out = c;
    } else {
        a.u = ub;
        a.out = ua;
    }
    return a;
}

// Work-item classes, as bits of a launch's mask: the bodies a stage kernel
// carries.  A batch runs each stage as one or more persistent launches whose
// masks cover the classes present:
//   one launch (the default): BM_ALL, so light (advection) and heavy (Euler)
//     items share the CUs;
//   per class: one launch per class, each with the register budget of its
//     own bodies.
// A kernel's register allocation, and with it the wavefronts a SIMD holds,
// is that of the bodies it carries (FP64: Euler tiles alone 2 wavefronts per
// SIMD, every body together 1).
enum : int {
    BM_LANE_ADV = 1,
    BM_LANE_EUL = 2,
    BM_TILE_ADV = 4,
    BM_TILE_EUL = 8,
    BM_LANES = BM_LANE_ADV | BM_LANE_EUL,
    BM_TILES = BM_TILE_ADV | BM_TILE_EUL,
    BM_ALL = BM_LANES | BM_TILES
};

// Work lists in `parts` (XCD x = blockIdx % 8), 9 bounds per list:
//   parts[9l + x] .. parts[9l + x + 1]   items of list l on XCD x
//   parts[36]                            resident slot count
// A launch is given parts + 9l: its lane items are list l, its tile items
// list l + 1 if it carries lane bodies too, else list l.  Workgroup bi of
// XCD x takes items parts[9l + x] + bi, + G8, + 2 G8, ... (G8 = workgroups
// per XCD).
constexpr int BATCH_LISTS = 4;
constexpr int BATCH_NSLOT_AT = 9 * BATCH_LISTS;
constexpr int BATCH_PARTS = BATCH_NSLOT_AT + 1;

// Lane items: 256 elements each, one per thread; ADV / EUL select the
// bodies compiled in (advection N = 1, Euler N = 4).
template <typename R, int KIND, bool ADV, bool EUL>
__device__ __forceinline__ void batch_lane_items(const SlotDesc<R>* __restrict__ slots,
                                                 const int2* __restrict__ items, int it, int i1, int G8,
                                                 double* sh) {
    constexpr bool LAST = (KIND == ST_RK4_4 || KIND == ST_RK2_2);
    for (; it < i1; it += G8) {
        const int2 item = items[it];
        const SlotDesc<R>* sd = slots + uni(item.x);
        if (!uni(uni(sd->ctrl)->active)) continue;
        const StageArgs<R> a = slot_args<R, KIND>(sd);
        const int N = uni(sd->N), p = uni(sd->p);
        const int e = uni(item.y) * BATCH_VALU_TILE + threadIdx.x;
        double lam = 0.0;
        if constexpr (EUL) {
            if (N == 4 && p == 2) lam = valu_process<4, 2, R>(a, e, KIND);
            else if (N == 4 && p == 1) lam = valu_process<4, 1, R>(a, e, KIND);
        }
        if constexpr (ADV) {
            if (N == 1 && p == 2) lam = valu_process<1, 2, R>(a, e, KIND);
            else if (N == 1 && p == 1) lam = valu_process<1, 1, R>(a, e, KIND);
        }
        if constexpr (LAST) {
            const double m = block_max(lam, sh);
            if (threadIdx.x == 0) lam_commit(m, a.lam_bits);
            __syncthreads();  // sh is reused by the next item
        }
    }
}

// Tile items of XCD x: runs of consecutive tiles of one solve through the
// pipelined matrix-core body.
template <typename R, int KIND, bool ADV, bool EUL>
__device__ __forceinline__ void batch_tile_items(const SlotDesc<R>* __restrict__ slots,
                                                 const int2* __restrict__ items, int it, int i1, int G8,
                                                 typename MfmaT<R>::V* fd, R* ts, typename MfmaT<R>::V* tabA,
                                                 double* sh) {
    constexpr bool LAST = (KIND == ST_RK4_4 || KIND == ST_RK2_2);
    while (it < i1) {
        const int slot = uni(items[it].x);
        const SlotDesc<R>* sd = slots + slot;
        if (!uni(uni(sd->ctrl)->active)) {
            it += G8;
            continue;
        }
        const StageArgs<R> a = slot_args<R, KIND>(sd);
        MfmaArgs mx;
        mx.fword = uni(sd->fword);
        mx.halo = uni(sd->halo);
        mx.face4 = uni(sd->face4);
        mx.metric4 = uni(sd->metric4);
        mx.hcap = uni(sd->hcap);
        mx.pad = 0;
        const int N = uni(sd->N), p = uni(sd->p);
        double lam = 0.0;
        int next = it + G8;
        if constexpr (EUL) {
            if (N == 4 && p == 4) next = batch_mfma_run<4, 4, KIND, R>(a, mx, slot, items, it, i1, G8, fd, ts, tabA, lam);
            else if (N == 4 && p == 3) next = batch_mfma_run<4, 3, KIND, R>(a, mx, slot, items, it, i1, G8, fd, ts, tabA, lam);
        }
        if constexpr (ADV) {
            if (N == 1 && p == 4) next = batch_mfma_run<1, 4, KIND, R>(a, mx, slot, items, it, i1, G8, fd, ts, tabA, lam);
            else if (N == 1 && p == 3) next = batch_mfma_run<1, 3, KIND, R>(a, mx, slot, items, it, i1, G8, fd, ts, tabA, lam);
        }
        it = next;
        if constexpr (LAST) {
            const double m = block_max(lam, sh);
            if (threadIdx.x == 0) lam_commit(m, a.lam_bits);
            __syncthreads();  // sh is reused by the next run
        }
    }
}

// Workgroups per CU each kernel's registers are held to, by mask.  One
// launch: FP32 2 (3 spills 76-220 B per lane and measured no faster); FP64
// 1, which lets the compiler keep its spills in the accumulation registers
// (AGPRs) of the single wavefront per SIMD (2 spill about 1 KB per lane to
// memory).  The others: the settings measured fastest on the ragged
// 64-solve batch among 2-6.
template <typename R, int MASK>
constexpr int batch_min_groups() {
    const bool f32 = sizeof(R) == 4;
    switch (MASK) {
        case BM_ALL: return f32 ? 2 : 1;
        case BM_LANE_ADV: return f32 ? 4 : 3;
        case BM_LANE_EUL: return f32 ? 3 : 1;
        case BM_TILE_ADV: return 4;
        default: return f32 ? 3 : 2;  // BM_TILE_EUL
    }
}

// One persistent launch per stage for each mask of the batch's form.  A
// workgroup takes its share of its XCD's items of the launch's list(s);
// lane and tile items are separate loops so the register allocation of one
// body does not burden the other.  Solves use lane items for p <= 2 and
// tiles for p >= 3 (the kernel choice of a single solve, so results match
// it bit for bit).
template <typename R, int KIND, int MASK>
__global__ void __launch_bounds__(256, (batch_min_groups<R, MASK>()))
    k_batch_stage(const SlotDesc<R>* __restrict__ slots, const int2* __restrict__ items,
                  const int* __restrict__ parts) {
    using V = typename MfmaT<R>::V;
    constexpr bool LANES = (MASK & BM_LANES) != 0, TILES = (MASK & BM_TILES) != 0;
    extern __shared__ float4 lds4[];
    __shared__ double sh[8];
    const int x = blockIdx.x & 7, G8 = gridDim.x >> 3, bi = blockIdx.x >> 3;
    if constexpr (LANES)
        batch_lane_items<R, KIND, (MASK & BM_LANE_ADV) != 0, (MASK & BM_LANE_EUL) != 0>(slots, items, parts[x] + bi,
                                                                                       parts[x + 1], G8, sh);
    if constexpr (TILES) {
        constexpr int T = LANES ? 9 : 0;  // the tile list follows the lane list
        V* fd = reinterpret_cast<V*>(lds4);
        R* ts = tile_traces<V, R>(fd);
        __shared__ V tabA[MFMA_A_SLOTS * 64];  // operand table of the current run's order
        batch_tile_items<R, KIND, (MASK & BM_TILE_ADV) != 0, (MASK & BM_TILE_EUL) != 0>(
            slots, items, parts[T + x] + bi, parts[T + x + 1], G8, fd, ts, tabA, sh);
    }
}

// the masks a batch's forms launch
constexpr int BATCH_MASKS[] = {BM_ALL, BM_LANE_ADV, BM_LANE_EUL, BM_TILE_ADV, BM_TILE_EUL};

template <typename R, int KIND>
const void* batch_stage_fn_k(int mask) {
    switch (mask) {
        case BM_ALL: return reinterpret_cast<const void*>(&k_batch_stage<R, KIND, BM_ALL>);
        case BM_LANE_ADV: return reinterpret_cast<const void*>(&k_batch_stage<R, KIND, BM_LANE_ADV>);
        case BM_LANE_EUL: return reinterpret_cast<const void*>(&k_batch_stage<R, KIND, BM_LANE_EUL>);
        case BM_TILE_ADV: return reinterpret_cast<const void*>(&k_batch_stage<R, KIND, BM_TILE_ADV>);
        case BM_TILE_EUL: return reinterpret_cast<const void*>(&k_batch_stage<R, KIND, BM_TILE_EUL>);
        default: return nullptr;
    }
}

// the stage kernel of RK4 stage `kind` (ST_RK4_1..4) and mask `mask`
template <typename R>
const void* batch_stage_fn(int kind, int mask) {
    switch (kind) {
        case ST_RK4_1: return batch_stage_fn_k<R, ST_RK4_1>(mask);
        case ST_RK4_2: return batch_stage_fn_k<R, ST_RK4_2>(mask);
        case ST_RK4_3: return batch_stage_fn_k<R, ST_RK4_3>(mask);
        case ST_RK4_4: return batch_stage_fn_k<R, ST_RK4_4>(mask);
        default: return nullptr;
    }
}

template <typename R>
__global__ void k_batch_dt(const SlotDesc<R>* __restrict__ slots, const int* __restrict__ nslots) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < *nslots) dt_update(slots[s].ctrl);
}

// every slot's control block into one array (one host copy)
template <typename R>
__global__ void k_batch_ctrls(const SlotDesc<R>* __restrict__ slots, int nslots, Ctrl* out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nslots) out[s] = *slots[s].ctrl;
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

// Device side of a batch: slot table, work lists, XCD partition, the
// persistent launches of a stage and the captured step.  A launch's grid is
// the workgroups its kernel keeps resident (or the set_max_groups cap) and
// its LDS fits the largest tile halo of its solves; the slot count lives on
// the device.  The step graph stays valid while set_slots changes which
// solves are resident, as long as the capacities and the launches (class,
// grid, LDS) stay the same.
template <typename R>
class EngineT : public BatchEngine {
   public:
    struct Launch {
        int list = 0;     // its first work list (lanes, then tiles when it carries both)
        int mask = BM_ALL;
        int grid = 0;     // workgroups launched
        int resident = 0; // workgroups of this kernel that fit on the device at once
        size_t lds = 0;   // dynamic LDS per workgroup
        bool operator==(const Launch& o) const {
            return list == o.list && mask == o.mask && grid == o.grid && lds == o.lds;
        }
        int tile_list() const { return list + ((mask & BM_LANES) ? 1 : 0); }
    };
    hipStream_t st = nullptr;
    SlotDesc<R>* d_slots = nullptr;
    int2* d_items = nullptr;
    int* d_xcd = nullptr;   // BATCH_PARTS ints: XCD chunk bounds of the work lists, slot count
    double* d_status = nullptr;
    Ctrl* d_ctrls = nullptr;
    int cap_slots = 0, cap_items = 0;
    int nslots = 0, nitems = 0, dt_grid = 0, cus = 0;
    int max_groups_req = 0;          // set_max_groups (0: no cap)
    // set_launch_form; the default is the form measured faster on the
    // 64-solve ragged mix in each precision (FP32 one launch: 1.15x the
    // per-class form; FP64 per class: 1.12x the one launch)
    int form = sizeof(R) == 8 ? DG_BATCH_PER_CLASS : DG_BATCH_ONE_LAUNCH;
    std::vector<Launch> launches;    // the batch kernels of a stage, in launch order
    std::vector<int2> raw_items[BATCH_LISTS];  // the work lists in solve order, and
    std::vector<double> raw_cost[BATCH_LISTS]; // each item's estimated cost
    std::vector<int2> h_items;       // host copy of the work lists and their XCD bounds
    std::vector<int> h_parts;
    hipGraphExec_t gexec = nullptr;
    int solo_min = 0;             // 0: every solve in the batch launch
    std::vector<Solver*> solo;    // solves launched with their own kernels
    std::vector<uint64_t> solo_uid;  // their ids: the captured step holds their buffers

    ~EngineT() override {
        if (st) (void)hipStreamSynchronize(st);
        if (gexec) hipGraphExecDestroy(gexec);
        void* b[] = {d_slots, d_items, d_xcd, d_status, d_ctrls};
        for (void* q : b)
            if (q) hipFree(q);
        if (st) hipStreamDestroy(st);
    }

    Status init(int max_slots, int max_items) {
        Status s0 = ensure_device_ready();
        if (!s0) return s0;
        HIPST(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIPST(hipMalloc((void**)&d_xcd, BATCH_PARTS * sizeof(int)));
        HIPST(hipMemsetAsync(d_xcd, 0, BATCH_PARTS * sizeof(int), st));
        int dev = 0;
        HIPST(hipGetDevice(&dev));
        HIPST(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        if constexpr (sizeof(R) == 8) {  // above 64 KB: raise the kernels' dynamic LDS limit
            int mxl = 0;
            HIPST(hipDeviceGetAttribute(&mxl, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
            for (int k : {ST_RK4_1, ST_RK4_2, ST_RK4_3, ST_RK4_4})
                for (int m : BATCH_MASKS)
                    HIPST(hipFuncSetAttribute(batch_stage_fn<R>(k, m), hipFuncAttributeMaxDynamicSharedMemorySize,
                                              mxl - 16 * 1024));
        }
        Status s1 = reserve(max_slots, max_items);
        if (!s1) return s1;
        HIPST(hipStreamSynchronize(st));
        return Status::ok();
    }

    // LDS and grid of a launch: LDS for the largest tile halo among its
    // solves (at the Euler size when it carries Euler tiles), the grid every
    // workgroup that fits at once for all four stage kernels, or the cap set
    // by set_max_groups.
    Status shape(Launch& L, int hcap) {
        const bool tiles = (L.mask & BM_TILES) != 0;
        L.lds = tiles ? mfma_tile_lds_bytes((L.mask & BM_TILE_EUL) ? 4 : 1, hcap, (int)sizeof(R)) : 0;
        int per_cu = 1 << 30;
        for (int k : {ST_RK4_1, ST_RK4_2, ST_RK4_3, ST_RK4_4}) {
            int q = 0;
            HIPST(hipOccupancyMaxActiveBlocksPerMultiprocessor(&q, batch_stage_fn<R>(k, L.mask), 256, L.lds));
            per_cu = std::min(per_cu, q);
        }
        L.resident = std::max(8, per_cu * cus / 8 * 8);
        L.grid = max_groups_req == 0 ? L.resident : std::min(L.resident, max_groups_req / 8 * 8);
        return Status::ok();
    }

    // grows the device tables; invalidates the captured step when it does
    Status reserve(int ns, int ni) {
        if (ns > cap_slots) {
            if (d_slots) hipFree(d_slots);
            if (d_status) hipFree(d_status);
            if (d_ctrls) hipFree(d_ctrls);
            d_slots = nullptr;
            d_status = nullptr;
            d_ctrls = nullptr;
            cap_slots = std::max(ns, 2 * cap_slots);
            HIPST(hipMalloc((void**)&d_slots, cap_slots * sizeof(SlotDesc<R>)));
            HIPST(hipMalloc((void**)&d_status, cap_slots * 3 * sizeof(double)));
            HIPST(hipMalloc((void**)&d_ctrls, cap_slots * sizeof(Ctrl)));
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

    void set_solo_threshold(int min_elems) override {
        if (min_elems != solo_min) drop_graph();
        solo_min = std::max(0, min_elems);
    }

    void set_launch_form(int f) override { form = f; }

    // the masks of the form's launches, in launch order (Euler tiles first:
    // the longest items start earliest)
    std::vector<int> form_masks() const {
        switch (form) {
            case DG_BATCH_PER_CLASS: return {BM_TILE_EUL, BM_TILE_ADV, BM_LANE_EUL, BM_LANE_ADV};
            default: return {BM_ALL};
        }
    }

    // Splits each list into 8 contiguous chunks of equal estimated cost, one
    // per XCD, and uploads the lists with their bounds (stream-ordered;
    // waits for the copies).
    Status upload_lists() {
        std::vector<int> xs(BATCH_PARTS, 0);
        std::vector<int2> hi;
        for (int l = 0; l < BATCH_LISTS; l++) {
            xcd_chunks(raw_items[l], raw_cost[l], hi, &xs[9 * l]);
        }
        xs[BATCH_NSLOT_AT] = nslots;
        const int ni = (int)hi.size();
        Status s0 = reserve(std::max(nslots, 1), std::max(ni, 1));
        if (!s0) return s0;
        if (ni) HIPST(hipMemcpyAsync(d_items, hi.data(), ni * sizeof(int2), hipMemcpyHostToDevice, st));
        HIPST(hipMemcpyAsync(d_xcd, xs.data(), BATCH_PARTS * sizeof(int), hipMemcpyHostToDevice, st));
        HIPST(hipStreamSynchronize(st));  // the host vectors go out of scope
        nitems = ni;
        h_items.swap(hi);
        h_parts.swap(xs);
        return Status::ok();
    }

    Status set_slots(const std::vector<Solver*>& solvers) override {
        const int ns = (int)solvers.size();
        std::vector<Solver*> new_solo;
        std::vector<SlotDesc<R>> hs(ns);
        std::vector<int2> list[BATCH_LISTS];
        std::vector<double> cost[BATCH_LISTS];
        int hmax[BATCH_LISTS] = {16, 16, 16, 16};  // largest tile halo per list
        // the form's launches and the work list of each class: a launch's
        // lane list, then its tile list
        std::vector<Launch> forms;
        int list_of[BM_ALL + 1] = {};
        int next = 0;
        for (int m : form_masks()) {
            Launch L;
            L.mask = m;
            L.list = next;
            if (m & BM_LANES) {
                for (int bit : {BM_LANE_ADV, BM_LANE_EUL})
                    if (m & bit) list_of[bit] = next;
                next++;
            }
            if (m & BM_TILES) {
                for (int bit : {BM_TILE_ADV, BM_TILE_EUL})
                    if (m & bit) list_of[bit] = next;
                next++;
            }
            forms.push_back(L);
        }
        for (int s = 0; s < ns; s++) {
            SlotInfo si;
            solvers[s]->slot_info(&si);
            const bool tiles = si.kind >= 2;
            // solves with boundary sides always run their own kernels (the
            // batched bodies carry no boundary code)
            const bool solo_here = si.boundary || (solo_min > 0 && si.K >= solo_min);
            if (!solo_here && tiles != (si.p >= 3))
                return Status::err(DG_ERR_STATE, "batch slots use matrix-core tiles exactly when p >= 3");
            if (!solo_here && tiles && si.hcap > MFMA_MAX_HALO)
                return Status::err(DG_ERR_UNSUPPORTED, "tile halo too large");
            if (!solo_here && si.N != 1 && si.N != 4)
                return Status::err(DG_ERR_STATE, "batch slots carry advection (1) or Euler (4) variables");
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
            d.face4 = si.face4;
            d.metric4 = si.metric4;
            d.ctrl = si.ctrl;
            d.K = si.K;
            d.Kp = si.Kp;
            d.N = si.N;
            d.p = si.p;
            d.hcap = si.hcap;
            d.path = tiles ? 2 : 1;
            d.phys = si.phys;
            if (solo_here) {
                new_solo.push_back(solvers[s]);  // slot kept for dt and status; no work items
                continue;
            }
            const int bit = tiles ? (si.N == 4 ? BM_TILE_EUL : BM_TILE_ADV) : (si.N == 4 ? BM_LANE_EUL : BM_LANE_ADV);
            const int l = list_of[bit];
            if (tiles) hmax[l] = std::max(hmax[l], si.hcap);
            const int te = tiles ? MFMA_TILE : BATCH_VALU_TILE;
            const int nt = (si.K + te - 1) / te;
            const double per_elem = (double)si.N * n_modes(si.p) * sizeof(R);
            for (int t = 0; t < nt; t++) {
                list[l].push_back(int2{s, t});
                cost[l].push_back(per_elem * std::min(te, si.K - t * te));
            }
        }
        // the launches of a stage: the form's masks that have items
        std::vector<Launch> nl;
        for (const Launch& L0 : forms) {
            const bool lanes = (L0.mask & BM_LANES) != 0, tiles = (L0.mask & BM_TILES) != 0;
            if ((!lanes || list[L0.list].empty()) && (!tiles || list[L0.tile_list()].empty())) continue;
            Launch L = L0;
            Status sh = shape(L, tiles ? hmax[L.tile_list()] : 16);
            if (!sh) return sh;
            nl.push_back(L);
        }
        if (nl != launches) {
            launches = nl;
            drop_graph();
        }
        for (int l = 0; l < BATCH_LISTS; l++) {
            raw_items[l].swap(list[l]);
            raw_cost[l].swap(cost[l]);
        }
        nslots = ns;
        Status s0 = reserve(std::max(ns, 1), 1);
        if (!s0) return s0;
        if (ns) HIPST(hipMemcpyAsync(d_slots, hs.data(), ns * sizeof(SlotDesc<R>), hipMemcpyHostToDevice, st));
        HIPST(hipStreamSynchronize(st));  // hs goes out of scope
        Status s1 = upload_lists();
        if (!s1) return s1;
        // The solo launches are part of the captured step, with the buffers of
        // the solves they were captured for; compare by id, since a new solve
        // can sit at the address of a deleted one.
        std::vector<uint64_t> new_uid;
        for (Solver* sv : new_solo) new_uid.push_back(sv->uid);
        if (new_uid != solo_uid) {
            solo = new_solo;
            solo_uid = new_uid;
            drop_graph();
        }
        return Status::ok();
    }

    Status launch_stage(int kind) {
        for (const Launch& L : launches) {
            const void* fn = batch_stage_fn<R>(kind, L.mask);
            if (!fn) return Status::err(DG_ERR_ARG, "unsupported batch stage");
            const SlotDesc<R>* slots = d_slots;
            const int2* items = d_items;
            const int* parts = d_xcd + 9 * L.list;
            void* args[] = {(void*)&slots, (void*)&items, (void*)&parts};
            HIPST(hipLaunchKernel(fn, dim3(L.grid), dim3(256), args, L.lds, st));
        }
        return Status::ok();
    }

    Status step_direct() {
        hipLaunchKernelGGL((k_batch_dt<R>), dim3(dt_grid), dim3(64), 0, st, d_slots, d_xcd + BATCH_NSLOT_AT);
        HIPST(hipGetLastError());
        Status s0;
        for (int k : {ST_RK4_1, ST_RK4_2, ST_RK4_3, ST_RK4_4}) {
            if (!(s0 = launch_stage(k))) return s0;
            for (Solver* sv : solo)
                if (!(s0 = sv->stage_on(k, st))) return s0;
        }
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

    Status pull_ctrls(std::vector<Ctrl>& out) override {
        out.assign((size_t)nslots, Ctrl{});
        if (nslots == 0) return sync();
        hipLaunchKernelGGL((k_batch_ctrls<R>), dim3((nslots + 63) / 64), dim3(64), 0, st, d_slots, nslots, d_ctrls);
        HIPST(hipGetLastError());
        HIPST(hipMemcpyAsync(out.data(), d_ctrls, out.size() * sizeof(Ctrl), hipMemcpyDeviceToHost, st));
        HIPST(hipStreamSynchronize(st));
        return Status::ok();
    }

    Status sync() override {
        HIPST(hipStreamSynchronize(st));
        return Status::ok();
    }
    hipStream_t stream() override { return st; }
    int max_tile_halo() const override { return MFMA_MAX_HALO; }

    Status set_max_groups(int n) override {
        if (n < 0 || (n > 0 && n < 8)) return Status::err(DG_ERR_ARG, "max_groups must be 0 or at least 8");
        max_groups_req = n;
        bool changed = false;
        for (Launch& L : launches) {
            const int g = n == 0 ? L.resident : std::min(L.resident, n / 8 * 8);
            changed = changed || g != L.grid;
            L.grid = g;
        }
        if (changed) drop_graph();
        return Status::ok();
    }

    // Workgroup bi of XCD x takes the positions parts[9l + x] + bi, + G8, ...
    // of list l; a run is a stretch of those that belong to one solve
    // (batch_mfma_run).
    void launch_info(int* o) const override {
        o[0] = o[1] = o[2] = o[3] = 0;
        if (h_parts.size() < (size_t)BATCH_PARTS) return;
        const auto count = [&](int l) { return (int)raw_items[l].size(); };
        int longest = 0;
        for (const Launch& L : launches) {
            o[0] = std::max(o[0], L.grid);
            if (L.mask & BM_LANES) o[2] += count(L.list);
            if (!(L.mask & BM_TILES)) continue;
            const int tl = L.tile_list(), G8 = L.grid / 8;
            o[1] += count(tl);
            for (int x = 0; x < 8; x++) {
                const int i0 = h_parts[9 * tl + x], i1 = h_parts[9 * tl + x + 1];
                for (int bi = 0; bi < G8; bi++) {
                    int run = 0, slot = -1;
                    for (int it = i0 + bi; it < i1; it += G8) {
                        run = (h_items[it].x == slot) ? run + 1 : 1;
                        slot = h_items[it].x;
                        longest = std::max(longest, run);
                    }
                }
            }
        }
        o[3] = longest;
    }
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
    // the batch's fixed per-order choice (no autotuning launches at setup);
    // a solve with boundary sides runs its own kernels (the same choice)
    // inside the batch step
    const int kind = p >= 3 ? 2 : 1;
    Solver* s = nullptr;
    Status st = Solver::create(m, prob, p, precision, &s, kind);
    if (!st) return st;
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
        // on the engine's stream, ahead of the steps: no wait per solve
        for (size_t s = 0; s < solvers.size(); s++)
            if (!(st = solvers[s]->begin_on(t_end[s], has_end, cfl, eng->stream()))) return st;
        return Status::ok();
    }

    // every control block back in one copy (t, steps and the wave speed the
    // next run starts from)
    Status finish_all() {
        std::vector<Ctrl> h;
        Status st = eng->pull_ctrls(h);
        if (!st) return st;
        if (h.size() != solvers.size()) return Status::err(DG_ERR_STATE, "batch slots out of date");
        for (size_t s = 0; s < solvers.size(); s++) {
            solvers[s]->set_host_ctrl(h[s]);
            solvers[s]->set_lambda_valid(true);
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

    Status set_max_groups(int n) override { return eng->set_max_groups(n); }
    void set_launch_form(int form) override {
        eng->set_launch_form(form);
        dirty = true;  // the work lists are rebuilt at the next run
    }
    // the work list is uploaded at the next run; describe that one
    void launch_info(int* out) const override { eng->launch_info(out); }

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
