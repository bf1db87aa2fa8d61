// Matrix-core (MFMA) stage kernel of one solve, FP32 or FP64 bulk (chosen by
// the autotuner per order and size; p = 1..4).
//
// Persistent: each workgroup walks a strided sequence of 64-element tiles
// inside the contiguous tile range of its XCD (so Morton neighbours share
// that XCD's L2), and the state of its next tile is in flight while it
// computes the current one.  The per-tile work is mfma_process_tile_pipe
// (dg_mfma_body.h), shared with the batched kernel.
#include "dg_mfma_body.h"

namespace dgmi {

DGMI_TABLE_UPLOADER(upload_tables_mfma)

bool mfma_supported(int N, int p) { return (N == 1 || N == 4) && p >= 1 && p <= 4; }

// In-bounds check of the builtin's lane layout on one wavefront:
// A[i][k] from lane i + 16k, B[k][j] from lane j + 16k, D[4(l/16)+r][l%16]
// in register r of lane l.
__global__ void k_mfma_layout_probe(float* out) {
    const int l = threadIdx.x;
    const int i = l & 15, k = l >> 4;
    const float a = (float)(i + 1) + 0.5f * (float)k;         // A[i][k]
    const float b = (float)(2 * k + 1) - 0.25f * (float)(i);  // B[k][j = i]
    floatx4 d = {0.f, 0.f, 0.f, 0.f};
    d = mfma4(a, b, d);
    for (int r = 0; r < 4; r++) out[(4 * (l >> 4) + r) * 16 + (l & 15)] = d[r];
}

hipError_t mfma_layout_probe(float* host_out) {
    float* d = nullptr;
    hipError_t e = hipMalloc(&d, 256 * sizeof(float));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mfma_layout_probe, dim3(1), dim3(64), 0, 0, d);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(host_out, d, 256 * sizeof(float), hipMemcpyDeviceToHost);
    hipFree(d);
    return e;
}

// The same check for v_mfma_f64_16x16x4f64, whose results sit at
// D[g + 4r][j] (register r of lane j + 16g).
__global__ void k_mfma_layout_probe_f64(double* out) {
    const int l = threadIdx.x;
    const int i = l & 15, k = l >> 4;
    const double a = (double)(i + 1) + 0.5 * (double)k;
    const double b = (double)(2 * k + 1) - 0.25 * (double)(i);
    doublex4 d = {0.0, 0.0, 0.0, 0.0};
    d = MfmaT<double>::mma(a, b, d);
    for (int r = 0; r < 4; r++) out[MfmaT<double>::pos(l >> 4, r) * 16 + (l & 15)] = d[r];
}

hipError_t mfma_layout_probe_f64(double* host_out) {
    double* d = nullptr;
    hipError_t e = hipMalloc(&d, 256 * sizeof(double));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mfma_layout_probe_f64, dim3(1), dim3(64), 0, 0, d);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(host_out, d, 256 * sizeof(double), hipMemcpyDeviceToHost);
    hipFree(d);
    return e;
}

// occupancy target (waves per SIMD) the register allocation is held to
// at least 3 workgroups (12 wavefronts) per CU in FP32: the register budget
// the p = 4 body fits without spills, and the faster setting on MI355X.
// FP64 doubles every value's registers: 2 workgroups per CU (256 registers).
#ifndef DGMI_MFMA_MINW
#define DGMI_MFMA_MINW 3
#endif
template <typename R>
constexpr int mfma_minw() { return sizeof(R) == 4 ? DGMI_MFMA_MINW : 2; }

// BND: the variant for meshes with boundary sides (Euler only)
template <int N, int P, int KIND, typename R, bool BND = false>
__global__ void __launch_bounds__(256, mfma_minw<R>()) k_stage_mfma(StageArgs<R> a, MfmaArgs mx) {
    using V = typename MfmaT<R>::V;
    constexpr bool LAST = (KIND == ST_RK4_4 || KIND == ST_RK2_2);
    extern __shared__ float4 lds4[];
    V* fd = reinterpret_cast<V*>(lds4);
    R* ts = tile_traces<V, R>(fd);
    __shared__ double shmax[8];

    if (!a.ctrl->active) return;
    const int ntiles = a.ntiles;
    int t, tend, stride;
    if (a.xcd_remap) {
        const int T8 = (ntiles + 7) >> 3, x = blockIdx.x & 7;
        t = x * T8 + (blockIdx.x >> 3);
        tend = min((x + 1) * T8, ntiles);
        stride = gridDim.x >> 3;
    } else {
        t = blockIdx.x;
        tend = ntiles;
        stride = gridDim.x;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int g = lane >> 4, j = lane & 15;
    double lam = 0.0;
    __shared__ V tabA[MFMA_A_SLOTS * 64];
    const R dt = (R)a.ctrl->dt;
    MfmaTileState<N, P, R> st;
    int hid[MfmaTileState<N, P, R>::HPW];
    TileGeo<R> gc;
    // tiles past the end of the walk read tile tend-1 instead (the loads
    // stay unconditional and in bounds; nothing is computed from them)
    auto walk = [&](int x) { return min(x, tend - 1); };
    if (t < tend) {
        MfmaTileState<N, P, R>::load_ids(mx.halo, mx.hcap, t, w, j, hid);
        gc.load(a, mx, t, w * 16 + j);
        st.load_data(a.u, a.K, t, w, g, j, hid);
        MfmaTileState<N, P, R>::load_ids(mx.halo, mx.hcap, walk(t + stride), w, j, hid);
    }
    if (w == 0) mfma_a_to_lds<N, P, R>(tabA, lane);
    // The first tile needs these loads at once.  Waiting for them here also
    // keeps the compiler's count of loads in flight at the loop head to the
    // steady state (loop back edge), where the next tile's operands were
    // issued before the previous tile's stores and need not wait for them.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (int tc = t; tc < tend; tc += stride)
        lam = fmax(lam, mfma_process_tile_pipe<N, P, KIND, R, BND>(a, mx, tabA, st, hid, gc, dt, tc, walk(tc + stride),
                                                              walk(tc + 2 * stride), fd, ts));

    if constexpr (LAST) {
        const double m2 = block_max(lam, shmax);
        if (threadIdx.x == 0) lam_commit(m2, a.lam_bits);
    }
}

size_t mfma_lds_bytes(int N, int hcap, int rs) { return mfma_tile_lds_bytes(N, hcap, rs); }

// Packs the geometry of the matrix-core kernel (one thread per element):
// a tile's 192 face records and 64 element maps become contiguous entries
// of 4 values (FP32: the neighbour word's bits in the 4th value, FP64: its
// value).
template <typename R>
__global__ void __launch_bounds__(256) k_mfma_pack(const R* __restrict__ geo_e, const R* __restrict__ geo_f,
                                                  const uint32_t* __restrict__ fword, int K, int Kp,
                                                  typename MfmaT<R>::V* __restrict__ face4,
                                                  typename MfmaT<R>::V* __restrict__ metric4) {
    using V = typename MfmaT<R>::V;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Kp) return;
    const size_t s = (size_t)min(e, K - 1), kp = (size_t)Kp;
    for (int f = 0; f < 3; f++) {
        const int wd = (int)(fword[f * kp + s] & (e < K ? ~0u : ~FW_OWNER));
        R wv;
        if constexpr (sizeof(R) == 4)
            wv = __int_as_float(wd);
        else
            wv = (R)wd;
        face4[3 * (size_t)e + f] = V{geo_f[3 * f * kp + s], geo_f[(3 * f + 1) * kp + s], geo_f[(3 * f + 2) * kp + s], wv};
    }
    metric4[e] = V{geo_e[s], geo_e[kp + s], geo_e[2 * kp + s], geo_e[3 * kp + s]};
}

template <typename R>
hipError_t launch_mfma_pack(const R* geo_e, const R* geo_f, const uint32_t* fword, int K, int Kp, void* face4,
                            void* metric4, hipStream_t st) {
    using V = typename MfmaT<R>::V;
    hipLaunchKernelGGL((k_mfma_pack<R>), dim3((Kp + 255) / 256), dim3(256), 0, st, geo_e, geo_f, fword, K, Kp,
                       (V*)face4, (V*)metric4);
    return hipGetLastError();
}
template hipError_t launch_mfma_pack<float>(const float*, const float*, const uint32_t*, int, int, void*, void*,
                                            hipStream_t);
template hipError_t launch_mfma_pack<double>(const double*, const double*, const uint32_t*, int, int, void*, void*,
                                             hipStream_t);

// workgroups of the matrix-core kernel that fit on the device at once
template <typename R, bool BND>
static const void* stage2_kernel(int N, int p) {
    if (N == 1 && p == 4) return reinterpret_cast<const void*>(&k_stage_mfma<1, 4, ST_RK4_2, R, BND>);
    if (N == 1 && p == 3) return reinterpret_cast<const void*>(&k_stage_mfma<1, 3, ST_RK4_2, R, BND>);
    if (N == 1 && p == 2) return reinterpret_cast<const void*>(&k_stage_mfma<1, 2, ST_RK4_2, R, BND>);
    if (N == 1 && p == 1) return reinterpret_cast<const void*>(&k_stage_mfma<1, 1, ST_RK4_2, R, BND>);
    if (N == 4 && p == 4) return reinterpret_cast<const void*>(&k_stage_mfma<4, 4, ST_RK4_2, R, BND>);
    if (N == 4 && p == 3) return reinterpret_cast<const void*>(&k_stage_mfma<4, 3, ST_RK4_2, R, BND>);
    if (N == 4 && p == 2) return reinterpret_cast<const void*>(&k_stage_mfma<4, 2, ST_RK4_2, R, BND>);
    if (N == 4 && p == 1) return reinterpret_cast<const void*>(&k_stage_mfma<4, 1, ST_RK4_2, R, BND>);
    return nullptr;
}
template <typename R>
static int resident_groups(int N, int p, int hcap, bool bnd) {
    const void* fn = bnd ? stage2_kernel<R, true>(N, p) : stage2_kernel<R, false>(N, p);
    if (!fn) return 0;
    int per_cu = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, mfma_lds_bytes(N, hcap, (int)sizeof(R))) !=
        hipSuccess)
        return 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return per_cu * cus;
}
int mfma_resident_groups(int N, int p, int hcap, int rs, bool bnd) {
    return rs == 8 ? resident_groups<double>(N, p, hcap, bnd) : resident_groups<float>(N, p, hcap, bnd);
}

// FP64 launches take up to 86 KB of dynamic LDS: raise each kernel's
// dynamic LDS limit to the device's per-workgroup maximum once
template <typename K>
static hipError_t allow_large_lds(K* fn) {
    int dev = 0, mx = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&mx, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (e == hipSuccess) e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize, mx - 16 * 1024);
    return e;
}

template <int N, int P, typename R, bool BND>
static hipError_t prepare_lds() {
    if constexpr (sizeof(R) == 4) {
        return hipSuccess;
    } else {
        static const hipError_t once = [] {
            hipError_t e = allow_large_lds(&k_stage_mfma<N, P, ST_RHS, R, BND>);
            if (e == hipSuccess) e = allow_large_lds(&k_stage_mfma<N, P, ST_RK4_1, R, BND>);
            if (e == hipSuccess) e = allow_large_lds(&k_stage_mfma<N, P, ST_RK4_2, R, BND>);
            if (e == hipSuccess) e = allow_large_lds(&k_stage_mfma<N, P, ST_RK4_3, R, BND>);
            if (e == hipSuccess) e = allow_large_lds(&k_stage_mfma<N, P, ST_RK4_4, R, BND>);
            if (e == hipSuccess) e = allow_large_lds(&k_stage_mfma<N, P, ST_RK2_1, R, BND>);
            if (e == hipSuccess) e = allow_large_lds(&k_stage_mfma<N, P, ST_RK2_2, R, BND>);
            return e;
        }();
        return once;
    }
}

template <int N, int P, typename R, bool BND = false>
static hipError_t launch_np(const StageArgs<R>& a, const MfmaArgs& mx, dim3 grid, hipStream_t st) {
    const size_t lds = mfma_lds_bytes(N, mx.hcap, (int)sizeof(R));
    const hipError_t pe = prepare_lds<N, P, R, BND>();
    if (pe != hipSuccess) return pe;
    switch (a.kind) {
        case ST_RHS: hipLaunchKernelGGL((k_stage_mfma<N, P, ST_RHS, R, BND>), grid, dim3(256), lds, st, a, mx); break;
        case ST_RK4_1: hipLaunchKernelGGL((k_stage_mfma<N, P, ST_RK4_1, R, BND>), grid, dim3(256), lds, st, a, mx); break;
        case ST_RK4_2: hipLaunchKernelGGL((k_stage_mfma<N, P, ST_RK4_2, R, BND>), grid, dim3(256), lds, st, a, mx); break;
        case ST_RK4_3: hipLaunchKernelGGL((k_stage_mfma<N, P, ST_RK4_3, R, BND>), grid, dim3(256), lds, st, a, mx); break;
        case ST_RK4_4: hipLaunchKernelGGL((k_stage_mfma<N, P, ST_RK4_4, R, BND>), grid, dim3(256), lds, st, a, mx); break;
        case ST_RK2_1: hipLaunchKernelGGL((k_stage_mfma<N, P, ST_RK2_1, R, BND>), grid, dim3(256), lds, st, a, mx); break;
        case ST_RK2_2: hipLaunchKernelGGL((k_stage_mfma<N, P, ST_RK2_2, R, BND>), grid, dim3(256), lds, st, a, mx); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// a.bnd set: the mesh has boundary sides and the launch takes the BND
// variant
template <typename R>
hipError_t launch_stage_mfma(int N, int p, const StageArgs<R>& a, const MfmaArgs& mx, dim3 grid, hipStream_t st) {
    if (a.bnd) {
        if (N == 4 && p == 4) return launch_np<4, 4, R, true>(a, mx, grid, st);
        if (N == 4 && p == 3) return launch_np<4, 3, R, true>(a, mx, grid, st);
        if (N == 4 && p == 2) return launch_np<4, 2, R, true>(a, mx, grid, st);
        if (N == 4 && p == 1) return launch_np<4, 1, R, true>(a, mx, grid, st);
        if (N == 1 && p == 4) return launch_np<1, 4, R, true>(a, mx, grid, st);
        if (N == 1 && p == 3) return launch_np<1, 3, R, true>(a, mx, grid, st);
        if (N == 1 && p == 2) return launch_np<1, 2, R, true>(a, mx, grid, st);
        if (N == 1 && p == 1) return launch_np<1, 1, R, true>(a, mx, grid, st);
        return hipErrorInvalidValue;
    }
    if (N == 4 && p == 4) return launch_np<4, 4, R>(a, mx, grid, st);
    if (N == 4 && p == 3) return launch_np<4, 3, R>(a, mx, grid, st);
    if (N == 4 && p == 2) return launch_np<4, 2, R>(a, mx, grid, st);
    if (N == 4 && p == 1) return launch_np<4, 1, R>(a, mx, grid, st);
    if (N == 1 && p == 4) return launch_np<1, 4, R>(a, mx, grid, st);
    if (N == 1 && p == 3) return launch_np<1, 3, R>(a, mx, grid, st);
    if (N == 1 && p == 2) return launch_np<1, 2, R>(a, mx, grid, st);
    if (N == 1 && p == 1) return launch_np<1, 1, R>(a, mx, grid, st);
    return hipErrorInvalidValue;
}
template hipError_t launch_stage_mfma<float>(int, int, const StageArgs<float>&, const MfmaArgs&, dim3, hipStream_t);
template hipError_t launch_stage_mfma<double>(int, int, const StageArgs<double>&, const MfmaArgs&, dim3, hipStream_t);

}  // namespace dgmi
