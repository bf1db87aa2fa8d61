// Fused element kernels on the vector ALUs: one lane per element.
//
// One launch per RK stage does what the reference does with four
// (eval_surface, eval_volume, eval_rhs, rk_tempstorage;
// src/2d/conservation/time_integrator.cu:274-307):
//   volume integral    (conserv_kernels.cu:766-811)
//   LLF surface terms  (conserv_kernels.cu:628-758) evaluated for the three
//                      faces of the element, each face flux computed in the
//                      side's canonical left->right orientation so both
//                      elements of a side obtain bit-identical fluxes
//                      (race-free, no atomics, no face buffers)
//   gather + M^-1      (time_integrator.cu:130-188, M = J*I)
//   RK stage update    (time_integrator.cu:194-218, 445-464)
//   CFL wave speeds    (euler.cu:252-278) of the new state on the last stage,
//                      reduced per workgroup with wavefront shuffles.
#include <algorithm>

#include "dg_valu_body.h"

namespace dgmi {

DGMI_TABLE_UPLOADER(upload_tables_kernels)

// BND: the variant for meshes with boundary sides (p <= VALU_BOUNDARY_MAX_P)
template <int N, int P, typename R, bool BND = false>
__global__ void __launch_bounds__(128) k_stage_valu(StageArgs<R> a) {
    __shared__ double sh[8];
    if (!a.ctrl->active) return;
    const int tile = tile_of_block(blockIdx.x, gridDim.x, a.xcd_remap);
    if (tile >= a.ntiles) return;
    const int kind = a.kind;
    const double lam = valu_process<N, P, R, BND>(a, tile * blockDim.x + threadIdx.x, kind);
    if (kind == ST_RK4_4 || kind == ST_RK2_2) {
        const double mx = block_max(lam, sh);
        if (threadIdx.x == 0) lam_commit(mx, a.lam_bits);
    }
}

// Maximum cell wave speed of a state without stepping it (first step of a
// run, dg_cfl_dt): grid-stride over elements, workgroup max, one atomic max.
template <int N, typename R>
__global__ void __launch_bounds__(256) k_lambda(const R* u, StateIx ix, int K, double b0, PhysP ph,
                                               unsigned long long* lam_bits) {
    __shared__ double sh[8];
    double mx = 0.0;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < K; e += gridDim.x * blockDim.x) {
        double U[4];
#pragma unroll
        for (int v = 0; v < N; v++) U[v] = (double)u[ix.at(v, 0, e)] * b0;
        mx = fmax(mx, cell_speed<N>(ph, U));
    }
    mx = block_max(mx, sh);
    if (threadIdx.x == 0) lam_commit(mx, lam_bits);
}

__global__ void k_reset_lam(Ctrl* ctrl) {
    if (threadIdx.x == 0 && blockIdx.x == 0) ctrl->lam_bits = 0ull;
}

// dt = cfl * min_r / max lambda / (2p+1), clamped to t_end
// (time_integrator.cu:254-270); reads and clears the accumulated maximum.
__global__ void k_dt(Ctrl* ctrl) {
    if (threadIdx.x == 0 && blockIdx.x == 0) dt_update(ctrl);
}

// host-chosen step: activate the stages with the given dt
__global__ void k_set_dt(Ctrl* ctrl, double dt, int advance) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctrl->dt = dt;
        ctrl->active = 1;
        ctrl->lam_bits = 0ull;
        if (advance) {
            ctrl->t += dt;
            ctrl->steps += 1;
        }
    }
}

// ---------------------------------------------------------------------------
// Initial conditions: c_i = sum_j w_j U0(x_j) phi_i(x_j) in FP64
// (conserv_kernels.cu:167-187, euler.cu:31-59), stored in the bulk precision.
// ---------------------------------------------------------------------------
template <int N, int P, typename R>
__global__ void __launch_bounds__(128) k_project_ic(R* c, StateIx ix, const double* V, int K, IcP ic) {
    constexpr int NP = n_modes(P), NQ = n_quad_vol(P);
    const int Kp = ix.Kp;
    const TabP<double, P>& T = tab<double, P>();
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= K) return;
    const double v1x = V[e], v1y = V[Kp + e], v2x = V[2 * Kp + e], v2y = V[3 * Kp + e];
    const double v3x = V[4 * Kp + e], v3y = V[5 * Kp + e];
    double u0[NQ][N];
    for (int j = 0; j < NQ; j++) {
        const double r1 = T.R1[j], r2 = T.R2[j];
        const double x = r1 * v2x + r2 * v3x + (1 - r1 - r2) * v1x;
        const double y = r1 * v2y + r2 * v3y + (1 - r1 - r2) * v1y;
        double U[4];
        ic_point(ic, x, y, U);
        for (int n = 0; n < N; n++) u0[j][n] = U[n];
    }
    for (int i = 0; i < NP; i++) {
        double U[N];
        for (int n = 0; n < N; n++) U[n] = 0.;
        for (int j = 0; j < NQ; j++)
            for (int n = 0; n < N; n++) U[n] += T.W[j] * u0[j][n] * T.B[i][j];
        for (int n = 0; n < N; n++) c[ix.at(n, i, e)] = (R)U[n];
    }
    if (ix.tiled)  // padding modes of the last group
        for (int i = NP; i < 4 * tile_groups(NP); i++)
            for (int n = 0; n < N; n++) c[ix.at(n, i, e)] = R(0);
}

// ---------------------------------------------------------------------------
// FP64 audit partials per workgroup: mass of each variable, minimum cell-mean
// density and pressure, count of non-finite / non-positive cells.
// part[blk*8 + 0..3] mass, +4 min rho, +5 min p, +6 bad count.
// ---------------------------------------------------------------------------
template <int N, typename R>
__global__ void __launch_bounds__(256) k_audit(const R* c, StateIx ix, const double* Jd, int K, double b0,
                                              double gamma, double* part) {
    __shared__ double sm[256];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    double mass[4] = {0, 0, 0, 0}, rho = 1e300, pr = 1e300, bad = 0;
    if (e < K) {
        double U[4];
        const double J = Jd[e];
        for (int v = 0; v < N; v++) {
            double c0 = (double)c[ix.at(v, 0, e)];
            mass[v] = J * c0 * b0 * 0.5;
            U[v] = c0 * b0;
        }
        rho = U[0];
        bool ok = isfinite(U[0]);
        if (N == 4) {
            for (int v = 1; v < 4; v++) ok = ok && isfinite(U[v]);
            pr = (gamma - 1.) * (U[3] - (U[1] * U[1] + U[2] * U[2]) / 2. / U[0]);
            ok = ok && U[0] > 0 && pr > 0;
        }
        bad = ok ? 0 : 1;
    }
    // fixed-shape tree reduction: deterministic
    double vals[7] = {mass[0], mass[1], mass[2], mass[3], rho, pr, bad};
    for (int k = 0; k < 7; k++) {
        sm[threadIdx.x] = vals[k];
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) {
                double x = sm[threadIdx.x], y = sm[threadIdx.x + s];
                sm[threadIdx.x] = (k == 4 || k == 5) ? fmin(x, y) : x + y;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) part[blockIdx.x * 8 + k] = sm[0];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int N, typename R>
static hipError_t launch_stage_valu_np(int p, const StageArgs<R>& a, dim3 grid, hipStream_t st) {
    if (a.bnd) {
        static_assert(VALU_BOUNDARY_MAX_P == 3, "boundary instantiations below are p = 0..3");
        switch (p) {
            case 0: hipLaunchKernelGGL((k_stage_valu<N, 0, R, true>), grid, dim3(128), 0, st, a); break;
            case 1: hipLaunchKernelGGL((k_stage_valu<N, 1, R, true>), grid, dim3(128), 0, st, a); break;
            case 2: hipLaunchKernelGGL((k_stage_valu<N, 2, R, true>), grid, dim3(128), 0, st, a); break;
            case 3: hipLaunchKernelGGL((k_stage_valu<N, 3, R, true>), grid, dim3(128), 0, st, a); break;
            default: return hipErrorInvalidValue;  // refused on the host before this
        }
        return hipGetLastError();
    }
    switch (p) {
        case 0: hipLaunchKernelGGL((k_stage_valu<N, 0, R>), grid, dim3(128), 0, st, a); break;
        case 1: hipLaunchKernelGGL((k_stage_valu<N, 1, R>), grid, dim3(128), 0, st, a); break;
        case 2: hipLaunchKernelGGL((k_stage_valu<N, 2, R>), grid, dim3(128), 0, st, a); break;
        case 3: hipLaunchKernelGGL((k_stage_valu<N, 3, R>), grid, dim3(128), 0, st, a); break;
        case 4: hipLaunchKernelGGL((k_stage_valu<N, 4, R>), grid, dim3(128), 0, st, a); break;
        case 5: hipLaunchKernelGGL((k_stage_valu<N, 5, R>), grid, dim3(128), 0, st, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename R>
hipError_t launch_stage_valu(int N, int p, const StageArgs<R>& a, dim3 grid, hipStream_t st) {
    if (N == 1) return launch_stage_valu_np<1, R>(p, a, grid, st);
    if (N == 4) return launch_stage_valu_np<4, R>(p, a, grid, st);
    return hipErrorInvalidValue;
}
template hipError_t launch_stage_valu<float>(int, int, const StageArgs<float>&, dim3, hipStream_t);
template hipError_t launch_stage_valu<double>(int, int, const StageArgs<double>&, dim3, hipStream_t);

template <typename R>
hipError_t launch_lambda(const R* u, const StateIx& ix, int K, double b0, const PhysP& ph, Ctrl* ctrl,
                         hipStream_t st) {
    hipLaunchKernelGGL(k_reset_lam, dim3(1), dim3(64), 0, st, ctrl);
    const dim3 grid(std::min((K + 255) / 256, 2048));
    unsigned long long* bits = &ctrl->lam_bits;
    if (ix.N == 1)
        hipLaunchKernelGGL((k_lambda<1, R>), grid, dim3(256), 0, st, u, ix, K, b0, ph, bits);
    else
        hipLaunchKernelGGL((k_lambda<4, R>), grid, dim3(256), 0, st, u, ix, K, b0, ph, bits);
    return hipGetLastError();
}
template hipError_t launch_lambda<float>(const float*, const StateIx&, int, double, const PhysP&, Ctrl*,
                                         hipStream_t);
template hipError_t launch_lambda<double>(const double*, const StateIx&, int, double, const PhysP&, Ctrl*,
                                          hipStream_t);

hipError_t launch_dt(Ctrl* ctrl, hipStream_t st) {
    hipLaunchKernelGGL(k_dt, dim3(1), dim3(64), 0, st, ctrl);
    return hipGetLastError();
}

hipError_t launch_set_dt(Ctrl* ctrl, double dt, int advance, hipStream_t st) {
    hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(64), 0, st, ctrl, dt, advance);
    return hipGetLastError();
}

template <int N, typename R>
static hipError_t launch_ic_np(int p, R* c, const StateIx& ix, const double* V, int K, const IcP& ic,
                               hipStream_t st) {
    dim3 grid((K + 127) / 128);
    switch (p) {
        case 0: hipLaunchKernelGGL((k_project_ic<N, 0, R>), grid, dim3(128), 0, st, c, ix, V, K, ic); break;
        case 1: hipLaunchKernelGGL((k_project_ic<N, 1, R>), grid, dim3(128), 0, st, c, ix, V, K, ic); break;
        case 2: hipLaunchKernelGGL((k_project_ic<N, 2, R>), grid, dim3(128), 0, st, c, ix, V, K, ic); break;
        case 3: hipLaunchKernelGGL((k_project_ic<N, 3, R>), grid, dim3(128), 0, st, c, ix, V, K, ic); break;
        case 4: hipLaunchKernelGGL((k_project_ic<N, 4, R>), grid, dim3(128), 0, st, c, ix, V, K, ic); break;
        case 5: hipLaunchKernelGGL((k_project_ic<N, 5, R>), grid, dim3(128), 0, st, c, ix, V, K, ic); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename R>
hipError_t launch_project_ic(int p, R* c, const StateIx& ix, const double* V, int K, const IcP& ic, hipStream_t st) {
    if (ix.N == 1) return launch_ic_np<1, R>(p, c, ix, V, K, ic, st);
    return launch_ic_np<4, R>(p, c, ix, V, K, ic, st);
}
template hipError_t launch_project_ic<float>(int, float*, const StateIx&, const double*, int, const IcP&,
                                             hipStream_t);
template hipError_t launch_project_ic<double>(int, double*, const StateIx&, const double*, int, const IcP&,
                                              hipStream_t);

template <typename R>
hipError_t launch_audit(const R* c, const StateIx& ix, const double* Jd, int K, double b0, double gamma,
                        double* part, int nblocks, hipStream_t st) {
    if (ix.N == 1)
        hipLaunchKernelGGL((k_audit<1, R>), dim3(nblocks), dim3(256), 0, st, c, ix, Jd, K, b0, gamma, part);
    else
        hipLaunchKernelGGL((k_audit<4, R>), dim3(nblocks), dim3(256), 0, st, c, ix, Jd, K, b0, gamma, part);
    return hipGetLastError();
}
template hipError_t launch_audit<float>(const float*, const StateIx&, const double*, int, double, double, double*,
                                        int, hipStream_t);
template hipError_t launch_audit<double>(const double*, const StateIx&, const double*, int, double, double,
                                         double*, int, hipStream_t);

// ---------------------------------------------------------------------------
// State conversion on the device between the boundary's reference layout
// (FP64, c[var][mode][elem] in the caller's element order) and a state
// array of the bulk precision in layout `ix`.  perm[i] (Kp entries) is the
// caller's element of internal element i; padding elements carry element
// K-1's id.  One thread per (variable, internal element).
// ---------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(256) k_ref_to_state(const double* __restrict__ ref, R* __restrict__ dst,
                                                     const int* __restrict__ perm, StateIx ix, int K) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ix.N * ix.Kp) return;
    const int v = t / ix.Kp, i = t - v * ix.Kp;
    const size_t src = (size_t)perm[i];
    // every slot of the layout is written: padding modes hold zeros
    const int mp = ix.tiled == 0 ? ix.NP : ix.tiled == 2 ? 16 : 4 * tile_groups(ix.NP);
    for (int m = 0; m < mp; m++)
        dst[ix.at(v, m, i)] = m < ix.NP ? (R)ref[((size_t)v * ix.NP + m) * K + src] : R(0);
}

template <typename R>
__global__ void __launch_bounds__(256) k_state_to_ref(const R* __restrict__ src, double* __restrict__ ref,
                                                     const int* __restrict__ perm, StateIx ix, int K) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ix.N * K) return;
    const int v = t / K, i = t - v * K;
    const size_t dst = (size_t)perm[i];
    for (int m = 0; m < ix.NP; m++) ref[((size_t)v * ix.NP + m) * K + dst] = (double)src[ix.at(v, m, i)];
}

template <typename R>
hipError_t launch_ref_to_state(const double* ref, R* dst, const int* perm, const StateIx& ix, int K, hipStream_t st) {
    const int n = ix.N * ix.Kp;
    hipLaunchKernelGGL((k_ref_to_state<R>), dim3((n + 255) / 256), dim3(256), 0, st, ref, dst, perm, ix, K);
    return hipGetLastError();
}
template <typename R>
hipError_t launch_state_to_ref(const R* src, double* ref, const int* perm, const StateIx& ix, int K, hipStream_t st) {
    const int n = ix.N * K;
    hipLaunchKernelGGL((k_state_to_ref<R>), dim3((n + 255) / 256), dim3(256), 0, st, src, ref, perm, ix, K);
    return hipGetLastError();
}
template hipError_t launch_ref_to_state<float>(const double*, float*, const int*, const StateIx&, int, hipStream_t);
template hipError_t launch_ref_to_state<double>(const double*, double*, const int*, const StateIx&, int, hipStream_t);
template hipError_t launch_state_to_ref<float>(const float*, double*, const int*, const StateIx&, int, hipStream_t);
template hipError_t launch_state_to_ref<double>(const double*, double*, const int*, const StateIx&, int, hipStream_t);

// Fill the constant-memory tables of every order from the FP64 host tables.
template <typename R, int P>
static void fill_tab(TabP<R, P>& d) {
    Tables t = build_tables(P);
    constexpr int NP = TabP<R, P>::NP, NQ = TabP<R, P>::NQ, NQ1 = TabP<R, P>::NQ1;
    for (int m = 0; m < NP; m++) {
        for (int q = 0; q < NQ; q++) {
            d.B[m][q] = (R)t.basis[m * NQ + q];
            d.Gx[m][q] = (R)t.gx[m * NQ + q];
            d.Gy[m][q] = (R)t.gy[m * NQ + q];
        }
        for (int f = 0; f < 3; f++)
            for (int j = 0; j < NQ1; j++) d.S[f][m][j] = (R)t.side[(f * NP + m) * NQ1 + j];
        for (int v = 0; v < 3; v++) d.V[m][v] = (R)t.vertex[m * 3 + v];
    }
    for (int j = 0; j < NQ1; j++) d.W1[j] = (R)t.w1[j];
    for (int q = 0; q < NQ; q++) {
        d.R1[q] = (R)t.r1[q];
        d.R2[q] = (R)t.r2[q];
        d.W[q] = (R)t.w[q];
    }
}

template <typename R>
static void fill_all(TabAll<R>& a) {
    fill_tab<R, 0>(a.p0);
    fill_tab<R, 1>(a.p1);
    fill_tab<R, 2>(a.p2);
    fill_tab<R, 3>(a.p3);
    fill_tab<R, 4>(a.p4);
    fill_tab<R, 5>(a.p5);
}

hipError_t upload_tables() {
    static TabAll<double> h64;
    static TabAll<float> h32;
    fill_all<double>(h64);
    fill_all<float>(h32);
    hipError_t e = upload_tables_kernels(h64, h32);
    if (e == hipSuccess) e = upload_tables_mfma(h64, h32);
    if (e == hipSuccess) e = upload_tables_batch(h64, h32);
    return e;
}

}  // namespace dgmi

namespace dgmi {

// Streaming copy (16-byte vector loads and stores), the bandwidth reference
// for the roofline and the scale of the FETCH_SIZE / WRITE_SIZE counters.
// Streaming copy used as the bandwidth reference.  One 16-byte load and
// store per thread and iteration over a grid of 1024 workgroups: on MI355X
// this simple form measured faster than deeper unrolling or larger grids
// (about 5.6 TB/s for 1 GiB, read + write).
__global__ void __launch_bounds__(256) k_stream_copy(const float4* __restrict__ src, float4* __restrict__ dst,
                                                    int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// The same copy with four independent 16-byte loads in flight per thread
// before their stores (64 bytes per thread and iteration).
__global__ void __launch_bounds__(256) k_stream_copy4(const float4* __restrict__ src, float4* __restrict__ dst,
                                                     int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const float4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
        dst[i] = a;
        dst[i + stride] = b;
        dst[i + 2 * stride] = c;
        dst[i + 3 * stride] = d;
    }
    for (; i < n; i += stride) dst[i] = src[i];
}

// variant 0: k_stream_copy on 1024 workgroups; 1..3: k_stream_copy4 on
// 1024, 2048, 4096 workgroups
hipError_t launch_stream_copy(const void* src, void* dst, int64_t bytes, hipStream_t st, int variant) {
    const int64_t n = bytes / 16;
    if (variant == 0)
        hipLaunchKernelGGL(k_stream_copy, dim3(1024), dim3(256), 0, st, (const float4*)src, (float4*)dst, n);
    else
        hipLaunchKernelGGL(k_stream_copy4, dim3(512 << variant), dim3(256), 0, st, (const float4*)src, (float4*)dst, n);
    return hipGetLastError();
}

}  // namespace dgmi
