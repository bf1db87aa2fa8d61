// Matrix-core (MFMA) processing of one 64-element tile, shared by the
// single-solve persistent kernel and the batched kernel.
//
// A workgroup of 4 wavefronts processes a tile of 64 consecutive (Morton)
// elements, one 16-element sub-tile per wavefront.  Per sub-tile and
// variable the element-local contractions are 16x16xK FP32 matrix products
// on v_mfma_f32_16x16x4f32 with elements as the N dimension:
//   volume interpolation  U[q][e]  = sum_m B[q][m]      u[m][e]      (K = 16)
//   face traces           T[pt][e] = sum_m S[pt][m]     u[m][e]      (K = 16)
//   volume projection     R[m][e]  = sum_q Gx[m][q] Fr[q][e] + Gy[m][q] Fs[q][e]  (K = 32)
//   surface projection    R[m][e] += sum_pt S[m][pt] w    G[pt][e]   (K = 16)
// where pt runs over the 3*(p+1) face points (padded to 16).  Each lane
// computes the fluxes of the points its accumulator registers hold; the
// reduction dimension of the projections is ordered the same way, so no
// data moves between lanes between the products.
//
// Neighbour traces come through LDS: every element's own face traces are
// written to LDS, and the tile's halo (neighbours outside the tile, listed
// per tile on the host) is traced into extra LDS slots.  Each face flux is
// evaluated by both of its elements in the side's canonical orientation with
// identical operands, so the two contributions are bit-identical with
// opposite signs: race-free, no atomics, no face buffers in HBM.
#pragma once
#include "dg_kernels.h"

namespace dgmi {

typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ floatx4 mfma4(float a, float b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

constexpr int MFMA_TILE = 64;

// LDS of a tile: face records [64][3] float4 (nx, ny, len, packed word) then
// traces [(64 + hcap)][17*N] floats (16 points + 1 padding row so
// consecutive slots start in different banks), N variables innermost.
__host__ __device__ constexpr size_t mfma_tile_lds_bytes(int N, int hcap) {
    return (size_t)MFMA_TILE * 3 * 16 + (size_t)(MFMA_TILE + hcap) * 17 * N * 4;
}

// Global addresses are split into a wavefront-uniform base (scalar
// registers) and one per-lane 32-bit unsigned offset, so loads take the
// scalar-base form and per-load address arithmetic stays off the VALU.
template <int N, int P>
__device__ __forceinline__ void load_state(const float* __restrict__ u, int Kp, unsigned lane_off, int g,
                                           float (&dst)[N][4]) {
    constexpr int NP = n_modes(P);
#pragma unroll
    for (int v = 0; v < N; v++)
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const DG_GLOBAL float* bp = gp(u) + (size_t)(v * NP + 4 * s) * Kp;  // uniform
            dst[v][s] = (4 * s + g < NP) ? bp[lane_off] : 0.f;   // lane_off = g*Kp + e
        }
}

// one variable of one halo sub-tile: the B operand of its trace product
template <int P>
__device__ __forceinline__ void load_state_var(const float* __restrict__ u, int Kp, unsigned lane_off, int g, int v,
                                               float (&dst)[4]) {
    constexpr int NP = n_modes(P);
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const DG_GLOBAL float* bp = gp(u) + (size_t)(v * NP + 4 * s) * Kp;
        dst[s] = (4 * s + g < NP) ? bp[lane_off] : 0.f;
    }
}

// per-lane A operands of the five products (lane layout: A[i][k] from lane
// i + 16k, B[k][j] from lane j + 16k, D[4(l/16)+r][l%16] in register r)
template <int N, int P>
struct MfmaA {
    float vol[4], tr[4], sp[4], gx[4], gy[4];
    // `z` is zero at run time; a caller that loops over several instantiations
    // passes a loop-variant zero so the table reads stay inside its loop
    // instead of being hoisted (and held in registers) for every variant.
    __device__ __forceinline__ void init(int g, int j, int z = 0) {
        constexpr int NP = n_modes(P), NQ = n_quad_vol(P), NQ1 = P + 1, NPT = 3 * NQ1;
        const TabP<float, P>& T =
            *reinterpret_cast<const TabP<float, P>*>(reinterpret_cast<const float*>(&tab<float, P>()) + z);
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const int m = 4 * s + g;
            vol[s] = (m < NP && j < NQ) ? T.B[m][j] : 0.f;
            tr[s] = (m < NP && j < NPT) ? T.S[j / NQ1][m][j % NQ1] : 0.f;
            const int q = 4 * g + s;
            gx[s] = (j < NP && q < NQ) ? T.Gx[j][q] : 0.f;
            gy[s] = (j < NP && q < NQ) ? T.Gy[j][q] : 0.f;
            sp[s] = (j < NP && q < NPT) ? T.S[q / NQ1][j][q % NQ1] * T.W1[q % NQ1] : 0.f;
        }
    }
};

// State operands of one tile held by one wavefront: its own sub-tile, and
// its share of the halo, dealt as (halo sub-tile, variable) pairs
// round-robin to the 4 wavefronts.
template <int N, int P>
struct MfmaTileState {
    static constexpr int HPW = (4 * N + 3) / 4;  // most halo pairs per wave (hcap <= 64)
    float ub[N][4];
    float hb[HPW][4];
    __device__ __forceinline__ void load(const float* __restrict__ u, int K, int Kp, const int* __restrict__ halo,
                                         int hcap, int tile, int w, int g, int j) {
        load_state<N, P>(u, Kp, (unsigned)(g * Kp + min(tile * MFMA_TILE + w * 16 + j, K - 1)), g, ub);
        const int npairs = (hcap >> 4) * N;
#pragma unroll
        for (int q = 0; q < HPW; q++) {
            const int pr = w + 4 * q;
            if (pr < npairs) {
                const int he = gp(halo)[(unsigned)(tile * hcap + (pr / N) * 16 + j)];
                load_state_var<P>(u, Kp, (unsigned)(g * Kp + he), g, pr % N, hb[q]);
            } else {
#pragma unroll
                for (int s = 0; s < 4; s++) hb[q][s] = 0.f;
            }
        }
    }
};

// Processes one tile (all 4 wavefronts of the workgroup call it).  Returns
// the lane's maximum cell wave speed of the new state for the last stage
// of a step (0 otherwise).  Contains two workgroup barriers.
template <int N, int P, int KIND>
__device__ __forceinline__ double mfma_process_tile(const StageArgs<float>& a, const MfmaArgs& mx,
                                                    const MfmaA<N, P>& A, const MfmaTileState<N, P>& st, int tile,
                                                    float4* fd, float* ts) {
    constexpr int NP = n_modes(P), NQ = n_quad_vol(P), NQ1 = P + 1, NPT = 3 * NQ1;
    constexpr int TILE = MFMA_TILE;
    constexpr int TS = 17 * N;
    constexpr bool LAST = (KIND == ST_RK4_4 || KIND == ST_RK2_2);
    constexpr bool NEED_ACC = (KIND == ST_RK4_2 || KIND == ST_RK4_3 || KIND == ST_RK4_4);
    constexpr bool NEED_C = (KIND != ST_RHS);
    constexpr int HPW = MfmaTileState<N, P>::HPW;
    static_assert(NP <= 16 && NQ <= 16 && NPT <= 16, "order too high for 16x16 tiles");
    const int K = a.K, Kp = a.Kp;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int g = lane >> 4, j = lane & 15;
    const int npairs = (mx.hcap >> 4) * N;
    const float dt = (float)a.ctrl->dt;
    const int base = tile * TILE;
    const int el = w * 16 + j;
    const int e = base + el;
    const bool valid = e < K;
    const int ee = valid ? e : K - 1;

    for (int i = threadIdx.x; i < TILE * 3; i += blockDim.x) {
        const int l = i / 3, f = i - 3 * l;
        const unsigned e2 = (unsigned)min(base + l, K - 1);
        const DG_GLOBAL float* gf = gp(a.geo_f) + (size_t)(3 * f) * Kp;
        float4 r;
        r.x = gf[e2];
        r.y = (gf + Kp)[e2];
        r.z = (gf + 2 * (size_t)Kp)[e2];
        r.w = __int_as_float((int)(gp(mx.fword) + (size_t)f * Kp)[e2]);
        fd[i] = r;
    }
    const unsigned ueu = (unsigned)ee;
    const DG_GLOBAL float* ge = gp(a.geo_e);
    const float xr = ge[ueu], yr = (ge + Kp)[ueu], xs = (ge + 2 * (size_t)Kp)[ueu];
    const float ys = (ge + 3 * (size_t)Kp)[ueu], Jv = (ge + 4 * (size_t)Kp)[ueu];

    // ---- phase 1: face traces of the own sub-tile and of this wave's halo pairs ----
#pragma unroll
    for (int v = 0; v < N; v++) {
        floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < 4; s++) acc = mfma4(A.tr[s], st.ub[v][s], acc);
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (4 * g + i < NPT) ts[el * TS + (4 * g + i) * N + v] = acc[i];
    }
#pragma unroll
    for (int q = 0; q < HPW; q++) {
        const int pr = w + 4 * q;
        if (pr < npairs) {
            const int slot = TILE + (pr / N) * 16 + j, v = pr % N;
            floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < 4; s++) acc = mfma4(A.tr[s], st.hb[q][s], acc);
#pragma unroll
            for (int i = 0; i < 4; i++)
                if (4 * g + i < NPT) ts[slot * TS + (4 * g + i) * N + v] = acc[i];
        }
    }
    __syncthreads();

    // ---- phase 2: volume, faces, inverse mass, RK update ----
    floatx4 U[N];
#pragma unroll
    for (int v = 0; v < N; v++) {
        floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < 4; s++) acc = mfma4(A.vol[s], st.ub[v][s], acc);
        U[v] = acc;
    }
    floatx4 R[N];
#pragma unroll
    for (int v = 0; v < N; v++) R[v] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < 4; i++) {
        // volume point q = 4g+i: flux, metric, reduction step i
        float Uq[N], fx[N], fy[N];
#pragma unroll
        for (int v = 0; v < N; v++) Uq[v] = U[v][i];
        const bool qv = (4 * g + i) < NQ;
        if (N == 4 && !qv) Uq[0] = 1.f;  // padding point: keep the flux finite
        flux<N, float>(a.phys, Uq, fx, fy);
#pragma unroll
        for (int v = 0; v < N; v++) {
            const float fr = qv ? fx[v] * ys - fy[v] * xs : 0.f;
            const float fs = qv ? fy[v] * xr - fx[v] * yr : 0.f;
            R[v] = mfma4(A.gx[i], fr, R[v]);
            R[v] = mfma4(A.gy[i], fs, R[v]);
        }
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        // own face point pt = 4g+i: LLF flux in the side's orientation
        const int pt = 4 * g + i;
        float G[N];
        if (pt < NPT) {
            const int f = pt / NQ1, jo = pt - f * NQ1;
            const float4 r = fd[el * 3 + f];
            const int wd = __float_as_int(r.w);
            const int slot = wd & 0xFFF, fn = (wd >> 12) & 3;
            const bool left = (wd >> 14) & 1;
            const float* own = ts + el * TS + pt * N;
            const float* nbt = ts + slot * TS + (fn * NQ1 + NQ1 - 1 - jo) * N;
            float UL[N], UR[N];
#pragma unroll
            for (int v = 0; v < N; v++) {
                const float uo = own[v], un = nbt[v];
                UL[v] = left ? uo : un;
                UR[v] = left ? un : uo;
            }
            float fxl[N], fyl[N], fxr[N], fyr[N];
            flux<N, float>(a.phys, UL, fxl, fyl);
            flux<N, float>(a.phys, UR, fxr, fyr);
            const float lm = llf_speed<N, float>(a.phys, UL, UR, r.x, r.y);
            const float coef = left ? -0.5f * r.z : 0.5f * r.z;
#pragma unroll
            for (int v = 0; v < N; v++)
                G[v] = coef * (0.5f * ((fxl[v] + fxr[v]) * r.x + (fyl[v] + fyr[v]) * r.y + lm * (UL[v] - UR[v])));
        } else {
#pragma unroll
            for (int v = 0; v < N; v++) G[v] = 0.f;
        }
#pragma unroll
        for (int v = 0; v < N; v++) R[v] = mfma4(A.sp[i], G[v], R[v]);
    }

    // inverse mass and the RK stage update; lane holds modes m = 4g+i
    const unsigned lane_rk = (unsigned)(4 * g * Kp + ee);
    const unsigned lane_st = (unsigned)(4 * g * Kp + e);
    const float cj = __builtin_amdgcn_rcpf(Jv) * dt;
    float c0[N];
    double lam = 0.0;
#pragma unroll
    for (int v = 0; v < N; v++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const bool mv = 4 * g + i < NP;
            const size_t uoff = (size_t)(v * NP + i) * Kp;
            const float c = (NEED_C && mv) ? (gp(a.c) + uoff)[lane_rk] : 0.f;
            const float aa = (NEED_ACC && mv) ? (gp(a.acc) + uoff)[lane_rk] : 0.f;
            if (mv && valid) {
                const float k = cj * R[v][i];
                float o;
                if constexpr (KIND == ST_RHS) {
                    o = k;
                } else if constexpr (KIND == ST_RK4_1) {
                    (gp(a.acc) + uoff)[lane_st] = k * (1.f / 6.f);
                    o = c + 0.5f * k;
                } else if constexpr (KIND == ST_RK4_2) {
                    (gp(a.acc) + uoff)[lane_st] = aa + k * (1.f / 3.f);
                    o = c + 0.5f * k;
                } else if constexpr (KIND == ST_RK4_3) {
                    (gp(a.acc) + uoff)[lane_st] = aa + k * (1.f / 3.f);
                    o = c + k;
                } else if constexpr (KIND == ST_RK4_4) {
                    o = c + (aa + k * (1.f / 6.f));
                } else if constexpr (KIND == ST_RK2_1) {
                    o = c + 0.5f * k;
                } else {
                    o = c + k;
                }
                (gp(a.out) + uoff)[lane_st] = o;
                if (i == 0) c0[v] = o;
            }
        }
    if constexpr (LAST) {
        if (g == 0 && valid) {
            double Uc[4];
            const double b0 = tab<double, P>().B[0][0];
#pragma unroll
            for (int v = 0; v < N; v++) Uc[v] = (double)c0[v] * b0;
            lam = cell_speed<N>(a.phys, Uc);
        }
    }
    __syncthreads();  // the next tile reuses the LDS
    return lam;
}

}  // namespace dgmi
