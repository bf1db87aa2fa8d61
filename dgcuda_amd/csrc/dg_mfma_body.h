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

// State arrays are in the tile layout (StateIx, tiled): lane j + 16g of a
// wavefront reads modes 4g..4g+3 of its element as one 16-byte access.
// Addresses are a wavefront-uniform base (scalar registers) plus one
// per-lane 32-bit float4 index.
__device__ __forceinline__ const DG_GLOBAL floatx4* gp4(const float* p) { return (const DG_GLOBAL floatx4*)p; }
__device__ __forceinline__ DG_GLOBAL floatx4* gp4(float* p) { return (DG_GLOBAL floatx4*)p; }

// float4 stride between the variables of one 16-element block
template <int P>
constexpr int tile_var_stride() { return 16 * tile_groups(n_modes(P)); }

// float4 index of (variable 0, modes 4g..4g+3, element e); variable v adds
// v * tile_var_stride<P>().  Lanes of groups past the last (p <= 3) read
// the last group again; their operands meet zero matrix entries and they
// store nothing.
template <int N, int P>
__device__ __forceinline__ unsigned tile_lane4(int e, int g) {
    constexpr int G = tile_groups(n_modes(P));
    const int gg = (G == 4) ? g : min(g, G - 1);
    return (unsigned)((e >> 4) * (N * 16 * G) + gg * 16 + (e & 15));
}

// per-lane A operands of the five products (lane layout: A[i][k] from lane
// i + 16k, B[k][j] from lane j + 16k, D[4(l/16)+r][l%16] in register r).
// The mode reductions (vol, tr) run mode 4g+s through product s in lane
// group g, so the B operand of a lane is the lane's own 4 consecutive modes
// (the state's 16-byte access) and the state in operand registers is
// already in the output layout of the projections.
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
            const int m = 4 * g + s;
            vol[s] = (m < NP && j < NQ) ? T.B[m][j] : 0.f;
            tr[s] = (m < NP && j < NPT) ? T.S[j / NQ1][m][j % NQ1] : 0.f;
            const int q = 4 * g + s;
            gx[s] = (j < NP && q < NQ) ? T.Gx[j][q] : 0.f;
            gy[s] = (j < NP && q < NQ) ? T.Gy[j][q] : 0.f;
            sp[s] = (j < NP && q < NPT) ? T.S[q / NQ1][j][q % NQ1] * T.W1[q % NQ1] : 0.f;
        }
    }
};

// Slots of the per-lane A operands in a workgroup's LDS table [5][64] of
// float4 (the pipelined kernel reads them where used instead of holding 20
// registers).
enum { MFMA_A_VOL = 0, MFMA_A_TR = 1, MFMA_A_SP = 2, MFMA_A_GX = 3, MFMA_A_GY = 4, MFMA_A_SLOTS = 5 };

template <int N, int P>
__device__ __forceinline__ void mfma_a_to_lds(float4* tabA, int lane) {
    MfmaA<N, P> A;
    A.init(lane >> 4, lane & 15);
    tabA[MFMA_A_VOL * 64 + lane] = make_float4(A.vol[0], A.vol[1], A.vol[2], A.vol[3]);
    tabA[MFMA_A_TR * 64 + lane] = make_float4(A.tr[0], A.tr[1], A.tr[2], A.tr[3]);
    tabA[MFMA_A_SP * 64 + lane] = make_float4(A.sp[0], A.sp[1], A.sp[2], A.sp[3]);
    tabA[MFMA_A_GX * 64 + lane] = make_float4(A.gx[0], A.gx[1], A.gx[2], A.gx[3]);
    tabA[MFMA_A_GY * 64 + lane] = make_float4(A.gy[0], A.gy[1], A.gy[2], A.gy[3]);
}

// State operands of one tile held by one wavefront: its own sub-tile, and
// its share of the halo, dealt as (halo sub-tile, variable) pairs
// round-robin to the 4 wavefronts.
template <int N, int P>
struct MfmaTileState {
    static constexpr int HPW = (4 * N + 3) / 4;  // most halo pairs per wave (hcap <= 64)
    static constexpr int VS = tile_var_stride<P>();
    floatx4 ub[N];
    floatx4 hb[HPW];
    // Halo element ids of this wave's pairs.  A halo state load depends on
    // its id, and waiting for an id waits for every load issued before it,
    // so the ids are requested ahead of the state loads that use them.
    // Ids and halo states are loaded for every slot q, pairs past the list
    // from its last entries (in bounds, unused): loads under a branch, with
    // registers zeroed on the other path, make the compiler wait for them
    // (and for everything issued before them) at the next write.
    __device__ __forceinline__ static void load_ids(const int* __restrict__ halo, int hcap, int tile, int w, int j,
                                                   int (&hid)[HPW]) {
        const int hlast = (hcap >> 4) - 1;
#pragma unroll
        for (int q = 0; q < HPW; q++) {
            const int hs = min((w + 4 * q) / N, hlast);
            hid[q] = gp(halo)[(unsigned)(tile * hcap + hs * 16 + j)];
        }
    }
    __device__ __forceinline__ void load_data(const float* __restrict__ u, int K, int tile, int w, int g, int j,
                                              const int (&hid)[HPW]) {
        const unsigned l4 = tile_lane4<N, P>(min(tile * MFMA_TILE + w * 16 + j, K - 1), g);
#pragma unroll
        for (int v = 0; v < N; v++) ub[v] = gp4(u)[v * VS + l4];
#pragma unroll
        for (int q = 0; q < HPW; q++) hb[q] = gp4(u)[((w + 4 * q) % N) * VS + tile_lane4<N, P>(hid[q], g)];
    }
    __device__ __forceinline__ void load(const float* __restrict__ u, int K, const int* __restrict__ halo, int hcap,
                                         int tile, int w, int g, int j) {
        int hid[HPW];
        load_ids(halo, hcap, tile, w, j, hid);
        load_data(u, K, tile, w, g, j, hid);
    }
};

// RK4 on the matrix-core path keeps acc = u_n + k1/6 + k2/3 + k3/3 (the
// base state folded in), so the last stage writes acc + k4/6 without
// reading u_n; classical RK4 with the FP32 sums grouped differently from
// the reference's c + (k1/6 + k2/3 + k3/3 + k4/6).
template <int KIND>
struct RkNeeds {
    static constexpr bool C_FROM_INPUT = (KIND == ST_RK4_1 || KIND == ST_RK2_1);
    static constexpr bool LOAD_C = (KIND == ST_RK4_2 || KIND == ST_RK4_3 || KIND == ST_RK2_2);
    static constexpr bool LOAD_ACC = (KIND == ST_RK4_2 || KIND == ST_RK4_3 || KIND == ST_RK4_4);
};

template <int KIND>
__device__ __forceinline__ float rk_update(float c, float aa, float k, float* acc_out) {
    if constexpr (KIND == ST_RHS) {
        return k;
    } else if constexpr (KIND == ST_RK4_1) {
        *acc_out = c + k * (1.f / 6.f);
        return c + 0.5f * k;
    } else if constexpr (KIND == ST_RK4_2) {
        *acc_out = aa + k * (1.f / 3.f);
        return c + 0.5f * k;
    } else if constexpr (KIND == ST_RK4_3) {
        *acc_out = aa + k * (1.f / 3.f);
        return c + k;
    } else if constexpr (KIND == ST_RK4_4) {
        return aa + k * (1.f / 6.f);
    } else if constexpr (KIND == ST_RK2_1) {
        return c + 0.5f * k;
    } else {
        return c + k;
    }
}
// Cell wave speed |u| + c of the new state's cell mean for the CFL step
// (euler.cu:252-278, in the bulk precision as the reference computes it):
// FP32 with the hardware reciprocal and square root, as the fluxes.
template <int N>
__device__ __forceinline__ float cell_speed_f32(const PhysP& ph, const float* U) {
    if constexpr (N == 1) {
        const float ax = (float)ph.ax, ay = (float)ph.ay;
        return __builtin_amdgcn_sqrtf(ax * ax + ay * ay);
    } else {
        const float g = (float)ph.gamma;
        const float ir = __builtin_amdgcn_rcpf(U[0]);
        const float m2 = U[1] * U[1] + U[2] * U[2];
        const float pr = (g - 1.f) * (U[3] - 0.5f * m2 * ir);
        return __builtin_amdgcn_sqrtf(m2) * ir + __builtin_amdgcn_sqrtf(g * pr * ir);
    }
}

template <int KIND>
constexpr bool rk_writes_acc() {
    return KIND == ST_RK4_1 || KIND == ST_RK4_2 || KIND == ST_RK4_3;
}

// Geometry of one tile from the packed arrays: the (face normal, length,
// packed neighbour word) record this thread stores to LDS (threads 0..191
// hold one of the 64 x 3 records; the others load record 191 again and
// store nothing, so the load is not under a branch), the affine map and
// the Jacobian of the lane's own element.  Padding elements replicate
// element K-1 in every geometry array.
struct TileGeo {
    float4 f;
    float xr, yr, xs, ys, J;
    __device__ __forceinline__ void load(const StageArgs<float>& a, const MfmaArgs& mx, int tile, int el) {
        const floatx4 r = ((const DG_GLOBAL floatx4*)mx.face4)[(unsigned)(tile * (MFMA_TILE * 3) +
                                                                            min((int)threadIdx.x, MFMA_TILE * 3 - 1))];
        f = make_float4(r[0], r[1], r[2], r[3]);
        const unsigned ue = (unsigned)(tile * MFMA_TILE + el);
        const floatx4 m = ((const DG_GLOBAL floatx4*)mx.metric4)[ue];
        xr = m[0];
        yr = m[1];
        xs = m[2];
        ys = m[3];
        J = (gp(a.geo_e) + 4 * (size_t)a.Kp)[ue];
    }
};

// Software-pipelined body of the persistent single-solve kernel.  On entry
// `st` and `gc` hold the state and geometry of `tile`, and `hid` the halo
// ids of `next`.  Once the volume interpolation has consumed the state, the
// tile's RK registers (c, acc) are requested and then the state of `next`
// into the same registers; after the face fluxes the geometry of `next` and
// the halo ids of `next2` are requested, ahead of this tile's stores.  The
// vector-memory counter retires in issue order and counts stores, so every
// load a later wait is for is issued before the stores that wait need not
// cover (the scheduling barriers keep that order).  `next` and `next2` are
// tiles of this launch (the caller repeats the last tile where the walk
// ends), so no load sits under a branch.  The A operands are read from the
// LDS table `tabA`.
template <int N, int P, int KIND>
__device__ __forceinline__ double mfma_process_tile_pipe(const StageArgs<float>& a, const MfmaArgs& mx,
                                                         const float4* tabA, MfmaTileState<N, P>& st,
                                                         int (&hid)[MfmaTileState<N, P>::HPW], TileGeo& gc,
                                                         float dt, int tile, int next, int next2, float4* fd,
                                                         float* ts) {
    constexpr int NP = n_modes(P), NQ = n_quad_vol(P), NQ1 = P + 1, NPT = 3 * NQ1;
    constexpr int TILE = MFMA_TILE;
    constexpr int TS = 17 * N;
    constexpr bool LAST = (KIND == ST_RK4_4 || KIND == ST_RK2_2);
    using RK = RkNeeds<KIND>;
    constexpr int HPW = MfmaTileState<N, P>::HPW;
    static_assert(NP <= 16 && NQ <= 16 && NPT <= 16, "order too high for 16x16 tiles");
    const int K = a.K;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int g = lane >> 4, j = lane & 15;
    const int npairs = (mx.hcap >> 4) * N;
    const int base = tile * TILE;
    const int el = w * 16 + j;
    const int e = base + el;
    const bool valid = e < K;
    const int ee = valid ? e : K - 1;

    const float xr = gc.xr, yr = gc.yr, xs = gc.xs, ys = gc.ys;

    // ---- phase 1: face traces of the own sub-tile and of this wave's halo pairs ----
    // (points 4g..4g+3 of every lane are written: a slot holds 16 points plus
    // a padding row, and points >= 3(p+1) are never read)
    const float4 Atr = tabA[MFMA_A_TR * 64 + lane];
    {
        floatx4 T[N];
#pragma unroll
        for (int v = 0; v < N; v++) {
            T[v] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < 4; s++) T[v] = mfma4(Atr[s], st.ub[v][s], T[v]);
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            float* dst = ts + el * TS + (4 * g + i) * N;
            if constexpr (N == 4)
                *reinterpret_cast<floatx4*>(dst) = floatx4{T[0][i], T[1][i], T[2][i], T[3][i]};
            else
                dst[0] = T[0][i];
        }
    }
#pragma unroll
    for (int q = 0; q < HPW; q++) {
        const int pr = w + 4 * q;
        if (pr < npairs) {
            const int slot = TILE + (pr / N) * 16 + j, v = pr % N;
            floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < 4; s++) acc = mfma4(Atr[s], st.hb[q][s], acc);
#pragma unroll
            for (int i = 0; i < 4; i++) ts[slot * TS + (4 * g + i) * N + v] = acc[i];
        }
    }
    if (threadIdx.x < TILE * 3) fd[threadIdx.x] = gc.f;
    __syncthreads();

    // ---- phase 2: volume, faces, inverse mass, RK update ----
    floatx4 U[N];
    const float4 Avol = tabA[MFMA_A_VOL * 64 + lane];
#pragma unroll
    for (int v = 0; v < N; v++) {
        floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < 4; s++) acc = mfma4(Avol[s], st.ub[v][s], acc);
        U[v] = acc;
    }
    // RK registers of this tile, then the next tile's state (st is dead
    // after this block).  For a stage whose input is the RK base state the
    // operand registers already hold it in the output layout.  c and acc are
    // read once per launch and the outputs are not read again in it: their
    // accesses carry the non-temporal hint, which keeps the 4 MB L2 of an
    // XCD for the stage input, whose lines neighbouring tiles read again as
    // halo (measured: 0.135 -> 0.122 ms per stage at 524,288 triangles, p = 4;
    // the hint on the geometry loads measured slower).
    // The wave runs at raised priority while it issues its loads, so a wave
    // with requests to make gets the issue slot before waves computing
    // (measured 1.5% faster per step at 524,288 triangles, p = 4).
    __builtin_amdgcn_s_setprio(2);
    constexpr int VS = tile_var_stride<P>();
    const unsigned l4 = tile_lane4<N, P>(ee, g);
    floatx4 cr[N], ar[N];
#pragma unroll
    for (int v = 0; v < N; v++) {
        cr[v] = floatx4{0.f, 0.f, 0.f, 0.f};
        ar[v] = floatx4{0.f, 0.f, 0.f, 0.f};
        if constexpr (RK::C_FROM_INPUT) cr[v] = st.ub[v];
        if constexpr (RK::LOAD_C) cr[v] = __builtin_nontemporal_load(&gp4(a.c)[v * VS + l4]);
        if constexpr (RK::LOAD_ACC) ar[v] = __builtin_nontemporal_load(&gp4(a.acc)[v * VS + l4]);
    }
    __builtin_amdgcn_sched_barrier(0);
    st.load_data(a.u, K, next, w, g, j, hid);
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_setprio(0);
    floatx4 R[N];
#pragma unroll
    for (int v = 0; v < N; v++) R[v] = floatx4{0.f, 0.f, 0.f, 0.f};
    const float4 Agx = tabA[MFMA_A_GX * 64 + lane], Agy = tabA[MFMA_A_GY * 64 + lane];
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
            R[v] = mfma4(Agx[i], fr, R[v]);
            R[v] = mfma4(Agy[i], fs, R[v]);
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
        const float4 Asp = tabA[MFMA_A_SP * 64 + lane];
#pragma unroll
        for (int v = 0; v < N; v++) R[v] = mfma4(Asp[i], G[v], R[v]);
    }

    // geometry of the next tile and halo ids of the one after, requested
    // before this tile's stores
    const float cj = __builtin_amdgcn_rcpf(gc.J) * dt;
    __builtin_amdgcn_s_setprio(2);
    __builtin_amdgcn_sched_barrier(0);
    gc.load(a, mx, next, el);
    MfmaTileState<N, P>::load_ids(mx.halo, mx.hcap, next2, w, j, hid);
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_setprio(0);

    // inverse mass and the RK stage update; lane holds modes m = 4g+i
    const unsigned s4 = tile_lane4<N, P>(e, g);
    float c0[N];
    double lam = 0.0;
#pragma unroll
    for (int v = 0; v < N; v++) {
        floatx4 o, an;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            float t = 0.f;
            o[i] = rk_update<KIND>(cr[v][i], ar[v][i], cj * R[v][i], &t);
            an[i] = t;
        }
        // padding elements (e >= K) compute a copy of element K-1 into their own slots
        if (tile_groups(NP) == 4 || g < tile_groups(NP)) {
            if constexpr (rk_writes_acc<KIND>()) __builtin_nontemporal_store(an, &gp4(a.acc)[v * VS + s4]);
            __builtin_nontemporal_store(o, &gp4(a.out)[v * VS + s4]);
        }
        c0[v] = o[0];
    }
    if constexpr (LAST) {
        if (g == 0 && valid) {
            float Uc[4];
            const float b0 = tab<float, P>().B[0][0];
#pragma unroll
            for (int v = 0; v < N; v++) Uc[v] = c0[v] * b0;
            lam = (double)cell_speed_f32<N>(a.phys, Uc);
        }
    }
    __syncthreads();  // the next tile reuses the LDS
    return lam;
}

}  // namespace dgmi
