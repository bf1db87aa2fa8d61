// Matrix-core (MFMA) processing of one 64-element tile, shared by the
// single-solve persistent kernel and the batched kernel.
//
// A workgroup of 4 wavefronts processes a tile of 64 consecutive (Morton)
// elements, one 16-element sub-tile per wavefront.  Per sub-tile and
// variable the element-local contractions are 16x16xK matrix products on
// v_mfma_f32_16x16x4f32 (FP32) or v_mfma_f64_16x16x4f64 (FP64) with elements
// as the N dimension:
//   volume interpolation  U[q][e]  = sum_m B[q][m]      u[m][e]      (K = 16)
//   face traces           T[pt][e] = sum_m S[pt][m]     u[m][e]      (K = 16)
//   volume projection     R[m][e]  = sum_q Gx[m][q] Fr[q][e] + Gy[m][q] Fs[q][e]  (K = 32)
//   surface projection    R[m][e] += sum_pt S[m][pt] w    G[pt][e]   (K = 16)
// where pt runs over the 3*(p+1) face points (padded to 16).  Each lane
// computes the volume fluxes of the points its accumulator registers hold;
// the reduction dimension of the projections is ordered the same way, so no
// data moves between lanes between the products.
//
// Neighbour traces come through LDS: every element's own face traces are
// written to LDS, and the tile's halo (neighbours outside the tile, listed
// per tile on the host) is traced into extra LDS slots.  The tile then
// evaluates the LLF flux of each of its sides once, spread evenly over its
// 256 lanes: a side between two of its elements is computed by one of them
// and written to both (the negation is exactly what the other would
// compute); a side to the halo is computed by the tile's element, and again
// by the neighbour's own tile, with identical operands, so the two
// contributions are bit-identical with opposite signs: race-free, no
// atomics, no face buffers in HBM.
#pragma once

#include "dg_kernels.h"

namespace dgmi {

typedef float floatx4 __attribute__((ext_vector_type(4)));

typedef double doublex4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ floatx4 mfma4(float a, float b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Per precision: the 16x16x4 product and where its operands and results
// sit.  Both take A[i][k] from lane i + 16k and B[k][j] from lane j + 16k.
// The FP32 product leaves D[4g + r][j] in register r of lane j + 16g, the
// FP64 product D[g + 4r][j]; pos(g, r) names that row.  Every reduction is
// ordered the same way (product s of lane group g takes index pos(g, s)),
// so a lane's B operands are the rows it holds in the output.  When at
// most 8 modes are in use, a mode step s whose modes pos(0..3, s) are all
// padding (FP64: modes 4s..4s+3) is left out: at p = 2 in FP64 that is
// two steps of four, 6.7% faster per stage; the one step of p = 3
// measured 2% slower left out, so p = 3 keeps it.
// Volume and face points are placed by the A tables of the interpolations:
// row i holds point qrow(i), so register r of lane group g holds point
// qpos(g, r) = g + 4r in both precisions.  Product step r of the
// projections then takes points 4r..4r+3, and a step whose points are all
// padding (4r >= the point count, p <= 3) is left out at compile time.
template <typename R>
struct MfmaT;
template <>
struct MfmaT<float> {
    using V = floatx4;
    __device__ __forceinline__ static V mma(float a, float b, V c) { return mfma4(a, b, c); }
    __host__ __device__ static constexpr int pos(int g, int r) { return 4 * g + r; }
    __host__ __device__ static constexpr int qpos(int g, int r) { return g + 4 * r; }
    __host__ __device__ static constexpr int qrow(int i) { return (i >> 2) + 4 * (i & 3); }
};
template <>
struct MfmaT<double> {
    using V = doublex4;
    __device__ __forceinline__ static V mma(double a, double b, V c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __host__ __device__ static constexpr int pos(int g, int r) { return g + 4 * r; }
    __host__ __device__ static constexpr int qpos(int g, int r) { return g + 4 * r; }
    __host__ __device__ static constexpr int qrow(int i) { return i; }
};

constexpr int MFMA_TILE = 64;

// whether mode step s of the mode reductions is issued (see above)
template <typename R, int NP>
__host__ __device__ constexpr bool mode_step_used(int s) {
    return NP > 8 || MfmaT<R>::pos(0, s) < NP;
}

// LDS of a tile: face records [64][3] of 4 values (nx, ny, len, packed
// word), the tile's side list (MFMA_TASK_BYTES: 192 entries of 8 bytes and
// 4 per-wavefront counts), then traces [(64 + hcap)][17*N] values (16
// points + 1 padding row so consecutive slots start in different banks), N
// variables innermost; `rs` = bytes per value (4 FP32, 8 FP64).
constexpr int MFMA_TASK_BYTES = 192 * 8 + 32;
__host__ __device__ constexpr size_t mfma_tile_lds_bytes(int N, int hcap, int rs = 4) {
    return (size_t)MFMA_TILE * 3 * 4 * rs + MFMA_TASK_BYTES + (size_t)(MFMA_TILE + hcap) * 17 * N * rs;
}

// the traces of a tile's LDS, after the face records `fd` and the side list
template <typename V, typename R>
__device__ __forceinline__ R* tile_traces(V* fd) {
    return reinterpret_cast<R*>(reinterpret_cast<unsigned char*>(fd + MFMA_TILE * 3) + MFMA_TASK_BYTES);
}

// State arrays are in the tile layout (StateIx, tiled): lane j + 16g of a
// wavefront reads the 4 modes pos(g, 0..3) of its element as one access of
// 4 values (16 bytes FP32, 32 bytes FP64).  Addresses are a wavefront-uniform
// base (scalar registers) plus one per-lane 32-bit index in units of 4
// values.
template <typename R>
__device__ __forceinline__ const DG_GLOBAL typename MfmaT<R>::V* gp4(const R* p) {
    return (const DG_GLOBAL typename MfmaT<R>::V*)p;
}
template <typename R>
__device__ __forceinline__ DG_GLOBAL typename MfmaT<R>::V* gp4(R* p) {
    return (DG_GLOBAL typename MfmaT<R>::V*)p;
}

// mode groups per variable of a 16-element block: ceil(n_p/4) in FP32, 4 in
// FP64 (a lane's modes g, g+4, g+8, g+12 span all groups)
template <typename R>
__host__ __device__ constexpr int tile_groups_of(int NP) {
    return sizeof(R) == 4 ? tile_groups(NP) : 4;
}

// stride (in units of 4 values) between the variables of one 16-element block
template <int P, typename R>
constexpr int tile_var_stride() { return 16 * tile_groups_of<R>(n_modes(P)); }

// index (in units of 4 values) of (variable 0, lane group g, element e);
// variable v adds v * tile_var_stride<P, R>().  Lanes of groups past the
// last (FP32, p <= 3) read the last group again; their operands meet zero
// matrix entries and they store nothing.
template <int N, int P, typename R>
__device__ __forceinline__ unsigned tile_lane4(int e, int g) {
    constexpr int G = tile_groups_of<R>(n_modes(P));
    const int gg = (G == 4) ? g : min(g, G - 1);
    return (unsigned)((e >> 4) * (N * 16 * G) + gg * 16 + (e & 15));
}

// per-lane A operands of the five products (lane layout: A[i][k] from lane
// i + 16k, B[k][j] from lane j + 16k, D[4(l/16)+r][l%16] in register r).
// The mode reductions (vol, tr) run mode 4g+s through product s in lane
// group g, so the B operand of a lane is the lane's own 4 consecutive modes
// (the state's 16-byte access) and the state in operand registers is
// already in the output layout of the projections.
template <int N, int P, typename R = float>
struct MfmaA {
    R vol[4], tr[4], sp[4], gx[4], gy[4];
    // `z` is zero at run time; a caller that loops over several instantiations
    // passes a loop-variant zero so the table reads stay inside its loop
    // instead of being hoisted (and held in registers) for every variant.
    __device__ __forceinline__ void init(int g, int j, int z = 0) {
        constexpr int NP = n_modes(P), NQ = n_quad_vol(P), NQ1 = P + 1, NPT = 3 * NQ1;
        const TabP<R, P>& T = *reinterpret_cast<const TabP<R, P>*>(reinterpret_cast<const R*>(&tab<R, P>()) + z);
        const int qj = MfmaT<R>::qrow(j);  // the point of A row j
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const int m = MfmaT<R>::pos(g, s);
            vol[s] = (m < NP && qj < NQ) ? T.B[m][qj] : R(0);
            tr[s] = (m < NP && qj < NPT) ? T.S[qj / NQ1][m][qj % NQ1] : R(0);
            const int q = MfmaT<R>::qpos(g, s);
            gx[s] = (j < NP && q < NQ) ? T.Gx[j][q] : R(0);
            gy[s] = (j < NP && q < NQ) ? T.Gy[j][q] : R(0);
            sp[s] = (j < NP && q < NPT) ? T.S[q / NQ1][j][q % NQ1] * T.W1[q % NQ1] : R(0);
        }
    }
};

// Slots of the per-lane A operands in a workgroup's LDS table [5][64] of
// 4 values (the pipelined kernel reads them where used instead of holding
// 20 registers).
enum { MFMA_A_VOL = 0, MFMA_A_TR = 1, MFMA_A_SP = 2, MFMA_A_GX = 3, MFMA_A_GY = 4, MFMA_A_SLOTS = 5 };

template <int N, int P, typename R>
__device__ __forceinline__ void mfma_a_to_lds(typename MfmaT<R>::V* tabA, int lane) {
    using V = typename MfmaT<R>::V;
    MfmaA<N, P, R> A;
    A.init(lane >> 4, lane & 15);
    tabA[MFMA_A_VOL * 64 + lane] = V{A.vol[0], A.vol[1], A.vol[2], A.vol[3]};
    tabA[MFMA_A_TR * 64 + lane] = V{A.tr[0], A.tr[1], A.tr[2], A.tr[3]};
    tabA[MFMA_A_SP * 64 + lane] = V{A.sp[0], A.sp[1], A.sp[2], A.sp[3]};
    tabA[MFMA_A_GX * 64 + lane] = V{A.gx[0], A.gx[1], A.gx[2], A.gx[3]};
    tabA[MFMA_A_GY * 64 + lane] = V{A.gy[0], A.gy[1], A.gy[2], A.gy[3]};
}

// State operands of one tile held by one wavefront: its own sub-tile, and
// its share of the halo, dealt as (halo sub-tile, variable) pairs
// round-robin to the 4 wavefronts.
template <int N, int P, typename R = float>
struct MfmaTileState {
    using V = typename MfmaT<R>::V;
    static constexpr int HPW = (4 * N + 3) / 4;  // most halo pairs per wave (hcap <= 64)
    static constexpr int VS = tile_var_stride<P, R>();
    V ub[N];
    V hb[HPW];
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
    __device__ __forceinline__ void load_data(const R* __restrict__ u, int K, int tile, int w, int g, int j,
                                              const int (&hid)[HPW]) {
        const unsigned l4 = tile_lane4<N, P, R>(min(tile * MFMA_TILE + w * 16 + j, K - 1), g);
#pragma unroll
        for (int v = 0; v < N; v++) ub[v] = gp4(u)[v * VS + l4];
#pragma unroll
        for (int q = 0; q < HPW; q++) hb[q] = gp4(u)[((w + 4 * q) % N) * VS + tile_lane4<N, P, R>(hid[q], g)];
    }
    __device__ __forceinline__ void load(const R* __restrict__ u, int K, const int* __restrict__ halo, int hcap,
                                         int tile, int w, int g, int j) {
        int hid[HPW];
        load_ids(halo, hcap, tile, w, j, hid);
        load_data(u, K, tile, w, g, j, hid);
    }
};

// Cell wave speed |u| + c of the new state's cell mean for the CFL step
// (euler.cu:252-278, in the bulk precision as the reference computes it):
// FP32 with the hardware reciprocal and square root, as the fluxes.
template <int N>
__device__ __forceinline__ float cell_speed_bulk(const PhysP& ph, const float* U) {
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
template <int N>
__device__ __forceinline__ double cell_speed_bulk(const PhysP& ph, const double* U) {
    return cell_speed<N>(ph, U);
}

template <int KIND>
constexpr bool rk_writes_acc() {
    return KIND == ST_RK4_2;
}

// Geometry of one tile from the packed arrays: the (face normal, length,
// packed neighbour word) record this thread stores to LDS (threads 0..191
// hold one of the 64 x 3 records; the others load record 191 again and
// store nothing, so the load is not under a branch), the affine map and
// the Jacobian of the lane's own element.  Padding elements replicate
// element K-1 in every geometry array.
template <typename R = float>
struct TileGeo {
    using V = typename MfmaT<R>::V;
    V f;
    R xr, yr, xs, ys, ij;  // ij: the inverse mass 1/J (M = J*I), precomputed in FP64
    __device__ __forceinline__ void load(const StageArgs<R>& a, const MfmaArgs& mx, int tile, int el) {
        f = ((const DG_GLOBAL V*)mx.face4)[(unsigned)(tile * (MFMA_TILE * 3) + min((int)threadIdx.x, MFMA_TILE * 3 - 1))];
        const unsigned ue = (unsigned)(tile * MFMA_TILE + el);
        const V m = ((const DG_GLOBAL V*)mx.metric4)[ue];
        xr = m[0];
        yr = m[1];
        xs = m[2];
        ys = m[3];
        ij = (gp(a.geo_e) + 4 * (size_t)a.Kp)[ue];
    }
};

// neighbour word of a face record (FP32: the bits of the 4th value, FP64:
// its value)
__device__ __forceinline__ int face_word(float x) { return __float_as_int(x); }
__device__ __forceinline__ int face_word(double x) { return (int)x; }

// LLF flux of one point of one side: face record `r`, the element's trace
// at the point `own` and the neighbour's `nbt`, in the orientation of the
// element; written over `own`, and its negation over `nbt` when the
// neighbour is an element of the tile.  For a boundary side, `nw` is the
// side's neighbour word (the ghost data's index) and jo the point.
template <int N, int NQ1, typename R, bool BND>
__device__ __forceinline__ void side_flux(const StageArgs<R>& a, const typename MfmaT<R>::V r, R* own, R* nbt,
                                          int nw, int jo) {
    using V = typename MfmaT<R>::V;
    const int wd = face_word(r[3]);
    const bool left = (wd >> 14) & 1;
    const bool inner = (wd & 0xFFF) < MFMA_TILE;
    // boundary side (the element is its left element, the slot its own): the
    // ghost state of the point replaces the neighbour trace, which is not
    // read.  Only the BND instantiation carries this code, so the kernels
    // of meshes without boundary sides keep their register allocation.
    const bool bside = BND && ((wd >> 15) & 1);
    if constexpr (BND)
        if (bside) nbt = own;
    R G[N];
    if constexpr (N == 4) {
        // F(UL).n + F(UR).n, max of the two speeds and UL - UR with the sign
        // of this side folded into the speed
        R uo[N], un[N], Fo[N], Fb[N], so, sb;
        const V qo = *reinterpret_cast<const V*>(own), qn = *reinterpret_cast<const V*>(nbt);
#pragma unroll
        for (int v = 0; v < N; v++) {
            uo[v] = qo[v];
            un[v] = qn[v];
        }
        if constexpr (BND)
            if (bside) boundary_ghost<N, NQ1, R>(a.bnd, nw, jo, uo, un);
        normal_flux_speed<R>(a.phys, uo, r[0], r[1], Fo, so);
        normal_flux_speed<R>(a.phys, un, r[0], r[1], Fb, sb);
        const R lm = fmax(so, sb);
        const R lms = left ? lm : -lm;
        const R cf = left ? R(-0.25) * r[2] : R(0.25) * r[2];
#pragma unroll
        for (int v = 0; v < N; v++) G[v] = cf * fma(lms, uo[v] - un[v], Fo[v] + Fb[v]);
        *reinterpret_cast<V*>(own) = V{G[0], G[1], G[2], G[3]};
        if (inner && !bside) *reinterpret_cast<V*>(nbt) = V{-G[0], -G[1], -G[2], -G[3]};
    } else {
        R UL[N], UR[N];
        {
            const R uo = own[0], un = nbt[0];
            UL[0] = left ? uo : un;
            UR[0] = left ? un : uo;
        }
        if constexpr (BND)
            if (bside) boundary_ghost<N, NQ1, R>(a.bnd, nw, jo, UL, UR);
        R fxl[N], fyl[N], fxr[N], fyr[N];
        flux<N, R>(a.phys, UL, fxl, fyl);
        flux<N, R>(a.phys, UR, fxr, fyr);
        const R lm = llf_speed<N, R>(a.phys, UL, UR, r[0], r[1]);
        const R coef = left ? R(-0.5) * r[2] : R(0.5) * r[2];
        G[0] = coef * (R(0.5) * ((fxl[0] + fxr[0]) * r[0] + (fyl[0] + fyr[0]) * r[1] + lm * (UL[0] - UR[0])));
        own[0] = G[0];
        if (inner && !bside) nbt[0] = -G[0];
    }
}

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
template <int N, int P, int KIND, typename R, bool BND = false>
__device__ __forceinline__ double mfma_process_tile_pipe(const StageArgs<R>& a, const MfmaArgs& mx,
                                                         const typename MfmaT<R>::V* tabA, MfmaTileState<N, P, R>& st,
                                                         int (&hid)[MfmaTileState<N, P, R>::HPW], TileGeo<R>& gc,
                                                         R dt, int tile, int next, int next2,
                                                         typename MfmaT<R>::V* fd, R* ts) {
    using M = MfmaT<R>;
    using V = typename M::V;
    constexpr int NP = n_modes(P), NQ = n_quad_vol(P), NQ1 = P + 1, NPT = 3 * NQ1;
    constexpr int TILE = MFMA_TILE;
    constexpr int TS = 17 * N;
    constexpr bool LAST = (KIND == ST_RK4_4 || KIND == ST_RK2_2);
    using RK = RkNeeds<KIND>;
    constexpr int HPW = MfmaTileState<N, P, R>::HPW;
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

    const R xr = gc.xr, yr = gc.yr, xs = gc.xs, ys = gc.ys;

    // ---- phase 1: face traces of the own sub-tile and of this wave's halo pairs ----
    // (points qpos(g, 0..3) of every lane are written: a slot holds 16 points
    // plus a padding row; points >= 3(p+1) are read by the surface
    // projection only to be replaced by zero)
    const V Atr = tabA[MFMA_A_TR * 64 + lane];
    {
        V T[N];
#pragma unroll
        for (int v = 0; v < N; v++) {
            T[v] = V{0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < 4; s++)
                if (mode_step_used<R, NP>(s)) T[v] = M::mma(Atr[s], st.ub[v][s], T[v]);
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            R* dst = ts + el * TS + M::qpos(g, i) * N;
            if constexpr (N == 4)
                *reinterpret_cast<V*>(dst) = V{T[0][i], T[1][i], T[2][i], T[3][i]};
            else
                dst[0] = T[0][i];
        }
    }
#pragma unroll
    for (int q = 0; q < HPW; q++) {
        const int pr = w + 4 * q;
        if (pr < npairs) {
            const int slot = TILE + (pr / N) * 16 + j, v = pr % N;
            V acc = {0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < 4; s++)
                if (mode_step_used<R, NP>(s)) acc = M::mma(Atr[s], st.hb[q][s], acc);
#pragma unroll
            for (int i = 0; i < 4; i++) ts[slot * TS + M::qpos(g, i) * N + v] = acc[i];
        }
    }
    // face records, and the side list: the thread of record t = 3 el + f of
    // a side this tile computes writes entry (word bits 16..23) of the list,
    // {offset of the element's trace at point 0 of the face | offset of the
    // neighbour's trace at its last point << 16, t}; per wavefront the count
    uint2* task = reinterpret_cast<uint2*>(fd + TILE * 3);
    int* ntask4 = reinterpret_cast<int*>(task + TILE * 3);
    {
        const int t = threadIdx.x;
        const int wd = face_word(gc.f[3]);
        const bool owner = t < TILE * 3 && (wd & FW_OWNER);
        if (t < TILE * 3) fd[t] = gc.f;
        if (owner) {
            const int eo = t / 3, f = t - 3 * eo;
            const int slot = wd & 0xFFF, fn = (wd >> 12) & 3;
            const unsigned oo = (unsigned)(eo * TS + f * NQ1 * N);
            const unsigned no = (unsigned)(slot * TS + (fn * NQ1 + NQ1 - 1) * N);
            task[(wd >> FW_TASK_SHIFT) & 0xFF] = make_uint2(oo | (no << 16), (unsigned)t);
        }
        const unsigned long long bm = __ballot(owner);
        if (lane == 0) ntask4[w] = __popcll(bm);
    }

    // ---- phase 2: volume, side fluxes, surface, inverse mass, RK update ----
    V U[N];
    const V Avol = tabA[MFMA_A_VOL * 64 + lane];
#pragma unroll
    for (int v = 0; v < N; v++) {
        V acc = {0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < 4; s++)
            if (mode_step_used<R, NP>(s)) acc = M::mma(Avol[s], st.ub[v][s], acc);
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
    constexpr int VS = tile_var_stride<P, R>();
    const unsigned l4 = tile_lane4<N, P, R>(ee, g);
    V cr[N], ar[N];
#pragma unroll
    for (int v = 0; v < N; v++) {
        cr[v] = V{0, 0, 0, 0};
        ar[v] = V{0, 0, 0, 0};
        if constexpr (RK::C_FROM_INPUT) cr[v] = st.ub[v];
        if constexpr (RK::ACC_FROM_INPUT) ar[v] = st.ub[v];
        if constexpr (sizeof(R) == 4) {
            if constexpr (RK::LOAD_C) cr[v] = __builtin_nontemporal_load(&gp4(a.c)[v * VS + l4]);
            if constexpr (RK::LOAD_ACC) ar[v] = __builtin_nontemporal_load(&gp4(a.acc)[v * VS + l4]);
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    st.load_data(a.u, K, next, w, g, j, hid);
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_setprio(0);
    V Rv[N];
#pragma unroll
    for (int v = 0; v < N; v++) Rv[v] = V{0, 0, 0, 0};
    const V Agx = tabA[MFMA_A_GX * 64 + lane], Agy = tabA[MFMA_A_GY * 64 + lane];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        // volume point q = qpos(g, i): flux, metric, reduction step i (steps
        // of padding points only are left out)
        if (4 * i >= NQ) continue;
        R Uq[N], fx[N], fy[N];
#pragma unroll
        for (int v = 0; v < N; v++) Uq[v] = U[v][i];
        const bool qv = M::qpos(g, i) < NQ;
        if (N == 4 && !qv) Uq[0] = R(1);  // padding point: keep the flux finite
        if constexpr (N == 4) {
            metric_flux<R>(a.phys, Uq, xr, yr, xs, ys, fx, fy);  // fx, fy hold Fr, Fs
#pragma unroll
            for (int v = 0; v < N; v++) {
                Rv[v] = M::mma(Agx[i], qv ? fx[v] : R(0), Rv[v]);
                Rv[v] = M::mma(Agy[i], qv ? fy[v] : R(0), Rv[v]);
            }
        } else {
            flux<N, R>(a.phys, Uq, fx, fy);
#pragma unroll
            for (int v = 0; v < N; v++) {
                const R fr = qv ? fx[v] * ys - fy[v] * xs : R(0);
                const R fs = qv ? fy[v] * xr - fx[v] * yr : R(0);
                Rv[v] = M::mma(Agx[i], fr, Rv[v]);
                Rv[v] = M::mma(Agy[i], fs, Rv[v]);
            }
        }
    }
    __syncthreads();

    // ---- phase 2b: the LLF flux of each side of the tile, once ----
    // (after the volume terms, which use only this wave's registers, so the
    // waves meet at the barrier with the next tile's loads long in flight)
    // Work item (side k of the list, point jo); the 64-item blocks of a
    // round go to the wavefronts in an order that rotates with the tile, so
    // the partial last round falls to each wavefront in turn.  The flux is
    // computed in the orientation of the record's element and written over
    // that element's trace at the point; a side between two elements of the
    // tile also writes its negation over the other element's trace, which is
    // bit for bit what that element would compute (the sums and maxima of the
    // two sides' terms commute and the sign flips are exact).  Each trace
    // entry of the tile's elements belongs to one side and one point, so the
    // item that reads it is the only one that writes it.
    {
        const int nitem = (ntask4[0] + ntask4[1] + ntask4[2]) * NQ1;
        const int wblk = ((w + tile) & 3) * 64;
        for (int r0 = 0; r0 < nitem; r0 += 256) {
            if (r0 + wblk >= nitem) continue;  // uniform over the wavefront
            const int it = r0 + wblk + lane;
            if (it < nitem) {
                const int k = it / NQ1, jo = it - k * NQ1;
                const uint2 en = task[k];
                int nw = 0;
                if constexpr (BND) {
                    const int eo = (int)en.y / 3, f = (int)en.y - 3 * eo;
                    nw = gp(a.nbr)[f * a.Kp + base + eo];
                }
                side_flux<N, NQ1, R, BND>(a, fd[en.y], ts + (en.x & 0xFFFF) + jo * N, ts + (en.x >> 16) - jo * N,
                                          nw, jo);
            }
        }
    }
    __syncthreads();

#pragma unroll
    for (int i = 0; i < 4; i++) {
        // own face point pt = qpos(g, i): its flux from phase 2b (point slots
        // past 3(p+1) hold stale traces and meet zero; steps of such slots
        // only are left out)
        if (4 * i >= NPT) continue;
        const int pt = M::qpos(g, i);
        const R* gq = ts + el * TS + pt * N;
        R G[N];
        if constexpr (N == 4) {
            const V q = *reinterpret_cast<const V*>(gq);
#pragma unroll
            for (int v = 0; v < N; v++) G[v] = pt < NPT ? q[v] : R(0);
        } else {
            G[0] = pt < NPT ? gq[0] : R(0);
        }
        const V Asp = tabA[MFMA_A_SP * 64 + lane];
#pragma unroll
        for (int v = 0; v < N; v++) Rv[v] = M::mma(Asp[i], G[v], Rv[v]);
    }

    // geometry of the next tile and halo ids of the one after, requested
    // before this tile's stores
    const R cj = gc.ij * dt;
    __builtin_amdgcn_s_setprio(2);
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (sizeof(R) == 8) {
        // FP64 requests c and acc here instead, after the fluxes: their 64
        // registers are then not live across the face fluxes, which lets two
        // workgroups share a CU (0.409 -> 0.323 ms per stage at 524,288
        // triangles, p = 4); FP32 measured no gain from the same move
#pragma unroll
        for (int v = 0; v < N; v++) {
            if constexpr (RK::LOAD_C) cr[v] = __builtin_nontemporal_load(&gp4(a.c)[v * VS + l4]);
            if constexpr (RK::LOAD_ACC) ar[v] = __builtin_nontemporal_load(&gp4(a.acc)[v * VS + l4]);
        }
    }
    gc.load(a, mx, next, el);
    MfmaTileState<N, P>::load_ids(mx.halo, mx.hcap, next2, w, j, hid);
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_setprio(0);

    // inverse mass and the RK stage update; lane holds modes m = pos(g, i)
    const unsigned s4 = tile_lane4<N, P, R>(e, g);
    R c0[N];
    double lam = 0.0;
#pragma unroll
    for (int v = 0; v < N; v++) {
        V o, an;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            R t = R(0);
            o[i] = rk_update<KIND, R>(cr[v][i], ar[v][i], cj * Rv[v][i], &t);
            an[i] = t;
        }
        // padding elements (e >= K) compute a copy of element K-1 into their own slots
        if (tile_groups_of<R>(NP) == 4 || g < tile_groups_of<R>(NP)) {
            if constexpr (rk_writes_acc<KIND>()) __builtin_nontemporal_store(an, &gp4(a.acc)[v * VS + s4]);
            __builtin_nontemporal_store(o, &gp4(a.out)[v * VS + s4]);
        }
        c0[v] = o[0];
    }
    if constexpr (LAST) {
        if (g == 0 && valid) {  // mode 0 = pos(0, 0) sits in register 0 of lane group 0
            R Uc[4];
            const R b0 = tab<R, P>().B[0][0];
#pragma unroll
            for (int v = 0; v < N; v++) Uc[v] = c0[v] * b0;
            lam = (double)cell_speed_bulk<N>(a.phys, Uc);
        }
    }
    __syncthreads();  // the next tile reuses the LDS
    return lam;
}

}  // namespace dgmi
