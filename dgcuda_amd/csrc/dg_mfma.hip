// Matrix-core (MFMA) element kernel for FP32 bulk at p = 3 and 4.
//
// A workgroup of 4 wavefronts processes tiles of 64 consecutive (Morton)
// elements, one 16-element sub-tile per wavefront.  The kernel is
// persistent: each workgroup walks a strided sequence of tiles inside the
// contiguous tile range of its XCD (so Morton neighbours share that XCD's
// L2), and while it computes one tile the state of its next tile and the RK
// registers of the current one are already in flight.  Per sub-tile and
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
#include "dg_kernels.h"

namespace dgmi {

DGMI_TABLE_UPLOADER(upload_tables_mfma)

typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ floatx4 mfma4(float a, float b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

bool mfma_supported(int N, int p) { return (N == 1 || N == 4) && (p == 3 || p == 4); }

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

// LDS layout: face records [64][3] float4 (nx, ny, len, packed word) then
// traces [(64 + hcap)][TS] floats, TS = 17*N (16 points + 1 padding row so
// consecutive slots start in different banks), N variables innermost.
constexpr int MFMA_TILE = 64;

template <int N, int P>
struct MfmaLane {
    static constexpr int NP = n_modes(P);
    // B operand of the state products: lane (g, j) holds u[v][4s+g][element j]
    __device__ __forceinline__ static void load_state(const float* __restrict__ u, int Kp, int e, int g,
                                                      float (&dst)[N][4]) {
#pragma unroll
        for (int v = 0; v < N; v++)
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const int m = 4 * s + g;
                dst[v][s] = (m < NP) ? u[(size_t)(v * NP + m) * Kp + e] : 0.f;
            }
    }
};

template <int N, int P>
__global__ void __launch_bounds__(256) k_stage_mfma(StageArgs<float> a, MfmaArgs mx) {
    constexpr int NP = n_modes(P), NQ = n_quad_vol(P), NQ1 = P + 1, NPT = 3 * NQ1;
    constexpr int TILE = MFMA_TILE;
    constexpr int TS = 17 * N;
    static_assert(NP <= 16 && NQ <= 16 && NPT <= 16, "order too high for 16x16 tiles");
    extern __shared__ float4 lds4[];
    float4* fd = lds4;                                      // [TILE*3]
    float* ts = reinterpret_cast<float*>(lds4 + TILE * 3);  // [(TILE+hcap)*TS]
    __shared__ double shmax[8];

    if (!a.ctrl->active) return;
    const int ntiles = a.ntiles, K = a.K, Kp = a.Kp;
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
    const int nhs = mx.hcap >> 4;  // halo sub-tiles (<= 4): wave w traces halo sub-tile w
    const TabP<float, P>& T = tab<float, P>();
    const int kind = a.kind;
    const bool last = (kind == ST_RK4_4 || kind == ST_RK2_2);
    const bool need_acc = (kind == ST_RK4_2 || kind == ST_RK4_3 || kind == ST_RK4_4);
    const float dt = (float)a.ctrl->dt;

    // per-lane A operands (see the layout comment of k_mfma_layout_probe)
    float Avol[4], Atr[4], Asp[4], Ax[4], Ay[4];
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const int m = 4 * s + g;
        Avol[s] = (m < NP && j < NQ) ? T.B[m][j] : 0.f;
        Atr[s] = (m < NP && j < NPT) ? T.S[j / NQ1][m][j % NQ1] : 0.f;
        const int q = 4 * g + s;
        Ax[s] = (j < NP && q < NQ) ? T.Gx[j][q] : 0.f;
        Ay[s] = (j < NP && q < NQ) ? T.Gy[j][q] : 0.f;
        Asp[s] = (j < NP && q < NPT) ? T.S[q / NQ1][j][q % NQ1] * T.W1[q % NQ1] : 0.f;
    }

    float ub[N][4], hb[N][4];
    if (t < tend) {
        MfmaLane<N, P>::load_state(a.u, Kp, min(t * TILE + w * 16 + j, K - 1), g, ub);
        if (w < nhs) MfmaLane<N, P>::load_state(a.u, Kp, mx.halo[(size_t)t * mx.hcap + w * 16 + j], g, hb);
    }
    double lam = 0.0;
    for (; t < tend; t += stride) {
        const int tn = t + stride;
        const int base = t * TILE;
        const int el = w * 16 + j;
        const int e = base + el;
        const bool valid = e < K;
        const int ee = valid ? e : K - 1;

        for (int i = threadIdx.x; i < TILE * 3; i += blockDim.x) {
            const int l = i / 3, f = i - 3 * l;
            const int e2 = min(base + l, K - 1);
            float4 r;
            r.x = a.geo_f[(3 * f + 0) * Kp + e2];
            r.y = a.geo_f[(3 * f + 1) * Kp + e2];
            r.z = a.geo_f[(3 * f + 2) * Kp + e2];
            r.w = __int_as_float((int)mx.fword[f * Kp + e2]);
            fd[i] = r;
        }
        const float xr = a.geo_e[ee], yr = a.geo_e[Kp + ee], xs = a.geo_e[2 * Kp + ee], ys = a.geo_e[3 * Kp + ee];
        const float Jv = a.geo_e[4 * Kp + ee];

        // ---- phase 1: face traces of the own sub-tile and of one halo sub-tile ----
#pragma unroll
        for (int v = 0; v < N; v++) {
            floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < 4; s++) acc = mfma4(Atr[s], ub[v][s], acc);
#pragma unroll
            for (int i = 0; i < 4; i++)
                if (4 * g + i < NPT) ts[el * TS + (4 * g + i) * N + v] = acc[i];
        }
        if (w < nhs) {
            const int slot = TILE + w * 16 + j;
#pragma unroll
            for (int v = 0; v < N; v++) {
                floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int s = 0; s < 4; s++) acc = mfma4(Atr[s], hb[v][s], acc);
#pragma unroll
                for (int i = 0; i < 4; i++)
                    if (4 * g + i < NPT) ts[slot * TS + (4 * g + i) * N + v] = acc[i];
            }
        }

        // ---- loads needed later: RK registers of this tile, state of the next ----
        float cc[N][4], aa[N][4];
#pragma unroll
        for (int v = 0; v < N; v++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int m = 4 * g + i;
                const size_t ix = (size_t)(v * NP + min(m, NP - 1)) * Kp + ee;
                cc[v][i] = (kind == ST_RHS) ? 0.f : a.c[ix];
                aa[v][i] = need_acc ? a.acc[ix] : 0.f;
            }
        float ubn[N][4], hbn[N][4];
        if (tn < tend) {
            MfmaLane<N, P>::load_state(a.u, Kp, min(tn * TILE + w * 16 + j, K - 1), g, ubn);
            if (w < nhs) MfmaLane<N, P>::load_state(a.u, Kp, mx.halo[(size_t)tn * mx.hcap + w * 16 + j], g, hbn);
        }
        __syncthreads();

        // ---- phase 2: volume, faces, inverse mass, RK update ----
        floatx4 U[N];
#pragma unroll
        for (int v = 0; v < N; v++) {
            floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < 4; s++) acc = mfma4(Avol[s], ub[v][s], acc);
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
                R[v] = mfma4(Ax[i], fr, R[v]);
                R[v] = mfma4(Ay[i], fs, R[v]);
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
            for (int v = 0; v < N; v++) R[v] = mfma4(Asp[i], G[v], R[v]);
        }

        // inverse mass and the RK stage update; lane holds modes m = 4g+i
        const float cj = 1.f / Jv * dt;
        float c0[N];
#pragma unroll
        for (int v = 0; v < N; v++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int m = 4 * g + i;
                if (m < NP && valid) {
                    const size_t ix = (size_t)(v * NP + m) * Kp + e;
                    const float k = cj * R[v][i];
                    const float c = cc[v][i];
                    float o;
                    if (kind == ST_RHS) {
                        o = k;
                    } else if (kind == ST_RK4_1) {
                        a.acc[ix] = k * (1.f / 6.f);
                        o = c + 0.5f * k;
                    } else if (kind == ST_RK4_2) {
                        a.acc[ix] = aa[v][i] + k * (1.f / 3.f);
                        o = c + 0.5f * k;
                    } else if (kind == ST_RK4_3) {
                        a.acc[ix] = aa[v][i] + k * (1.f / 3.f);
                        o = c + k;
                    } else if (kind == ST_RK4_4) {
                        o = c + (aa[v][i] + k * (1.f / 6.f));
                    } else if (kind == ST_RK2_1) {
                        o = c + 0.5f * k;
                    } else {
                        o = c + k;
                    }
                    a.out[ix] = o;
                    if (i == 0) c0[v] = o;
                }
            }
        if (last && g == 0 && valid) {
            double Uc[4];
            const double b0 = tab<double, P>().B[0][0];
#pragma unroll
            for (int v = 0; v < N; v++) Uc[v] = (double)c0[v] * b0;
            lam = fmax(lam, cell_speed<N>(a.phys, Uc));
        }
        __syncthreads();  // the next tile reuses the LDS
#pragma unroll
        for (int v = 0; v < N; v++)
#pragma unroll
            for (int s = 0; s < 4; s++) {
                ub[v][s] = ubn[v][s];
                hb[v][s] = hbn[v][s];
            }
    }
    if (last) {
        const double m2 = block_max(lam, shmax);
        if (threadIdx.x == 0) lam_commit(m2, a.lam_bits);
    }
}

size_t mfma_lds_bytes(int N, int hcap) {
    return (size_t)MFMA_TILE * 3 * sizeof(float4) + (size_t)(MFMA_TILE + hcap) * 17 * N * sizeof(float);
}

template <int N, int P>
static const void* mfma_fn() {
    return reinterpret_cast<const void*>(&k_stage_mfma<N, P>);
}

static const void* mfma_kernel(int N, int p) {
    if (N == 4 && p == 4) return mfma_fn<4, 4>();
    if (N == 4 && p == 3) return mfma_fn<4, 3>();
    if (N == 1 && p == 4) return mfma_fn<1, 4>();
    if (N == 1 && p == 3) return mfma_fn<1, 3>();
    return nullptr;
}

// workgroups of the matrix-core kernel that fit on the device at once
int mfma_resident_groups(int N, int p, int hcap) {
    const void* fn = mfma_kernel(N, p);
    if (!fn) return 0;
    int per_cu = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, mfma_lds_bytes(N, hcap)) != hipSuccess)
        return 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return per_cu * cus;
}

hipError_t launch_stage_mfma(int N, int p, const StageArgs<float>& a, const MfmaArgs& mx, dim3 grid,
                             hipStream_t st) {
    const size_t lds = mfma_lds_bytes(N, mx.hcap);
    if (N == 4 && p == 4)
        hipLaunchKernelGGL((k_stage_mfma<4, 4>), grid, dim3(256), lds, st, a, mx);
    else if (N == 4 && p == 3)
        hipLaunchKernelGGL((k_stage_mfma<4, 3>), grid, dim3(256), lds, st, a, mx);
    else if (N == 1 && p == 4)
        hipLaunchKernelGGL((k_stage_mfma<1, 4>), grid, dim3(256), lds, st, a, mx);
    else if (N == 1 && p == 3)
        hipLaunchKernelGGL((k_stage_mfma<1, 3>), grid, dim3(256), lds, st, a, mx);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace dgmi
