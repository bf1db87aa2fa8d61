// Device-side data structures shared by the HIP kernels: per-order basis
// tables in constant memory, the per-solve control block, kernel argument
// blocks and the physics (Euler, linear advection) as device functions.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dg_tables.h"

namespace dgmi {

// ---------------------------------------------------------------------------
// Basis tables per order p, in the layouts the kernels index with
// compile-time subscripts.  B[m][q] = phi_m at volume point q; Gx/Gy[m][q] =
// w_q * dphi_m/dr, w_q * dphi_m/ds; S[f][m][j] = phi_m at point j of face f;
// W1[j] Gauss weight; R1/R2/W volume points and weights (for projection).
// ---------------------------------------------------------------------------
template <typename R, int P>
struct TabP {
    static constexpr int NP = n_modes(P), NQ = n_quad_vol(P), NQ1 = n_quad_face(P);
    R B[NP][NQ];
    R Gx[NP][NQ];
    R Gy[NP][NQ];
    R S[3][NP][NQ1];
    R W1[NQ1];
    R R1[NQ], R2[NQ], W[NQ];
    R V[NP][3];
};

template <typename R>
struct TabAll {
    TabP<R, 0> p0;
    TabP<R, 1> p1;
    TabP<R, 2> p2;
    TabP<R, 3> p3;
    TabP<R, 4> p4;
    TabP<R, 5> p5;
};

// Each kernel translation unit holds its own copy (code objects are not
// linked across units) and uploads it through its DGMI_TABLE_UPLOADER.
static __constant__ TabAll<double> c_tab64;
static __constant__ TabAll<float> c_tab32;

#define DGMI_TABLE_UPLOADER(fname)                                                        \
    hipError_t fname(const TabAll<double>& h64, const TabAll<float>& h32) {               \
        hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_tab64), &h64, sizeof(h64));          \
        if (e != hipSuccess) return e;                                                    \
        return hipMemcpyToSymbol(HIP_SYMBOL(c_tab32), &h32, sizeof(h32));                 \
    }

template <typename R, int P>
__device__ __forceinline__ const TabP<R, P>& tab();

#define DGMI_TAB_SPEC(RT, SYM, PP)                                              \
    template <>                                                                 \
    __device__ __forceinline__ const TabP<RT, PP>& tab<RT, PP>() { return SYM.p##PP; }
DGMI_TAB_SPEC(double, c_tab64, 0)
DGMI_TAB_SPEC(double, c_tab64, 1)
DGMI_TAB_SPEC(double, c_tab64, 2)
DGMI_TAB_SPEC(double, c_tab64, 3)
DGMI_TAB_SPEC(double, c_tab64, 4)
DGMI_TAB_SPEC(double, c_tab64, 5)
DGMI_TAB_SPEC(float, c_tab32, 0)
DGMI_TAB_SPEC(float, c_tab32, 1)
DGMI_TAB_SPEC(float, c_tab32, 2)
DGMI_TAB_SPEC(float, c_tab32, 3)
DGMI_TAB_SPEC(float, c_tab32, 4)
DGMI_TAB_SPEC(float, c_tab32, 5)
#undef DGMI_TAB_SPEC

// ---------------------------------------------------------------------------
// Per-solve control block in device memory: the time step is computed and
// consumed on the device, so a step needs no host round trip.
// ---------------------------------------------------------------------------
struct Ctrl {
    double dt;       // step size of the current step
    double t;        // simulated time after the current step's advance
    double t_end;    // end time (used when has_end)
    double cfl;      // CFL number
    double min_r;    // minimum inscribed-circle diameter of the mesh
    double two_p1;   // 2p + 1
    int64_t steps;   // steps taken
    // maximum cell wave speed of the current state, as the bit pattern of a
    // non-negative double (ordered like the value), accumulated with one
    // atomic max per workgroup; reset by the dt kernel after it is read
    unsigned long long lam_bits;
    int active;      // the stages of the current step run
    int done;        // t reached t_end
    int has_end;     // clamp to t_end
    int pad;
};

__device__ __forceinline__ void lam_commit(double mx, unsigned long long* bits) {
    if (mx > 0.0) atomicMax(bits, (unsigned long long)__double_as_longlong(mx));
}

// Start of a step on the device: dt = cfl * min_r / max lambda / (2p+1),
// clamped to t_end, t advanced before the stages (time_integrator.cu:254-270);
// reads the accumulated maximum wave speed and clears it for the step's last
// stage to fill.  A step not taken (t_end reached) leaves it: it still holds
// the wave speed of the current state, which a later run starts from.  One
// thread.
__device__ __forceinline__ void dt_update(Ctrl* ctrl) {
    if (ctrl->done) {
        ctrl->active = 0;
        return;
    }
    double t = ctrl->t;
    if (ctrl->has_end && !(t < ctrl->t_end)) {
        ctrl->active = 0;
        ctrl->done = 1;
        ctrl->dt = 0.0;
        return;
    }
    const double mx = __longlong_as_double((long long)ctrl->lam_bits);
    ctrl->lam_bits = 0ull;
    double dt = ctrl->cfl * ctrl->min_r / mx / ctrl->two_p1;
    if (ctrl->has_end && t + dt > ctrl->t_end) {
        dt = ctrl->t_end - t;
        t = ctrl->t_end;
    } else {
        t += dt;
    }
    ctrl->dt = dt;
    ctrl->t = t;
    ctrl->steps += 1;
    ctrl->active = 1;
}

struct PhysP {
    double gamma;
    double ax, ay;
    int quirks;      // DG_QUIRK_* bits
    int flux;        // advection: DG_FLUX_LLF (0) or DG_FLUX_CENTRAL (1, lambda = 0)
};

// Packed neighbour word per (element, face): bits 0..27 neighbour element,
// bits 28..29 neighbour's local face, bit 30 set when this element is the
// side's left element.  A boundary side (the element is its left element)
// has bit 31 set (a negative word), bits 28..29 the boundary kind (1
// reflecting, 2 outflow, 3 inflow: minus the reference's right_elem code)
// and bits 0..27 the index of the side in the boundary data.
constexpr int NB_ELEM_MASK = (1 << 28) - 1;
__host__ __device__ inline int nb_pack(int e, int f, int is_left) { return e | (f << 28) | (is_left << 30); }
__host__ __device__ inline int nb_pack_boundary(int bfi, int kind) {
    return (int)((unsigned)bfi | ((unsigned)kind << 28) | (1u << 30) | (1u << 31));
}
enum { BC_REFLECTING = 1, BC_OUTFLOW = 2, BC_INFLOW = 3 };
// Highest order whose lane kernel carries the boundary path.  The first
// run of the p = 4 FP64 instance with it faulted the GPU (memory aperture
// violation) and the cause is not found (a host build of the body checks
// every global access of that instance in bounds on the shipped boundary
// meshes, tests/test_lane_kernel_host.py), so p >= 4 keeps the kernel
// without it; boundary meshes at p = 4 run on the matrix-core kernels'
// boundary variant (Euler), and p = 5 is refused.  The host build raises it.
#ifndef DGMI_VALU_BOUNDARY_MAX_P
#define DGMI_VALU_BOUNDARY_MAX_P 3
#endif
constexpr int VALU_BOUNDARY_MAX_P = DGMI_VALU_BOUNDARY_MAX_P;

// k = dt M^-1 R(u) of the stage's input u; c = the step's base state u_n
enum StageKind {
    ST_RHS = 0,   // out = k
    ST_RK4_1 = 1, // u = c:  out = u2 = c + k/2
    ST_RK4_2 = 2, // u = u2: out = u3 = c + k/2; acc = W (rk_update)
    ST_RK4_3 = 3, // u = u3: out = u4 = c + k
    ST_RK4_4 = 4, // u = u4: c = acc + u4/3 + k/6; wave-speed partials
    ST_RK2_1 = 5, // out = c + k/2
    ST_RK2_2 = 6, // c = c + k; wave-speed partials
};

// Classical RK4 (time_integrator.cu:221) on every stage kernel, written
// with the stage states instead of the slopes.  With u2 = u_n + dt k1/2,
// u3 = u_n + dt k2/2 and u4 = u_n + dt k3, the slopes are dt k1 = 2(u2 - u_n),
// dt k3 = u4 - u_n, so
//   u_n+1 = u_n + dt (k1 + 2 k2 + 2 k3 + k4)/6 = W + u4/3 + dt k4/6,
//   W = (2 u_n + (u2 - u_n) + dt k2)/3 = 2 u_n/3 + dt k1/6 + dt k2/3.
// The stage that makes u3 also has u2 (its input) and u_n, and writes W to
// the accumulator; the last stage reads only its input u4 and W.  A step
// moves 12 state arrays (reads 1, 2, 2, 2; writes 1, 2, 1, 1) instead of 16
// with a running sum of the slopes.  The same RK4 polynomial; the sums are
// grouped differently from the reference's u_n + (k1/6 + k2/3 + k3/3 + k4/6),
// with the small terms summed first and one rounding per state-sized sum.
// Divisions by 3 of state-sized values are rounded to nearest (div3): a
// product with the rounded 1/3 would bias every step's state by 3e-8 in
// FP32, which shows as drift of the conserved totals.
// Per stage the lane holds two operands of the update, `c` and `aa`:
//   stage 1: c = input u_n                 -> u2
//   stage 2: c = u_n (read), aa = input u2  -> u3, and W to the accumulator
//   stage 3: c = u_n (read)                 -> u4
//   stage 4: c = input u4, aa = W (read)    -> u_n+1
template <int KIND>
struct RkNeeds {
    static constexpr bool C_FROM_INPUT = (KIND == ST_RK4_1 || KIND == ST_RK4_4 || KIND == ST_RK2_1);
    static constexpr bool ACC_FROM_INPUT = (KIND == ST_RK4_2);
    static constexpr bool LOAD_C = (KIND == ST_RK4_2 || KIND == ST_RK4_3 || KIND == ST_RK2_2);
    static constexpr bool LOAD_ACC = (KIND == ST_RK4_4);
};

// x / 3: the product with the rounded 1/3, corrected once with the exact
// residual x - 3q (one fused multiply-add), which rounds to nearest
template <typename R>
__device__ __forceinline__ R div3(R x) {
    const R r3 = R(1) / R(3);
    const R q = x * r3;
    return fma(fma(R(-3), q, x), r3, q);
}

template <int KIND, typename R>
__device__ __forceinline__ R rk_update(R c, R aa, R k, R* acc_out) {
    if constexpr (KIND == ST_RHS) {
        return k;
    } else if constexpr (KIND == ST_RK4_1 || KIND == ST_RK2_1) {
        return c + R(0.5) * k;
    } else if constexpr (KIND == ST_RK4_2) {
        *acc_out = div3(fma(R(2), c, (aa - c) + k));
        return c + R(0.5) * k;
    } else if constexpr (KIND == ST_RK4_3) {
        return c + k;
    } else if constexpr (KIND == ST_RK4_4) {
        return fma(k, R(1) / R(6), aa + div3(c));
    } else {
        return c + k;
    }
}

// Placement of entry (variable v, mode m, element e) of a state array in HBM.
//  * Mode-major (the lane-per-element kernels): (v*NP + m)*Kp + e, so the
//    lanes of a wavefront read consecutive elements of one mode.
//  * Tile layout of the matrix-core kernels: per 16-element block and
//    variable, groups g of [16 elements][4 modes] with lane j + 16g of a
//    wavefront holding its element's 4 modes in one access (the modes the
//    lane's matrix products leave in its registers):
//    - FP32 (tiled = 1): G = ceil(NP/4) groups, group g = modes 4g..4g+3,
//      (((e/16)*N + v)*G + m/4)*64 + (e%16)*4 + m%4;
//    - FP64 (tiled = 2): 4 groups, group g = modes g, g+4, g+8, g+12,
//      ((e/16)*N + v)*256 + (m%4)*64 + (e%16)*4 + m/4.
//    The wavefront's 64 lanes cover the block's variable contiguously.
//    Modes >= NP are padding and hold zeros.
constexpr int tile_groups(int NP) { return (NP + 3) / 4; }

struct StateIx {
    int N, NP, Kp, tiled;
    __host__ __device__ size_t at(int v, int m, int e) const {
        if (!tiled) return ((size_t)v * NP + m) * Kp + e;
        if (tiled == 2) return (((size_t)(e >> 4) * N + v) << 8) + (size_t)(((m & 3) << 6) + ((e & 15) << 2) + (m >> 2));
        const int G = tile_groups(NP);
        return ((((size_t)(e >> 4) * N + v) * G + (m >> 2)) << 6) + (size_t)((e & 15) << 2) + (m & 3);
    }
    // entries of one array: N*NP*Kp, or the tile layout's padded count
    __host__ __device__ size_t elems() const {
        if (!tiled) return (size_t)N * NP * Kp;
        return (size_t)N * (tiled == 2 ? 16 : 4 * tile_groups(NP)) * Kp;
    }
};

template <typename R>
struct StageArgs {
    const R* u;       // stage input [N][n_p][Kp] in the kernel's state layout (StateIx)
    const R* c;       // RK base state
    R* acc;           // RK accumulator
    R* out;           // next stage input, k, or c (in place) for the last stage
    const R* geo_e;   // [5][Kp]: xr, yr, xs, ys, J
    const R* geo_f;   // [9][Kp]: (nx, ny, len) for faces 0..2
    const int* nbr;   // [3][Kp]
    const Ctrl* ctrl;
    unsigned long long* lam_bits; // &ctrl->lam_bits (last stage)
    int K, Kp;        // elements, and the padded count (multiple of 64) of every [..][Kp] array
    int kind;
    int ntiles;       // tiles (workgroups) that carry elements
    int xcd_remap;    // map consecutive tiles to one XCD
    int pad;
    PhysP phys;
    // boundary sides, [side][point][4] FP64 (lane kernel only): the ghost
    // state of inflow / outflow points, the wall normal (Nx, Ny, 0, 0) of
    // reflecting points; null for a mesh without boundary sides
    const double* bnd;
};

// Workgroup -> tile mapping.  The dispatcher places workgroup b on XCD b % 8;
// with remapping, XCD x receives the contiguous tile range
// [x*per, (x+1)*per), so Morton neighbours share that XCD's L2.  The grid is
// launched with 8*per workgroups; tiles >= ntiles exit.
__device__ __forceinline__ int tile_of_block(int b, int nblocks, int remap) {
    if (!remap) return b;
    int per = nblocks >> 3;
    return (b & 7) * per + (b >> 3);
}

// ---------------------------------------------------------------------------
// Physics.  The formulas and their operation order are those of the
// reference's euler.cu:61-168 (pressure, eval_c, eval_flux, eval_lambda) and
// of the advection flux F = (ax u, ay u) with the LLF speed |a.n|.
// ---------------------------------------------------------------------------
// Pointers that arrive through memory (the batch slot table) are generic
// to the compiler, which then issues FLAT accesses with 64-bit per-lane
// addresses.  Reading and writing through a global-address-space view
// gives scalar-base global loads whatever the pointer's origin.
#ifndef DG_GLOBAL  // a host build of these headers (tests/host_check) defines it empty
#define DG_GLOBAL __attribute__((address_space(1)))
#endif
template <typename T>
__device__ __forceinline__ DG_GLOBAL T* gp(T* p) {
    return (DG_GLOBAL T*)p;
}

template <typename R>
__device__ __forceinline__ R rdiv(R a, R b);
template <>
__device__ __forceinline__ double rdiv<double>(double a, double b) { return a / b; }
template <>
__device__ __forceinline__ float rdiv<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

template <typename R>
__device__ __forceinline__ R qsqrt(R x, int float_sqrt);
template <>
__device__ __forceinline__ double qsqrt<double>(double x, int fq) {
    return fq ? (double)__builtin_sqrtf((float)x) : __builtin_sqrt(x);
}
template <>
__device__ __forceinline__ float qsqrt<float>(float x, int) { return __builtin_sqrtf(x); }

template <int N, typename R>
__device__ __forceinline__ void flux(const PhysP& ph, const R* U, R* fx, R* fy) {
    if constexpr (N == 1) {
        fx[0] = (R)ph.ax * U[0];
        fy[0] = (R)ph.ay * U[0];
    } else {
        const R rho = U[0], ru = U[1], rv = U[2], E = U[3];
        const R g = (R)ph.gamma;
        if constexpr (sizeof(R) == 8) {
            R p = (g - R(1)) * (E - (ru * ru + rv * rv) / R(2) / rho);
            fx[0] = ru;
            fy[0] = rv;
            fx[1] = ru * ru / rho + p;
            fy[1] = ru * rv / rho;
            fx[2] = ru * rv / rho;
            fy[2] = rv * rv / rho + p;
            fx[3] = ru * (E + p) / rho;
            fy[3] = rv * (E + p) / rho;
        } else {
            R ir = __builtin_amdgcn_rcpf(rho);
            R u = ru * ir, v = rv * ir;
            R p = (g - R(1)) * (E - R(0.5) * (ru * u + rv * v));
            fx[0] = ru;
            fy[0] = rv;
            fx[1] = ru * u + p;
            fy[1] = rv * u;
            fx[2] = ru * v;
            fy[2] = rv * v + p;
            fx[3] = (E + p) * u;
            fy[3] = (E + p) * v;
        }
    }
}

template <typename R>
__device__ __forceinline__ R sound_speed(const PhysP& ph, const R* U) {
    // rounded as written, like the oracle (no contraction into fused
    // multiply-adds): under DG_QUIRK_FLOAT_SQRT the float square root must
    // see the same double
#pragma clang fp contract(off)
    const R rho = U[0], ru = U[1], rv = U[2], E = U[3];
    const R g = (R)ph.gamma;
    R p = (g - R(1)) * (E - (ru * ru + rv * rv) / R(2) / rho);
    return qsqrt<R>(g * p / rho, ph.quirks & 1);
}

template <int N, typename R>
__device__ __forceinline__ R llf_speed(const PhysP& ph, const R* UL, const R* UR, R nx, R ny) {
    if constexpr (N == 1) {
        // central flux: no dissipation term (the advection prototype's
        // riemann = (u_L + u_R)/2, 2dadvec_kernels.cu:162-164)
        return ph.flux ? R(0) : fabs((R)ph.ax * nx + (R)ph.ay * ny);
    } else {
        if constexpr (sizeof(R) == 8) {
            R cl = sound_speed<R>(ph, UL), cr = sound_speed<R>(ph, UR);
            R ul = UL[1] / UL[0], vl = UL[2] / UL[0];
            R ur = UR[1] / UR[0], vr = UR[2] / UR[0];
            R sl = nx * ul + ny * vl;
            R sr = nx * ur + ny * vr;
            R lm = (sl > R(0)) ? sl + cl : -sl + cl;
            R rm = (sr > R(0)) ? sr + cr : -sr + cr;
            return (fabs(lm) > fabs(rm)) ? fabs(lm) : fabs(rm);
        } else {
            const R g = (R)ph.gamma;
            R il = __builtin_amdgcn_rcpf(UL[0]), ir = __builtin_amdgcn_rcpf(UR[0]);
            R ul = UL[1] * il, vl = UL[2] * il, ur = UR[1] * ir, vr = UR[2] * ir;
            R pl = (g - R(1)) * (UL[3] - R(0.5) * (UL[1] * ul + UL[2] * vl));
            R pr = (g - R(1)) * (UR[3] - R(0.5) * (UR[1] * ur + UR[2] * vr));
            // hardware square root (about 1 ulp): the FP32 path does not
            // reproduce the reference's rounding bit for bit
            R cl = __builtin_amdgcn_sqrtf(g * pl * il), cr = __builtin_amdgcn_sqrtf(g * pr * ir);
            R lm = fabsf(nx * ul + ny * vl) + cl;
            R rm = fabsf(nx * ur + ny * vr) + cr;
            return fmaxf(lm, rm);
        }
    }
}

// Reciprocal and the sound speed's square root of the normal-flux forms:
// FP32 the hardware instructions (about 1 ulp), FP64 IEEE division and
// square root (a float square root under DG_QUIRK_FLOAT_SQRT, euler.cu:80).
__device__ __forceinline__ float nf_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double nf_rcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float nf_sqrt(float x, int) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double nf_sqrt(double x, int quirks) { return qsqrt<double>(x, quirks & 1); }

// Euler face terms of one side: the normal flux F(U).n and the wave speed
// |u.n| + c of the local Lax-Friedrichs flux (euler.cu flux and
// eval_lambda).  F.n is formed directly (u.n first) instead of from the two
// Cartesian fluxes, with one reciprocal of rho: 17 operations per side
// where the reference's form divides by rho seven times.
template <typename R>
__device__ __forceinline__ void normal_flux_speed(const PhysP& ph, const R* U, R nx, R ny, R* F, R& spd) {
    const R g = (R)ph.gamma;
    const R rho = U[0], ru = U[1], rv = U[2], E = U[3];
    const R ir = nf_rcp(rho);
    const R u = ru * ir, v = rv * ir;
    const R p = (g - R(1)) * (E - R(0.5) * (ru * u + rv * v));
    const R un = u * nx + v * ny;
    F[0] = ru * nx + rv * ny;
    F[1] = ru * un + p * nx;
    F[2] = rv * un + p * ny;
    F[3] = (E + p) * un;
    if constexpr (sizeof(R) == 8) {
        // float-rounded sound speed (DG_QUIRK_FLOAT_SQRT): its argument in
        // the reference's order (eval_c, euler.cu:78-81), so the float
        // rounding of sqrtf sees the same double as the oracle's
        if (ph.quirks & 1) {
            spd = fabs(un) + sound_speed<R>(ph, U);
            return;
        }
    }
    spd = fabs(un) + nf_sqrt(g * p * ir, ph.quirks);
}

// Euler volume terms of one quadrature point: the fluxes contracted with the
// two rows of the element's inverse map, Fr = F.(ys, -xs) and
// Fs = F.(-yr, xr) (the F.n form of normal_flux_speed for both directions,
// sharing u, v, p and E + p).
template <typename R>
__device__ __forceinline__ void metric_flux(const PhysP& ph, const R* U, R xr, R yr, R xs, R ys, R* Fr, R* Fs) {
    const R g = (R)ph.gamma;
    const R rho = U[0], ru = U[1], rv = U[2], E = U[3];
    const R ir = nf_rcp(rho);
    const R u = ru * ir, v = rv * ir;
    const R p = (g - R(1)) * (E - R(0.5) * (ru * u + rv * v));
    const R Ep = E + p;
    const R u1 = u * ys - v * xs, u2 = v * xr - u * yr;
    Fr[0] = ru * ys - rv * xs;
    Fr[1] = ru * u1 + p * ys;
    Fr[2] = rv * u1 - p * xs;
    Fr[3] = Ep * u1;
    Fs[0] = rv * xr - ru * yr;
    Fs[1] = ru * u2 - p * yr;
    Fs[2] = rv * u2 + p * xr;
    Fs[3] = Ep * u2;
}

// CFL wave speed of one cell from its mean state (euler.cu:252-278), FP64.
template <int N>
__device__ __forceinline__ double cell_speed(const PhysP& ph, const double* U) {
    if constexpr (N == 1) {
        double a = ph.ax * ph.ax + ph.ay * ph.ay;
        return (ph.quirks & 1) ? (double)__builtin_sqrtf((float)a) : __builtin_sqrt(a);
    } else {
        double c = sound_speed<double>(ph, U);
        double m2 = U[1] * U[1] + U[2] * U[2];
        double s = ((ph.quirks & 1) ? (double)__builtin_sqrtf((float)m2) : __builtin_sqrt(m2)) / U[0];
        return (s > 0) ? s + c : -s + c;
    }
}

// Maximum over a workgroup of blockDim.x lanes (multiple of 64): wavefront
// butterfly with cross-lane shuffles, then one value per wave through LDS.
__device__ __forceinline__ double block_max(double v, double* sh) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double r = sh[0];
    for (int i = 1; i < nw; i++) r = fmax(r, sh[i]);
    return r;
}

}  // namespace dgmi
