// Host-visible launchers of the HIP kernels.
#pragma once
#include "dg_device.h"

namespace dgmi {

// initial-condition parameters passed by value to the projection kernel;
// also the state of inflow and outflow boundary sides
struct IcP {
    int kind;   // 0 constant, 1 Gaussian pulse, 2 vortex superposition, 3 supersonic vortex, 4 (x - y)^alpha
    int nv;
    double gamma;
    double gx, gy, gk;
    double Lx, Ly;
    double rho_inf, u_inf, v_inf, p_inf;
    double beta[4], vx[4], vy[4];
    double constant[4];
    double mach;  // supersonic vortex: Mach number at r = 1
    int fpow;     // supersonic vortex: powers rounded through float
    int alpha;    // (x - y)^alpha
};

// Initial condition U0(x, y) in FP64 (conserv_kernels.cu:167-187,
// euler.cu:31-59, supersonic.cu:28-38), on the host and the device.
__host__ __device__ inline double wrap_d(double d, double L) { return (L > 0) ? d - L * rint(d / L) : d; }

// a^b, or with fpow the reference's powf on double arguments: arguments and
// result rounded to float (the power itself here in FP64)
__host__ __device__ inline double ic_pow(double a, double b, int fpow) {
    return fpow ? (double)(float)pow((double)(float)a, (double)(float)b) : pow(a, b);
}

__host__ __device__ inline void ic_point(const IcP& ic, double x, double y, double* U) {
    if (ic.kind == 0) {
        for (int n = 0; n < 4; n++) U[n] = ic.constant[n];
        return;
    }
    if (ic.kind == 1) {
        double dx = wrap_d(x - ic.gx, ic.Lx), dy = wrap_d(y - ic.gy, ic.Ly);
        U[0] = exp(-ic.gk * (dx * dx + dy * dy));
        return;
    }
    if (ic.kind == 4) {
        // the advection prototype's test function (test_all.cu:21-51)
        const double u = pow(x - y, (double)ic.alpha);
        for (int n = 0; n < 4; n++) U[n] = u;
        return;
    }
    if (ic.kind == 3) {
        // supersonic vortex between two arcs (supersonic.cu:28-38): density
        // (1 + (gamma-1)/2 M^2 (1 - 1/r^2))^(1/(gamma-1)) with the reference's
        // literals 1.0125 and 2.5, speed M / r along the arcs, p = rho^gamma / gamma
        const double r = sqrt(x * x + y * y);
        U[0] = ic_pow(1 + 1.0125 * (1. - 1. / (r * r)), 2.5, ic.fpow);
        U[1] = U[0] * sin(atan(y / x)) * ic.mach / r;
        U[2] = U[0] * -cos(atan(y / x)) * ic.mach / r;
        const double pr = (1.0 / ic.gamma) * ic_pow(U[0], ic.gamma, ic.fpow);
        U[3] = 0.5 * U[0] * (ic.mach * ic.mach / (r * r)) + pr * (1. / (ic.gamma - 1.));
        return;
    }
    const double PI = 3.14159265358979323846;
    const double g = ic.gamma;
    double du = 0, dv = 0, dT = 0;
    for (int k = 0; k < ic.nv; k++) {
        double dx = wrap_d(x - ic.vx[k], ic.Lx), dy = wrap_d(y - ic.vy[k], ic.Ly);
        double r2 = dx * dx + dy * dy;
        double ex = exp(0.5 * (1. - r2));
        double beta = ic.beta[k];
        du += -beta / (2. * PI) * dy * ex;
        dv += beta / (2. * PI) * dx * ex;
        dT += -(g - 1.) * beta * beta / (8. * g * PI * PI) * ex * ex;
    }
    double T_inf = ic.p_inf / ic.rho_inf;
    double T = T_inf + dT;
    double rho = ic.rho_inf * pow(T / T_inf, 1. / (g - 1.));
    double pp = ic.p_inf * pow(rho / ic.rho_inf, g);
    double uu = ic.u_inf + du, vv = ic.v_inf + dv;
    U[0] = rho;
    U[1] = rho * uu;
    U[2] = rho * vv;
    U[3] = pp / (g - 1.) + 0.5 * rho * (uu * uu + vv * vv);
}

hipError_t upload_tables();
hipError_t upload_tables_kernels(const TabAll<double>& h64, const TabAll<float>& h32);
hipError_t upload_tables_mfma(const TabAll<double>& h64, const TabAll<float>& h32);
hipError_t upload_tables_batch(const TabAll<double>& h64, const TabAll<float>& h32);
// streaming copy variants of the bandwidth probe (0..STREAM_COPY_VARIANTS-1)
constexpr int STREAM_COPY_VARIANTS = 4;
hipError_t launch_stream_copy(const void* src, void* dst, int64_t bytes, hipStream_t st, int variant = 0);

template <typename R>
hipError_t launch_stage_valu(int N, int p, const StageArgs<R>& a, dim3 grid, hipStream_t st);

// Extra per-solve data of the matrix-core kernel: per (element, face) a
// 32-bit word {bits 0..11 LDS slot of the neighbour (own tile element or
// halo entry), bits 12..13 neighbour's local face, bit 14 left element,
// bit 15 boundary side, bits 16..23 the side's index in its tile's side
// list, bit 24 set if the element computes the side's flux (FW_OWNER)},
// stored [3][Kp]; per tile the halo element list [ntiles][hcap]; and the
// packed geometry in the bulk precision: per (element, face) 4 values
// (nx, ny, len, word) [Kp][3], per element 4 values (xr, yr, xs, ys) [Kp],
// padding elements replicating element K-1 (without FW_OWNER).
// A tile computes each of its sides once: the boundary sides, the sides to
// its halo, and of the sides between two of its elements the one whose
// record has the left element.
constexpr int FW_TASK_SHIFT = 16;
constexpr uint32_t FW_OWNER = 1u << 24;
struct MfmaArgs {
    const uint32_t* fword;
    const int* halo;
    const void* face4;
    const void* metric4;
    int hcap;
    int pad;
};
template <typename R>
hipError_t launch_mfma_pack(const R* geo_e, const R* geo_f, const uint32_t* fword, int K, int Kp, void* face4,
                            void* metric4, hipStream_t st);

// matrix-core element kernel (FP32 or FP64, p = 1..4, state in the tile
// layout): 64-element tiles, 4 wavefronts per workgroup, persistent over the
// tiles of its XCD's range when the grid is smaller than the tile count;
// returns hipErrorInvalidValue for unsupported (N, p)
constexpr int MFMA_TILE_ELEMS = 64;
constexpr int MFMA_MAX_HALO = 64;
template <typename R>
hipError_t launch_stage_mfma(int N, int p, const StageArgs<R>& a, const MfmaArgs& mx, dim3 grid, hipStream_t st);
bool mfma_supported(int N, int p);
// dynamic LDS of a launch, `rs` = bytes per value of the bulk precision
size_t mfma_lds_bytes(int N, int hcap, int rs);
int mfma_resident_groups(int N, int p, int hcap, int rs, bool bnd);
hipError_t mfma_layout_probe(float* host_out);
hipError_t mfma_layout_probe_f64(double* host_out);
constexpr int VALU_TILE_ELEMS = 128;

// resets ctrl->lam_bits and accumulates the state's maximum cell wave speed into it
template <typename R>
hipError_t launch_lambda(const R* u, const StateIx& ix, int K, double b0, const PhysP& ph, Ctrl* ctrl,
                         hipStream_t st);
hipError_t launch_dt(Ctrl* ctrl, hipStream_t st);
// reference layout (FP64, caller's order) <-> state array in layout ix
template <typename R>
hipError_t launch_ref_to_state(const double* ref, R* dst, const int* perm, const StateIx& ix, int K, hipStream_t st);
template <typename R>
hipError_t launch_state_to_ref(const R* src, double* ref, const int* perm, const StateIx& ix, int K, hipStream_t st);
hipError_t launch_set_dt(Ctrl* ctrl, double dt, int advance, hipStream_t st);
template <typename R>
hipError_t launch_project_ic(int p, R* c, const StateIx& ix, const double* V, int K, const IcP& ic, hipStream_t st);
template <typename R>
hipError_t launch_audit(const R* c, const StateIx& ix, const double* Jd, int K, double b0, double gamma,
                        double* part, int nblocks, hipStream_t st);

// Ghost state UR of point j of a boundary side from the left trace UL
// (eval_left_right, conserv_kernels.cu:647-671): reflecting keeps rho and E
// and the velocity's component along the wall (supersonic.cu:64-108, wall
// normal precomputed per point); inflow and outflow take the problem's
// state at the point (euler.cu:177-241, precomputed in FP64).
template <int N, int NQ1, typename R>
__device__ __forceinline__ void boundary_ghost(const double* bnd, int w, int j, const R* UL, R* UR) {
    // the point's four values go to scalars first: indexing the pointer in a
    // loop made the compiler keep the ghost in a private (scratch) array
    const double* g = bnd + ((size_t)(w & NB_ELEM_MASK) * NQ1 + j) * 4;
    const double g0 = g[0], g1 = g[1], g2 = g[2], g3 = g[3];
    if (((w >> 28) & 3) == BC_REFLECTING) {
        UR[0] = UL[0];
        if constexpr (N == 4) {
            const R Nx = (R)g0, Ny = (R)g1;
            const R t = UL[1] * Ny - UL[2] * Nx;
            UR[1] = t * Ny;
            UR[2] = -t * Nx;
            UR[3] = UL[3];
        }
    } else {
        UR[0] = (R)g0;
        if constexpr (N == 4) {
            UR[1] = (R)g1;
            UR[2] = (R)g2;
            UR[3] = (R)g3;
        }
    }
}

}  // namespace dgmi
