// Host-visible launchers of the HIP kernels.
#pragma once
#include "dg_device.h"

namespace dgmi {

// initial-condition parameters passed by value to the projection kernel
struct IcP {
    int kind;   // 0 constant, 1 Gaussian pulse, 2 vortex superposition
    int nv;
    double gamma;
    double gx, gy, gk;
    double Lx, Ly;
    double rho_inf, u_inf, v_inf, p_inf;
    double beta[4], vx[4], vy[4];
    double constant[4];
};

hipError_t upload_tables();
hipError_t upload_tables_kernels(const TabAll<double>& h64, const TabAll<float>& h32);
hipError_t upload_tables_mfma(const TabAll<double>& h64, const TabAll<float>& h32);
hipError_t upload_tables_batch(const TabAll<double>& h64, const TabAll<float>& h32);
hipError_t launch_stream_copy(const void* src, void* dst, int64_t bytes, hipStream_t st);

template <typename R>
hipError_t launch_stage_valu(int N, int p, const StageArgs<R>& a, dim3 grid, hipStream_t st);

// Extra per-solve data of the matrix-core kernel: per (element, face) a
// 16-bit word {bits 0..11 LDS slot of the neighbour (own tile element or
// halo entry), bits 12..13 neighbour's local face, bit 14 left element},
// stored [3][Kp]; per tile the halo element list [ntiles][hcap]; and the
// packed geometry in the bulk precision: per (element, face) 4 values
// (nx, ny, len, word) [Kp][3], per element 4 values (xr, yr, xs, ys) [Kp],
// padding elements replicating element K-1.
struct MfmaArgs {
    const uint16_t* fword;
    const int* halo;
    const void* face4;
    const void* metric4;
    int hcap;
    int pad;
};
template <typename R>
hipError_t launch_mfma_pack(const R* geo_e, const R* geo_f, const uint16_t* fword, int K, int Kp, void* face4,
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
int mfma_resident_groups(int N, int p, int hcap, int rs);
hipError_t mfma_layout_probe(float* host_out);
hipError_t mfma_layout_probe_f64(double* host_out);
constexpr int VALU_TILE_ELEMS = 128;

// resets ctrl->lam_bits and accumulates the state's maximum cell wave speed into it
template <typename R>
hipError_t launch_lambda(const R* u, const StateIx& ix, int K, double b0, const PhysP& ph, Ctrl* ctrl,
                         hipStream_t st);
hipError_t launch_dt(Ctrl* ctrl, hipStream_t st);
hipError_t launch_set_dt(Ctrl* ctrl, double dt, int advance, hipStream_t st);
template <typename R>
hipError_t launch_project_ic(int p, R* c, const StateIx& ix, const double* V, int K, const IcP& ic, hipStream_t st);
template <typename R>
hipError_t launch_audit(const R* c, const StateIx& ix, const double* Jd, int K, double b0, double gamma,
                        double* part, int nblocks, hipStream_t st);

}  // namespace dgmi
