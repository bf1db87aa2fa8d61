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

template <typename R>
hipError_t launch_stage_valu(int N, int p, const StageArgs<R>& a, dim3 grid, hipStream_t st);

// matrix-core element kernels (FP32, p = 3, 4); returns hipErrorInvalidValue
// for unsupported (N, p)
hipError_t launch_stage_mfma(int N, int p, const StageArgs<float>& a, dim3 grid, hipStream_t st);
bool mfma_supported(int N, int p);
// elements per workgroup of the matrix-core kernel
constexpr int MFMA_TILE_ELEMS = 64;
constexpr int VALU_TILE_ELEMS = 128;

template <typename R>
hipError_t launch_lambda(int N, const R* u, int K, int Kp, int NP, int tile_elems, int ntiles, int xcd_remap,
                         double b0, const PhysP& ph, double* lam_part, dim3 grid, hipStream_t st);
hipError_t launch_dt(Ctrl* ctrl, const double* lam_part, hipStream_t st);
hipError_t launch_set_dt(Ctrl* ctrl, double dt, int advance, hipStream_t st);
template <typename R>
hipError_t launch_project_ic(int N, int p, R* c, const double* V, int K, int Kp, const IcP& ic, hipStream_t st);
template <typename R>
hipError_t launch_audit(int N, const R* c, const double* Jd, int K, int Kp, int NP, double b0, double gamma,
                        double* part, int nblocks, hipStream_t st);

}  // namespace dgmi
