// One DG solve resident on the GPU.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/dgmi.h"
#include "dg_device.h"
#include "dg_geometry.h"
#include "dg_kernels.h"
#include "dg_mesh.h"

namespace dgmi {

struct Status {
    int code = DG_OK;
    std::string msg;
    static Status ok() { return Status(); }
    static Status err(int c, std::string m) {
        Status s;
        s.code = c;
        s.msg = std::move(m);
        return s;
    }
    explicit operator bool() const { return code == DG_OK; }
};

Status hip_status(hipError_t e, const char* what);
// Sets up the current device once per process (tables, the scratch address).
Status ensure_device_ready();

// Bytes of the device scratch behind the allocation-free right-hand side,
// and its address on device `dev` (null before the device is set up).
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
void* noalloc_scratch(int dev);

// Device memory and streams for solves, kept for reuse within the process:
// hipFree and hipStreamDestroy synchronise the whole device, which a serving
// loop that retires solves while others run cannot afford.  Blocks are
// reused best-fit (at most 1.5x the request); at most 16 GiB of idle blocks
// are kept per device.  Callers return memory only when no queued work uses
// it (each solver synchronises its own stream first).
Status pool_alloc(void** p, size_t bytes);
void pool_free(void* p);
Status pool_stream(hipStream_t* s);
void pool_release_stream(hipStream_t s);

PhysP make_phys(const dg_problem_desc& d);
IcP make_ic(const dg_problem_desc& d, double Lx, double Ly);
Status check_problem(const dg_problem_desc& d, int p, int precision);

// Kernel selection and launch shape of the stage kernels.
struct LaunchShape {
    int kind = 1;        // 1 = VALU (lane per element), 2 / 3 = matrix cores, 64 / 128-element tiles
    int sub = 0;         // matrix cores: 16-element sub-tiles per wavefront
    int tile_elems = 0;  // elements per workgroup
    int ntiles = 0;      // workgroups carrying elements
    int grid = 0;        // workgroups launched (multiple of 8 when remapped)
    int xcd_remap = 1;
};

// Device-side view of one solve, for the batched kernels.
struct SlotInfo {
    void *c, *acc, *ua, *ub, *geo_e, *geo_f;
    const int* nbr;
    const uint32_t* fword;
    const int* halo;
    const void* face4;      // packed geometry of the matrix-core kernels
    const void* metric4;
    Ctrl* ctrl;
    int K, Kp, N, p, hcap, kind;
    int boundary;           // the mesh has boundary sides (batched as a solo solve)
    PhysP phys;
};

class Solver {
   public:
    // process-unique, never reused (a new solve may reuse a freed one's address)
    const uint64_t uid = next_uid();
    static uint64_t next_uid() {
        static std::atomic<uint64_t> n{0};
        return ++n;
    }
    virtual ~Solver() {}
    virtual void slot_info(SlotInfo* out) = 0;
    // pushes the run parameters to the control block and computes the wave
    // speed of the resident state, then waits for the solver's stream
    virtual Status begin(double t_end, bool has_end, double cfl) = 0;
    // the same, enqueued on stream `on` without waiting (the caller's later
    // work on `on` sees the control block)
    virtual Status begin_on(double t_end, bool has_end, double cfl, hipStream_t on) = 0;
    // reads the control block back (t, steps) after work issued elsewhere
    virtual Status refresh() = 0;
    // takes a copy of the control block read back by the caller
    virtual void set_host_ctrl(const Ctrl& c) = 0;
    // the last stage of a step issued elsewhere left the wave speed in ctrl
    virtual void set_lambda_valid(bool v) = 0;
    virtual hipStream_t get_stream() const = 0;
    // enqueue RK4 stage `kind` (ST_RK4_1..4) of the current step with this
    // solve's own kernel on stream `st` (its dt is already in the control block)
    virtual Status stage_on(int kind, hipStream_t st) = 0;
    // kernel: 0 = automatic (timed by the autotuner), 1..3 = that kind
    static Status create(const Mesh& m, const dg_problem_desc& prob, int p, int precision, Solver** out,
                         int kernel = 0);

    int K = 0, S = 0, Kp = 0, p = 0, NP = 0, NQ = 0, NQ1 = 0, N = 0, precision = DG_FP64;
    double min_r = 0;
    double t = 0;
    int64_t steps = 0;
    dg_problem_desc prob{};

    virtual Status project_ic() = 0;
    virtual Status set_state(const double* c_ref) = 0;
    virtual Status get_state(double* c_ref) = 0;
    virtual Status rhs(const double* u_ref, double* k_ref, double dt) = 0;
    // rhs on device buffers (reference layout, FP64) through the device's
    // scratch: no allocation, enqueued on `on` (null: the solver's stream)
    virtual Status rhs_device_noalloc(const double* u_dev, double* k_dev, double dt, hipStream_t on) = 0;
    virtual Status rk_step(double dt, int scheme) = 0;
    virtual Status cfl_dt(double cfl, double* dt) = 0;
    virtual Status integrate(double t_end, double cfl, int scheme, int64_t* nsteps) = 0;
    virtual Status run_steps(int nsteps, double cfl, int scheme, bool use_graph) = 0;
    virtual Status audit(dg_audit_out* out) = 0;
    virtual Status eval_vertices(int var, double* uv) = 0;
    virtual Status time_stage_kernels(int nsteps, double cfl, double* out) = 0;
    // nsteps graph-replayed steps between one event pair: out[0] ms/step, out[1] total ms
    virtual Status time_step_graph(int nsteps, double cfl, double* out) = 0;
    virtual Status set_kernel(int kind) = 0;
    virtual int kernel_kind() const = 0;
    // caps the persistent matrix-core grid at max_groups workgroups (a
    // multiple of 8, at least 8; 0 = every resident workgroup)
    virtual Status set_max_groups(int max_groups) = 0;
    // out[0] kernel kind, out[1] tiles, out[2] workgroups launched,
    // out[3] most tiles one workgroup walks in a launch
    virtual void launch_info(int* out) const = 0;
    virtual int64_t device_bytes() const = 0;
    virtual Status sync() = 0;
};

}  // namespace dgmi
