// Many independent solves of mixed size and order advanced together:
// persistent launches per RK stage (one per kernel class, or one in all)
// over device-side work lists.
#pragma once
#include "dg_solver.h"

namespace dgmi {

// Device side of a batch (slot table, work list, persistent launches, step
// graph), shared by Batch and the serving loop.
class BatchEngine {
   public:
    static Status create(int precision, int max_slots, int max_items, BatchEngine** out);
    virtual ~BatchEngine() {}
    // uploads the slot table and work list of these solves (stream-ordered)
    virtual Status set_slots(const std::vector<Solver*>& solvers) = 0;
    // Solves with at least this many elements (0: none) keep their slot
    // (per-solve dt and status) but run their own single-solve stage
    // kernels inside the captured step, after the batch launch.
    virtual void set_solo_threshold(int min_elems) = 0;
    // DG_BATCH_ONE_LAUNCH (default) or DG_BATCH_PER_CLASS (dgmi.h): which
    // persistent launches a stage runs.
    // Takes effect at the next set_slots.
    virtual void set_launch_form(int form) = 0;
    // enqueues one step (per-solve dt + 4 stages) for every resident solve
    virtual Status step(bool use_graph) = 0;
    // per slot: done flag, steps, t (waits for the engine's stream)
    virtual Status status(std::vector<double>& out) = 0;
    // every slot's control block, in one copy (waits for the engine's stream)
    virtual Status pull_ctrls(std::vector<Ctrl>& out) = 0;
    virtual Status sync() = 0;
    virtual hipStream_t stream() = 0;
    virtual int max_tile_halo() const = 0;
    // caps the persistent grid at max_groups workgroups (a multiple of 8, at
    // least 8; 0 = every resident workgroup)
    virtual Status set_max_groups(int max_groups) = 0;
    // out[0] workgroups of the largest launch, out[1] tile items, out[2]
    // lane items, out[3] longest run of consecutive tiles of one solve a
    // workgroup takes
    virtual void launch_info(int* out) const = 0;
};

// a solver set up for batched stepping (FP32: matrix-core tiling at every p)
Status make_batch_solver(const Mesh& m, const dg_problem_desc& prob, int p, int precision, Solver** out);

class Batch {
   public:
    static Status create(int precision, Batch** out);
    virtual ~Batch() {}
    virtual Status add(const Mesh& m, const dg_problem_desc& prob, int p, double t_end, int* slot) = 0;
    virtual Status project_ic() = 0;
    virtual Status run_steps(int nsteps, double cfl, bool use_graph) = 0;
    virtual Status integrate(double cfl, bool use_graph, int64_t* steps) = 0;
    virtual Status get_state(int slot, double* c_ref) = 0;
    virtual Status audit(int slot, dg_audit_out* out) = 0;
    virtual void info(int* n_solves, int64_t* total_elems, int64_t* device_bytes) const = 0;
    virtual Status set_max_groups(int max_groups) = 0;
    virtual void launch_info(int* out) const = 0;
    virtual void set_launch_form(int form) = 0;
};

}  // namespace dgmi
