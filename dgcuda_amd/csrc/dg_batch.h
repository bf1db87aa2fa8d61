// Many independent solves of mixed size and order advanced together: one
// persistent launch per RK stage over a device-side work list.
#pragma once
#include "dg_solver.h"

namespace dgmi {

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
};

}  // namespace dgmi
