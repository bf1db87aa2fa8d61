// Continuous-batching solve server: requests are admitted into resident
// slots as slots and device memory free up, and all resident solves advance
// together through the batched stage kernels.
#pragma once
#include "dg_solver.h"

namespace dgmi {

class Server {
   public:
    static Status create(const dg_server_config& cfg, Server** out);
    virtual ~Server() {}
    virtual Status submit(const dg_request& req, uint64_t* id) = 0;
    virtual Status poll(dg_result* out, int max, int* n) = 0;
    virtual Status get_state(uint64_t id, double* c_ref, int64_t n) = 0;
    virtual void keep_states(bool keep) = 0;
    virtual Status drain() = 0;
    virtual double now() = 0;
};

}  // namespace dgmi
