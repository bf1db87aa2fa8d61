/* supersonic.c -- the reference's supersonic-vortex problem translation unit
 * (src/2d/conservation/euler/supersonic/supersonic.cu) on this library: the
 * vortex between the arcs r = 1 and r = 1.384 at Mach 2.25 as initial, inflow
 * and outflow state, reflecting arcs about the origin, RK4. */
#include <string.h>

#include "dgmi.h"

int run_dgcuda(int argc, char* argv[]);

int problem_desc(dg_problem_desc* prob) {
    memset(prob, 0, sizeof *prob);
    prob->physics = DG_PHYS_EULER; /* local_N = 4 */
    prob->ic = DG_IC_SUPERSONIC;
    prob->gamma = 1.4;             /* get_GAMMA() */
    prob->mach = 2.25;
    return DG_RK4;                 /* time_integrator = RK4 */
}

int main(int argc, char* argv[]) { return run_dgcuda(argc, argv); }
