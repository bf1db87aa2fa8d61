/* run_dgcuda.c -- the body the reference's run_dgcuda
 * (src/2d/conservation/main.cu:8) takes on this library: read the mesh,
 * create the solver, project the initial condition, integrate to the end
 * time with CFL steps, write the density at the vertices as a gmsh view.
 *
 * The reference's problem translation units (euler/supersonic/supersonic.cu,
 * euler/cylinder/cylinder.cu, ...) keep their role: they supply the problem
 * through problem_desc() here, where they supplied U0, U_inflow, U_outflow
 * and reflecting_boundary as device functions, and their main() calls
 * run_dgcuda(argc, argv) as before.
 *
 *   cc -std=c99 -I<repo>/include run_dgcuda.c problem.c -L<repo>/dgcuda_amd -ldgmi
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dgmi.h"

/* supplied by the problem translation unit; returns the integrator
 * (DG_RK4 or DG_RK2, the reference's time_integrator global) */
int problem_desc(dg_problem_desc* prob);

static int check(int st) {
    if (st != DG_OK) fprintf(stderr, "dgmi: %s\n", dg_last_error());
    return st;
}

/* the reference's get_input (conserv.cu:386-444): -n order (0..5),
 * -T end time, the mesh file last */
static int get_input(int argc, char* argv[], int* n, double* endtime, const char** mesh) {
    int i;
    if (argc < 5) return 1;
    for (i = 0; i < argc; i++) {
        if (strcmp(argv[i], "-n") == 0 && i + 1 < argc) {
            *n = atoi(argv[i + 1]);
            if (*n < 0 || *n > 5) return 1;
        }
        if (strcmp(argv[i], "-T") == 0 && i + 1 < argc) {
            *endtime = atof(argv[i + 1]);
            if (*endtime < 0) return 1;
        }
    }
    *mesh = argv[argc - 1];
    return 0;
}

int run_dgcuda(int argc, char* argv[]) {
    int p = 1, K = 0, S = 0, nb = 0, scheme;
    double endtime = 1.0, t = 0.0, *uv;
    int64_t steps = 0;
    const char* mesh_file = NULL;
    dg_mesh* mesh = NULL;
    dg_solver* s = NULL;
    dg_problem_desc prob;

    if (get_input(argc, argv, &p, &endtime, &mesh_file)) {
        fprintf(stderr, "usage: %s -n <order 0..5> -T <end time> <mesh.pmsh>\n", argv[0]);
        return 1;
    }
    memset(&prob, 0, sizeof prob);
    scheme = problem_desc(&prob);
    if (check(dg_mesh_read_pmsh(mesh_file, &mesh))) return 1;
    if (check(dg_solver_create(mesh, &prob, p, DG_FP64, &s))) {
        dg_mesh_destroy(mesh);
        return 1;
    }
    /* the reference's CFL constant 0.7 (time_integrator.cu:263) */
    if (check(dg_project_ic(s)) || check(dg_integrate(s, endtime, 0.7, scheme, &steps, &t))) {
        dg_solver_destroy(s);
        dg_mesh_destroy(mesh);
        return 1;
    }
    printf("%lld steps to t = %.17g\n", (long long)steps, t);
    dg_mesh_info(mesh, &K, &S, &nb);
    uv = malloc(sizeof(double) * 3 * (size_t)K);
    /* density at the vertices (eval_u) and its view, as main.cu:187-209 writes output/U0.msh */
    if (!uv || check(dg_eval_vertices(s, 0, uv)) || check(dg_write_msh(mesh, 0, uv, "U0.msh"))) {
        free(uv);
        dg_solver_destroy(s);
        dg_mesh_destroy(mesh);
        return 1;
    }
    free(uv);
    dg_solver_destroy(s);
    dg_mesh_destroy(mesh);
    return 0;
}
