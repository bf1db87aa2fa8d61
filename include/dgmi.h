/*
 * dgmi.h -- C ABI of the MI355X-native DG right-hand side and explicit RK step
 * for 2D advection and Euler on unstructured triangle meshes.
 *
 * Every entry point returns a dg_status (0 = DG_OK).  On failure the reason is
 * available from dg_last_error() (thread-local, valid until the next call on
 * the same thread).  No entry point calls exit() or aborts.
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   read_mesh            src/2d/conservation/conserv.cu:118   -> dg_mesh_read_pmsh / dg_mesh_from_arrays
 *   init_gpu             src/2d/conservation/conserv.cu:194   -> dg_solver_create
 *   set_quadrature       src/2d/conservation/conserv.cu:37    -> dg_solver_create (internal), dg_tables
 *   preval_basis         src/2d/basis.cu:150                  -> dg_solver_create (internal), dg_tables
 *   preval_* geometry    src/2d/conservation/conserv_kernels.cu:338-514, main.cu:84-149
 *                                                             -> dg_solver_create (internal), dg_mesh_geometry
 *   init_conditions      src/2d/conservation/conserv_kernels.cu:167 -> dg_project_ic
 *   eval_surface/eval_volume/eval_rhs
 *                        src/2d/conservation/conserv_kernels.cu:697,766; time_integrator.cu:130
 *                                                             -> dg_rhs
 *   rk_tempstorage/rk4/rk2 time_integrator.cu:194,212,445,458  -> dg_rk_step
 *   eval_global_lambda + host max + dt
 *                        src/2d/conservation/euler/euler.cu:252; time_integrator.cu:254-263
 *                                                             -> dg_cfl_dt
 *   time_integrate_rk4 / time_integrate_rk2
 *                        src/2d/conservation/time_integrator.cu:221,466 -> dg_integrate
 *   eval_u               src/2d/conservation/conserv_kernels.cu:818 -> dg_eval_vertices
 *   free_gpu             src/2d/conservation/conserv.cu:323   -> dg_solver_destroy
 *   run_dgcuda           src/2d/conservation/main.cu:8        -> the caller's own driver: read or generate
 *                        the mesh, dg_solver_create, dg_project_ic, dg_integrate (no
 *                        command-line driver is part of the library)
 *
 * State layout at the boundary ("reference layout"): c[var][mode][elem] in
 * the caller's element numbering, index num_elem*n_p*var + mode*num_elem + elem
 * (conserv_kernels.cu:183), always FP64 on the host side.
 */
#ifndef DGMI_H
#define DGMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    DG_OK = 0,
    DG_ERR_ARG = 1,         /* invalid argument */
    DG_ERR_IO = 2,          /* file could not be read or written */
    DG_ERR_PARSE = 3,       /* malformed mesh file */
    DG_ERR_HIP = 4,         /* HIP runtime error or no GPU */
    DG_ERR_UNSUPPORTED = 5, /* valid but not supported (e.g. p = 5 in a batch) */
    DG_ERR_NUMERIC = 6,     /* non-finite state or non-positive density / pressure */
    DG_ERR_BUDGET = 7,      /* request exceeds the memory budget */
    DG_ERR_STATE = 8        /* call out of order */
} dg_status;

/* physics: the value is the number of conserved variables (the reference's local_N) */
enum { DG_PHYS_ADVECTION = 1, DG_PHYS_EULER = 4 };
/* bulk arithmetic of the state and the kernels; geometry is always computed
 * in FP64 and stored in the bulk precision.  DG_GEOM_FP32 may be or'ed into
 * DG_FP64 for dg_solver_create: the geometric factors are then rounded
 * through FP32 before use, which isolates the share of an FP32 run's error
 * that FP32-stored geometry causes (a measurement aid; no counterpart in the
 * reference). */
enum { DG_FP64 = 0, DG_FP32 = 1, DG_GEOM_FP32 = 0x100 };
/* time integrators, numbered as the reference's RK4 / RK2 (conserv.cu:19-20) */
enum { DG_RK4 = 1, DG_RK2 = 2 };
/* initial conditions; DG_IC_SUPERSONIC is the supersonic vortex between two
 * arcs (euler/supersonic/supersonic.cu:28-38), also the state its inflow and
 * outflow boundaries impose.  The cylinder problem (euler/cylinder/cylinder.cu:
 * 26-111) is DG_IC_CONSTANT with the state (gamma, gamma M, 0,
 * gamma M^2/2 + 1/(gamma - 1)), which its inflow and outflow sides also
 * impose, and the wall centre (0.5, 0.5); the reference runs it with RK2. */
enum { DG_IC_CONSTANT = 0, DG_IC_GAUSSIAN = 1, DG_IC_VORTEX = 2, DG_IC_SUPERSONIC = 3, DG_IC_POLY = 4 };
/* DG_IC_POLY: u = (x - y)^poly_alpha in every variable, the advection
 * prototype's projection and time-step test function (src/2d/test_all.cu:
 * 21-51, 275-498): with a = (1, 1) it is a stationary solution, and its
 * value is also the state of inflow / outflow (exact Dirichlet) sides. */
/* numerical flux of advection: DG_FLUX_LLF the local Lax-Friedrichs flux
 * (lambda = |a.n|, upwind); DG_FLUX_CENTRAL the advection prototype's
 * central flux F(s).n with s = (u_L + u_R)/2 (2dadvec_kernels.cu:162-164,
 * 533-570), evaluated as the LLF form with lambda = 0.  Euler takes LLF. */
enum { DG_FLUX_LLF = 0, DG_FLUX_CENTRAL = 1 };
/* optional reproduction of the reference's float roundings and quirks:
 *  FLOAT_SQRT:     sound speed, cell wave speed and the reflecting wall's
 *                  radial length through a float square root (euler.cu:80,
 *                  269; supersonic.cu:99);
 *  FLOAT_GEOM:     side lengths, normals and inscribed diameters in float
 *                  (conserv_kernels.cu:346, 364-371, 434);
 *  FLOAT_POW:      the supersonic vortex's powers through powf
 *                  (supersonic.cu:31, 35);
 *  REFLECT_MIRROR: the reflecting boundary maps face 1's point j to the
 *                  mirrored point of the side (supersonic.cu:86-89), where
 *                  inflow and outflow use the trace's own point
 *                  (euler.cu:192-195). */
enum { DG_QUIRK_FLOAT_SQRT = 1, DG_QUIRK_FLOAT_GEOM = 2, DG_QUIRK_FLOAT_POW = 4, DG_QUIRK_REFLECT_MIRROR = 8 };

#define DG_MAX_VORTEX 4

/* Problem description (replaces the compile-time plug-in of the reference's
 * problem translation units, e.g. euler/supersonic/supersonic.cu:10-38). */
typedef struct {
    int physics;        /* DG_PHYS_ADVECTION or DG_PHYS_EULER */
    int ic;             /* DG_IC_* */
    int quirks;         /* DG_QUIRK_* bit set */
    int n_vortex;       /* DG_IC_VORTEX: number of superposed vortices (1..DG_MAX_VORTEX) */
    double gamma;       /* ratio of specific heats (Euler) */
    double ax, ay;      /* advection velocity */
    double gauss_x, gauss_y, gauss_k; /* DG_IC_GAUSSIAN: exp(-k |x - x_c|^2) */
    double rho_inf, u_inf, v_inf, p_inf; /* DG_IC_VORTEX background state */
    double vortex_beta[DG_MAX_VORTEX], vortex_x[DG_MAX_VORTEX], vortex_y[DG_MAX_VORTEX];
    double constant[4]; /* DG_IC_CONSTANT: conserved state */
    double mach;        /* DG_IC_SUPERSONIC: Mach number at the inner arc r = 1 (the reference's 2.25) */
    double wall_x, wall_y; /* reflecting sides: the wall normal at a face point is the direction from
                              this centre to the point (the supersonic vortex's arcs: (0, 0),
                              supersonic.cu:92-99; the cylinder: (0.5, 0.5), cylinder.cu:84-90) */
    int flux;           /* DG_FLUX_LLF or DG_FLUX_CENTRAL (advection only) */
    int poly_alpha;     /* DG_IC_POLY: the exponent alpha (0..8) */
} dg_problem_desc;

typedef struct {
    double mass[4];         /* integral of each conserved variable (FP64 sum) */
    double min_density;     /* minimum cell-mean density (Euler) or value (advection) */
    double min_pressure;    /* minimum cell-mean pressure (Euler; +inf for advection) */
    int64_t n_bad;          /* cells with non-finite or non-positive density / pressure */
    double t;               /* simulated time reached */
    int64_t steps;          /* RK steps taken since the last dg_project_ic / dg_set_state */
} dg_audit_out;

typedef struct dg_mesh dg_mesh;
typedef struct dg_solver dg_solver;
typedef struct dg_batch dg_batch;
typedef struct dg_server dg_server;

const char* dg_last_error(void);
/* library version string */
const char* dg_version(void);
/* number of visible GPUs (0 when none) */
int dg_device_count(void);

/* ---------------- meshes (host only) ---------------- */
/* reads the conservation-framework .pmsh dialect (conserv.cu:118-192) */
int dg_mesh_read_pmsh(const char* path, dg_mesh** out);
/* gmsh 2.2 ASCII .msh converted as the reference's genmesh.py (src/2d/genmesh.py:11-310)
 * builds a .pmsh: sides numbered as first met over (element, face 0..2), the
 * first element meeting a side is its left element, line elements tagged
 * 10000 / 20000 / 30000 make reflecting / outflow / inflow sides, boundary
 * sides sorted to the front.  flags = DG_GMSH_QUIRKS reproduces the converter
 * exactly, including its two latent bugs (a negative-Jacobian element swaps
 * shared vertex coordinates; the front-sort writes through a negative
 * element index); 0 swaps the element's own vertices instead, leaves
 * boundary sides' missing right elements alone and rejects untagged open
 * sides.  DG_GMSH_PERIODIC (not with DG_GMSH_QUIRKS: genmesh.py has no
 * periodic support) first pairs the open untagged sides on the edges of the
 * nodes' bounding box with their translates on the opposite edge, which sets
 * the mesh's period (dg_mesh_period); the partner side is removed and the
 * others keep their order.  Unknown flags: DG_ERR_ARG; an open side with
 * no tag and no partner: DG_ERR_PARSE. */
enum { DG_GMSH_QUIRKS = 1, DG_GMSH_PERIODIC = 2 };
int dg_mesh_read_gmsh(const char* path, int flags, dg_mesh** out);
/* reads the advection prototype's .pmsh dialect (a line with num_elem, then
 * V1x V1y V2x V2y V3x V3y per element) and builds its sides as the
 * prototype's read_mesh does (src/2d/2dadvec/2dadvec.cu:114-252): a clockwise
 * element has V1 and V2 swapped; faces V1V2, V2V3, V3V1 are matched with the
 * sides met so far by their end points' coordinates; an unmatched face makes
 * a new side with this element on its left; a matched one gets this element
 * on its right.  Faces met once are boundary sides; they get right element
 * `open_kind` (-1 reflecting, -2 outflow, -3 inflow): the prototype's
 * boundary sides take the exact solution (boundary_exact,
 * 2dadvec_kernels.cu:184), i.e. -3 with a stationary initial condition. */
int dg_mesh_read_vertex_pmsh(const char* path, int open_kind, dg_mesh** out);
/* the mesh's periodic extent in x and y (0 for a direction that does not
 * wrap); the initial conditions measure distances to the nearest image over
 * it.  No counterpart in the reference, whose meshes are not periodic. */
int dg_mesh_period(const dg_mesh* m, double* period_x, double* period_y);
/* builds a mesh from the arrays read_mesh returns, in its order:
 * V[6K] = V1x|V1y|V2x|V2y|V3x|V3y, elem_s[3K] = elem_s1|elem_s2|elem_s3,
 * sides[4S] = x1|y1|x2|y2, side_int[4S] = left_elem|right_elem|left_side_number|right_side_number.
 * period_x/period_y: periodic extent (0 if not periodic). */
int dg_mesh_from_arrays(int num_elem, int num_sides, const double* V, const int* elem_s, const double* sides,
                        const int* side_int, double period_x, double period_y, dg_mesh** out);
/* periodic n_x x n_y jittered-square mesh on [x0,x1]x[y0,y1], 2*n_x*n_y triangles
 * (no counterpart in the reference, which reads meshes made by gmsh) */
int dg_mesh_periodic_jittered(int nx, int ny, double x0, double x1, double y0, double y1, double jitter,
                              uint64_t seed, dg_mesh** out);
int dg_mesh_write_pmsh(const dg_mesh* m, const char* path);
int dg_mesh_info(const dg_mesh* m, int* num_elem, int* num_sides, int* num_boundary_sides);
/* the read_mesh arrays (any pointer may be NULL) */
int dg_mesh_get_arrays(const dg_mesh* m, double* V, int* elem_s, double* sides, int* side_int);
/* FP64 geometry as the reference precomputes it: geo_e[5K] = J|xr|yr|xs|ys,
 * geo_s[3S] = len|Nx|Ny, *min_r = minimum inscribed-circle diameter */
int dg_mesh_geometry(const dg_mesh* m, int quirks, double* geo_e, double* geo_s, double* min_r);
void dg_mesh_destroy(dg_mesh* m);

/* gmsh view of one variable's vertex values as the reference writes
 * output/U<var>.msh (conservation/main.cu:187-209): "View \"U<var> \" {",
 * one "ST (x1,y1,0,x2,y2,0,x3,y3,0) {u1,u2,u3};" line per element in the
 * mesh's order (%lf), "};".  uv[3*K] as dg_eval_vertices returns it. */
int dg_write_msh(const dg_mesh* m, int var, const double* uv, const char* path);
/* a named gmsh view as the reference's CPU Euler driver writes its output
 * (euler/cpu/main.c:201-263): "View \"<name>\" {", then per element in the
 * mesh's order a scalar line "ST (x1,y1,0,x2,y2,0,x3,y3,0) {f1,f2,f3};"
 * (ncomp = 1, vals[3*K]) or a vector line
 * "VT (x1,y1,0,...) {u1,v1,0,u2,v2,0,u3,v3,0};" (ncomp = 2, vals[6*K] =
 * u1,v1,u2,v2,u3,v3 per element), all numbers %lf, then "};".  The
 * reference's names are "Density ", "u ", "E " and, for its pressure file,
 * "E " again. */
int dg_write_msh_view(const dg_mesh* m, const char* name, int ncomp, const double* vals, const char* path);

/* L2 error of a state (reference layout, FP64, the mesh's element order)
 * against the problem's initial condition U0, per variable:
 * err[v] = sqrt(sum_K J_K sum_q w_q (u_h(x_q) - U0(x_q))^2) with the volume
 * rule of order min(p + 2, 5) (degree 2 min(p + 2, 5)).  The exact solution where U0 is steady (the
 * supersonic vortex) or after whole periods of a periodic flow; the
 * reference's eval_error is a stub.  Host only. */
int dg_l2_error(const dg_mesh* m, const dg_problem_desc* prob, int p, const double* c_ref, double* err);

/* FP64 basis and quadrature tables for order p in the reference's preval_basis
 * layouts (basis.cu:150-226): r1,r2,w[n_q]; s_r,w1[n_q1]; basis,gx,gy[n_p*n_q];
 * side[3*n_p*n_q1]; vertex[3*n_p].  sizes[0..2] = n_p, n_q, n_q1. */
int dg_tables(int p, int* sizes, double* r1, double* r2, double* w, double* s_r, double* w1, double* basis,
              double* gx, double* gy, double* side, double* vertex);

/* ---------------- one solve on the GPU ---------------- */
/* p in 0..5 as the reference's -n option; precision DG_FP64 / DG_FP32.
 * Boundary sides (right element -1 reflecting, -2 outflow, -3 inflow, as the
 * reference's eval_left_right, conserv_kernels.cu:647-671) take ghost states:
 * inflow and outflow the problem's initial state at the face point
 * (euler.cu:177-241), reflecting the interior density and energy and the
 * velocity's component along the wall whose normal is the radial direction
 * (supersonic.cu:64-108).  Such meshes run on the lane-per-element kernel
 * up to p = 3 and on the matrix-core kernels up to p = 4 (advection and
 * Euler); p = 5 on them returns DG_ERR_UNSUPPORTED. */
int dg_solver_create(const dg_mesh* m, const dg_problem_desc* prob, int p, int precision, dg_solver** out);
void dg_solver_destroy(dg_solver* s);
/* sizes[0..5] = num_elem, num_sides, n_p, n_q, n_q1, N */
int dg_solver_info(const dg_solver* s, int* sizes, double* min_r);
/* L2 projection of the problem's initial condition (init_conditions); resets t and steps */
int dg_project_ic(dg_solver* s);
/* state I/O in the reference layout, FP64 host buffers of N*n_p*num_elem; set resets t and steps */
int dg_set_state(dg_solver* s, const double* c_ref);
int dg_get_state(dg_solver* s, double* c_ref);
/* k = dt * M^-1 R(u) exactly as eval_surface + eval_volume + eval_rhs compute it
 * (M = J*I); u and k are host buffers in the reference layout; t is kept for
 * the reference's signature (the boundary states do not depend on it) */
int dg_rhs(dg_solver* s, const double* u_ref, double* k_ref, double t, double dt);
/* dg_rhs for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc); no
 * counterpart in the reference.  u_dev and k_dev are the caller's device
 * buffers of N*n_p*num_elem doubles in the reference layout; k = dt*M^-1 R(u)
 * as dg_rhs computes it, with the same kernels.  The state in between goes
 * through a fixed 64 MiB scratch array per device (two arrays of the
 * solve's state layout: up to about 65,000 triangles for FP64 Euler at
 * p = 4, 131,000 for FP32); a larger solve returns DG_ERR_BUDGET before any
 * work.  No allocation and no wait: the work is enqueued on `stream` (a
 * hipStream_t; NULL = the solver's own stream) and the caller synchronises.
 * Every call on a device shares that device's scratch, so calls that
 * overlap on one device (on different streams or threads) race; the
 * solver's other entry points must not run meanwhile either. */
int dg_rhs_device_noalloc(dg_solver* s, const double* u_dev, double* k_dev, double dt, void* stream);
/* one RK step of size dt on the resident state (scheme DG_RK4 / DG_RK2) */
int dg_rk_step(dg_solver* s, double dt, int scheme);
/* dt = cfl * min_r / max_K lambda_K / (2p+1) from the resident state */
int dg_cfl_dt(dg_solver* s, double cfl, double* dt);
/* time_integrate_rk4/rk2: from the current t to t_end with CFL steps, the last
 * one clamped to t_end; dt is computed on the device every step */
int dg_integrate(dg_solver* s, double t_end, double cfl, int scheme, int64_t* steps, double* t_out);
/* nsteps CFL steps with no end time (benchmarking); use_graph replays a
 * captured hipGraph of one step */
int dg_run_steps(dg_solver* s, int nsteps, double cfl, int scheme, int use_graph);
int dg_audit(dg_solver* s, dg_audit_out* out);
/* uv[3*num_elem]: variable var at the three vertices of each element (eval_u) */
int dg_eval_vertices(dg_solver* s, int var, double* uv);
/* Euler: the velocity at the three vertices of each element, u = (rho u)/rho
 * and v = (rho v)/rho from the vertex values (eval_u_velocity,
 * euler/cpu/euler_kernels.c:1131-1172); uv[6*K] = u1,v1,u2,v2,u3,v3 per element */
int dg_eval_vertices_velocity(dg_solver* s, double* uv);
/* Euler: the pressure at the three vertices of each element,
 * (gamma - 1)(E - (u^2 + v^2)/2 rho) from the vertex values (eval_p and
 * pressure, euler/cpu/euler_kernels.c:124-140, 1177-1240); pv[3*K] */
int dg_eval_vertices_pressure(dg_solver* s, double* pv);
int dg_solver_sync(dg_solver* s);
/* kernel selection for the stage kernels: 0 = automatic (timed per
 * physics, order, precision and size), 1 = one element per lane (VALU),
 * 2 = persistent matrix-core tiles of 64 elements, 3 = one matrix-core tile
 * per workgroup (2 and 3: FP32 or FP64, p = 1..4) */
int dg_solver_set_kernel(dg_solver* s, int kind);
int dg_solver_get_kernel(const dg_solver* s, int* kind);
/* Caps the grid of the persistent matrix-core kernel (kind 2) at max_groups
 * workgroups (rounded down to a multiple of 8; 0 = as many as are resident
 * on the device, the default), so each workgroup walks more tiles.  Results
 * do not depend on it.  Test and tuning aid; no counterpart in the reference,
 * whose kernels launch one thread per element or side. */
int dg_solver_set_max_groups(dg_solver* s, int max_groups);
/* Launch shape of the stage kernels: out[0] kernel kind, out[1] tiles (64
 * elements for kinds 2 and 3, 128 for kind 1), out[2] workgroups launched,
 * out[3] the most tiles one workgroup processes in a launch.  No counterpart
 * in the reference. */
int dg_solver_launch_info(const dg_solver* s, int* out4);
/* Times nsteps RK4 steps with HIP events on the solver's stream, one event
 * pair around every stage-kernel launch.  out[0] = mean stage-kernel ms,
 * out[1] = mean dt-kernel ms, out[2] = total ms, out[3] = number of stage
 * launches timed. */
int dg_time_stage_kernels(dg_solver* s, int nsteps, double cfl, double* out);
/* Runs nsteps RK4 CFL steps replayed from the captured step graph (the
 * form dg_integrate runs) with one HIP event pair on the solver's stream
 * around all of them: out[0] = ms per step, out[1] = total ms.  Warm up
 * first (e.g. dg_solver_run_steps with the graph): a graph not yet captured
 * is captured inside the timed span.  Measurement aid; no counterpart in
 * the reference. */
int dg_time_step_graph(dg_solver* s, int nsteps, double cfl, double* out);
/* device buffer bytes held by the solver */
int dg_solver_device_bytes(const dg_solver* s, int64_t* bytes);
/* Device-to-device streaming copy of `bytes` (multiple of 16), iters times
 * for each of 4 copy kernels (one or four 16-byte loads in flight per thread,
 * 1024 to 4096 workgroups), timed with HIP events; the fastest counts:
 * out[0] = mean ms per copy, out[1] = GB/s counting read + write bytes.
 * The bandwidth reference of the roofline (no counterpart in the
 * reference). */
int dg_bandwidth_probe(int64_t bytes, int iters, double* out);
/* dg_bandwidth_probe with out[2] = the fastest copy kernel's number (0: one
 * load per thread on 1024 workgroups, 1..3: four loads on 1024, 2048, 4096
 * workgroups) when nout >= 3. */
int dg_bandwidth_probe_ex(int64_t bytes, int iters, double* out, int nout);
/* Runs one 16x16x4 FP32 and one FP64 matrix-core product on the GPU and
 * compares them with the host product under the lane layouts the kernels
 * assume; *max_err = max absolute difference (no counterpart in the
 * reference). */
int dg_selftest_mfma(double* max_err);

/* ---------------- many solves, persistent launches per RK stage ---------------- */
int dg_batch_create(int precision, dg_batch** out);
void dg_batch_destroy(dg_batch* b);
/* adds one solve (its own mesh, problem, order p = 1..4, end time); returns
 * its slot id.  A mesh with boundary sides runs its own stage kernels inside
 * the batch's captured step (the batched work items carry no boundary
 * code): p >= 3 on the matrix-core kernel's boundary variant, p <= 2 on the
 * lane kernel. */
int dg_batch_add(dg_batch* b, const dg_mesh* m, const dg_problem_desc* prob, int p, double t_end, int* slot);
/* projects the initial conditions of every solve and resets their clocks */
int dg_batch_project_ic(dg_batch* b);
/* nsteps RK4 steps on every unfinished solve (each with its own CFL dt),
 * persistent launches per stage (dg_batch_set_launch_mode), replayed from a
 * hipGraph when use_graph */
int dg_batch_run_steps(dg_batch* b, int nsteps, double cfl, int use_graph);
/* runs every solve to its end time; *steps = number of batch steps issued */
int dg_batch_integrate(dg_batch* b, double cfl, int use_graph, int64_t* steps);
int dg_batch_get_state(dg_batch* b, int slot, double* c_ref);
int dg_batch_audit(dg_batch* b, int slot, dg_audit_out* out);
int dg_batch_info(const dg_batch* b, int* n_solves, int64_t* total_elems, int64_t* device_bytes);
/* Caps the batch's persistent grid at max_groups workgroups (rounded down to
 * a multiple of 8; 0 = every resident workgroup, the default).  Results do
 * not depend on it.  Test and tuning aid; no counterpart in the reference. */
int dg_batch_set_max_groups(dg_batch* b, int max_groups);
/* How a stage of the batch is launched; takes effect at the next run.
 * DG_BATCH_ONE_LAUNCH (the FP32 default): a single persistent launch per
 * stage carrying every body, at the register budget of the largest;
 * workgroups mix light and heavy items.  DG_BATCH_PER_CLASS (the FP64
 * default): one persistent launch per kernel class present (Euler tiles,
 * advection tiles, Euler lane items, advection lane items), each compiled
 * with its own bodies only and so holding the wavefronts per SIMD those
 * bodies allow (the FP64 one-launch kernel holds one wavefront per SIMD).
 * Results are the same bit for bit.  No counterpart in the reference (its
 * solves run one at a time). */
enum { DG_BATCH_ONE_LAUNCH = 0, DG_BATCH_PER_CLASS = 1 };
int dg_batch_set_launch_mode(dg_batch* b, int mode);
/* Launch shape of the last uploaded work list (uploaded by the first run
 * after dg_batch_add): out[0] workgroups (of the largest launch), out[1]
 * matrix-core tile items,
 * out[2] lane items, out[3] the longest run of consecutive tiles of one solve
 * that one workgroup takes through the pipelined tile body.  No counterpart
 * in the reference. */
int dg_batch_launch_info(const dg_batch* b, int* out4);

/* ---------------- serving: continuous batching ---------------- */
typedef struct {
    int precision;          /* DG_FP32 or DG_FP64 */
    int max_slots;          /* solves resident at once */
    int64_t hbm_budget;     /* bytes of device memory the resident solves may hold */
    int steps_per_round;    /* batch steps between admissions / completions */
    double cfl;
    int use_graph;
    int solo_min_elems;     /* solves with at least this many elements run their own stage
                               kernels inside the step graph (0: all in the batch launch) */
    int prep_threads;       /* host threads preparing admitted solves (mesh, setup, projection) */
    int fifo_min_elems;     /* solves with at least this many elements advance one at a time, in
                               arrival order, the smaller ones beside them (0: every resident
                               solve advances at every step) */
} dg_server_config;

typedef struct {
    int nx, ny;             /* jittered periodic mesh of 2*nx*ny triangles */
    double x0, x1, y0, y1, jitter;
    uint64_t mesh_seed;
    int p;
    double t_end;
    dg_problem_desc prob;
    double arrival_s;       /* arrival time, seconds on the server's clock (0 = now) */
} dg_request;

typedef struct {
    uint64_t id;
    int status;             /* dg_status of this request */
    int64_t steps;
    double t;
    double arrival_s, admit_s, finish_s; /* server clock, seconds */
    double mass[4];         /* FP64 audit of the final state */
    double mass0[4];        /* FP64 audit of the initial state */
    int64_t n_bad;
} dg_result;

int dg_server_create(const dg_server_config* cfg, dg_server** out);
/* queues a request; it is admitted when a slot and budget are free */
int dg_server_submit(dg_server* sv, const dg_request* req, uint64_t* id);
/* copies up to max finished results (oldest first) and removes them */
int dg_server_poll(dg_server* sv, dg_result* out, int max, int* n);
/* final state of a finished request, reference layout (requires keep_states) */
int dg_server_get_state(dg_server* sv, uint64_t id, double* c_ref, int64_t n);
int dg_server_keep_states(dg_server* sv, int keep);
/* blocks until every submitted request has finished */
int dg_server_drain(dg_server* sv);
/* seconds on the server's clock */
double dg_server_now(dg_server* sv);
void dg_server_destroy(dg_server* sv);

#ifdef __cplusplus
}
#endif
#endif
