/*
 * pda_c_api.h -- C ABI of the MI355X-native pressio-demoapps engine.
 *
 * This is the drop-in boundary for the reference's problem API
 * (include/pressiodemoapps/adapter_cpp.hpp:60-264 and the per-family
 * create_problem_eigen factories).  Every entry point takes plain pointers
 * and sizes; state, velocity and Jacobian values live in HIP device memory.
 * Streams are passed as void* and are hipStream_t values (NULL = default).
 *
 * Conventions:
 *   - return 0 on success, a negative PDA_E* code on failure;
 *     pda_last_error() returns a thread-local message for the last failure;
 *   - a problem handle holds a reference on its mesh;
 *   - one problem handle is not re-entrant (as the reference, whose
 *     evaluation methods mutate internal ghost buffers); calls are ordered on
 *     the stream they are given; distinct handles are independent.
 */
#ifndef PDA_C_API_H
#define PDA_C_API_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pda_mesh_s *pda_mesh;
typedef struct pda_problem_s *pda_problem;
typedef void *pda_stream; /* hipStream_t */

enum {
  PDA_OK = 0,
  PDA_E_INVALID_ARG = -1,
  PDA_E_UNSUPPORTED = -2,
  PDA_E_IO = -3,
  PDA_E_HIP = -4,
  PDA_E_STATE = -5
};

/* families: one per reference problem header */
enum {
  PDA_EULER1D = 1,    /* include/pressiodemoapps/euler1d.hpp:65-70  */
  PDA_EULER2D = 2,    /* include/pressiodemoapps/euler2d.hpp:67-77  */
  PDA_EULER3D = 3,    /* include/pressiodemoapps/euler3d.hpp:65-68  */
  PDA_SWE2D = 4,      /* include/pressiodemoapps/swe2d.hpp:66-69    */
  PDA_BURGERS2D = 5,  /* include/pressiodemoapps/advection_diffusion2d.hpp:65-68 */
  PDA_ADR2D = 6       /* include/pressiodemoapps/advection_diffusion_reaction2d.hpp:61-62 */
};

/* InviscidFluxReconstruction, include/pressiodemoapps/schemes_info.hpp:61-63 */
enum { PDA_FIRSTORDER = 0, PDA_WENO3 = 1, PDA_WENO5 = 2 };

const char *pda_last_error(void);
const char *pda_version(void);

/* ---------------------------------------------------------------- mesh */

/* Reads a mesh directory in the reference text format (info.dat,
   coordinates.dat, connectivity.dat).  Replaces
   load_cellcentered_uniform_mesh_eigen, include/pressiodemoapps/mesh.hpp:87-91
   and impl/mesh_ccu.hpp:359-448; returns PDA_E_IO instead of exit(). */
int pda_mesh_load(const char *dir, pda_mesh *out);

/* Builds the natural-order structured mesh that
   meshing_scripts/create_full_mesh.py:32-213 writes (GID = gk*Nx*Ny+gj*Nx+gi,
   neighbour columns [l0 f0 r0 b0 (d0 u0)] per layer, periodic wrap per axis),
   including its %.14f round trip of dx and of the cell centres.
   n[dim], bounds = {xmin,xmax,ymin,ymax,zmin,zmax}; periodic bit a = axis a. */
int pda_mesh_create_structured(int dim, const int64_t *n, const double *bounds,
                               int stencil, int periodic_mask, pda_mesh *out);

/* Builds a sample mesh from a full mesh and sample-cell ids, as
   meshing_scripts/create_sample_mesh.py:28-137 does.  stencil_gids_out
   (nullable, size >= returned stencil size) receives the stencil-mesh cells
   as full-mesh ids, sample_gids_out (nullable) the sorted unique sample
   cells. */
int pda_mesh_create_sample(pda_mesh full, const int32_t *sample_gids, int64_t n,
                           pda_mesh *out);
int pda_mesh_sample_gids(pda_mesh sample, int32_t *sample_gids_out,
                         int32_t *stencil_gids_out);

/* Writes info.dat / coordinates.dat / connectivity.dat in the reference
   format (meshing_scripts/create_full_mesh.py:152-213). */
int pda_mesh_write(pda_mesh mesh, const char *dir);

int pda_mesh_info(pda_mesh mesh, int *dim, int *stencil, int64_t *n_sample,
                  int64_t *n_stencil, double *d /*[3]*/);
/* host views valid for the mesh lifetime */
const int32_t *pda_mesh_graph(pda_mesh mesh);      /* n_sample x ncols */
const double *pda_mesh_coords(pda_mesh mesh, int axis); /* n_stencil */
/* counts of inner / near-boundary graph rows, impl/mesh_ccu.hpp:385-447 */
int pda_mesh_row_classes(pda_mesh mesh, int64_t *n_inner, int64_t *n_near);
/* graph rows of a row class, in row order, as the reference mesh lists them
   (impl/mesh_ccu.hpp:385-447): which = 0 graphRowsOfCellsNearBd (a -1 within
   the mesh stencil width), 1 graphRowsOfCellsAwayFromBd, 2
   graphRowsOfCellsStrictlyOnBd (a missing first-order neighbour; 2D meshes
   only, else PDA_E_UNSUPPORTED).  rows (nullable) receives *count ids. */
int pda_mesh_boundary_rows(pda_mesh mesh, int which, int32_t *rows, int64_t *count);
void pda_mesh_destroy(pda_mesh mesh);

/* ------------------------------------------------------------- problem */

/* Replaces create_problem_eigen(mesh, ProblemEnum, InviscidFluxReconstruction
   [, icFlag][, parameter map]) for the six families (e.g.
   include/pressiodemoapps/euler2d.hpp:88-173, swe2d.hpp:143-195,
   advection_diffusion2d.hpp:80-179, advection_diffusion_reaction2d.hpp:111-180).
   ic_flag <= 0 selects the family default.  Parameter names are the
   reference's ("gamma", "gravity", "coriolis", "pulseMagnitude",
   "diffusion", "riemannTopRightPressure", ...).  Unknown names and
   unsupported problem enums return PDA_E_INVALID_ARG / PDA_E_UNSUPPORTED. */
int pda_problem_create(pda_mesh mesh, int family, int prob, int scheme,
                       int ic_flag, const char *const *keys, const double *vals,
                       int nparams, pda_problem *out);
void pda_problem_destroy(pda_problem p);

/* numDofPerCell / totalDofSampleMesh / totalDofStencilMesh and nonZeros of
   the Jacobian, adapter_cpp.hpp:93-103 */
int pda_problem_sizes(pda_problem p, int *ndpc, int64_t *n_sample_dofs,
                      int64_t *n_stencil_dofs, int64_t *nnz);
/* parameter lookup, e.g. gamma() / gravity() / coriolis() / queryParameter */
int pda_problem_parameter(pda_problem p, const char *name, double *value);

/* initialCondition(), into device memory of n_stencil_dofs doubles */
int pda_initial_condition(pda_problem p, double *d_U, pda_stream stream);
/* the same into host memory (no device needed) */
int pda_initial_condition_host(pda_problem p, double *U);

/* Jacobian sparsity, fixed at creation: scalar CSR with the reference's row
   order and sorted columns (Eigen setFromTriplets + makeCompressed,
   impl/euler_2d_prob_class.hpp:223-237).  Device pointers, valid for the
   problem lifetime.  The host variant copies into caller arrays of
   n_sample_dofs+1 and nnz entries. */
int pda_jacobian_pattern(pda_problem p, int64_t *n_rows, int64_t *nnz,
                         const int64_t **d_rowptr, const int32_t **d_colidx);
int pda_jacobian_pattern_host(pda_problem p, int64_t *rowptr, int32_t *colidx);

/* rightHandSide(U, t, V): adapter_cpp.hpp:193-199 */
int pda_velocity(pda_problem p, const double *d_U, double t, double *d_V,
                 pda_stream stream);

/* Velocity tuning (no counterpart in the reference): times the candidate
   strip heights (rows per workgroup) of the structured velocity kernel on
   d_U and keeps the fastest for this problem; synchronizes the stream.
   Without a call the height follows the grid size.  *rows_per_strip = 0
   when the problem's velocity does not use the structured kernel; *us is
   the time of one velocity() call with the kept height.  The values of
   velocity() do not depend on the height. */
int pda_problem_autotune(pda_problem p, const double *d_U, pda_stream stream,
                         int *rows_per_strip, double *us);

/* Sets the strip height of the structured velocity kernel (no counterpart
   in the reference), for callers that tune it in their own loop: 0 restores
   the grid-size rule, 1..64 fixes the height; other values are
   PDA_E_INVALID_ARG.  Problems whose velocity does not use the structured
   kernel accept and ignore it.  The values of velocity() do not depend on
   the height. */
int pda_problem_set_velocity_strip_height(pda_problem p, int rows_per_strip);

/* Sets the strip height (grid rows marched by one workgroup) of the
   first-order 2D and 3D Euler Jacobian kernels used by jacobian() without V
   on a structured mesh (no counterpart in the reference): 0 restores the
   grid-size rule (about 16k workgroups in 2D, 64k in 3D), 1..64 fixes the
   height; other values are PDA_E_INVALID_ARG.  Problems that do not use
   those kernels accept and ignore it.  The Jacobian values do not depend on
   the height. */
int pda_problem_set_jacobian_strip_height(pda_problem p, int rows_per_strip);

/* rightHandSideAndJacobian(U, t, V, J) (adapter_cpp.hpp:201-208) when d_V is
   non-NULL, jacobian(U, t, J) (adapter_cpp.hpp:223-229) when it is NULL.
   d_Jvals has nnz entries in the order of pda_jacobian_pattern. */
int pda_jacobian(pda_problem p, const double *d_U, double t, double *d_Jvals,
                 double *d_V, pda_stream stream);

/* applyJacobian(U, B, t, AJ): adapter_cpp.hpp:231-265.  B has n_stencil_dofs
   rows and ncols columns; layout 0 = column-major, 1 = row-major; AJ has
   n_sample_dofs rows in the same layout. */
int pda_apply_jacobian(pda_problem p, const double *d_U, double t,
                       const double *d_B, int64_t ncols, int layout,
                       double *d_AJ, pda_stream stream);

/* applyJacobian for callers that cannot allocate device memory.  No
   counterpart in the reference: it is for code under HIP graph capture,
   tight serving loops and harnesses that forbid hipMalloc.  The same
   computation as pda_apply_jacobian, with the Jacobian values held in one
   fixed device array of PDA_NOALLOC_MAX_NNZ doubles (64 MiB, reserved on
   every device the library runs on until the process ends) instead of a
   per-problem buffer allocated on first use; it calls no allocating
   routine.  Refuses (PDA_E_INVALID_ARG, before any device work) a problem
   with more nonzeros than the array holds, and (PDA_E_STATE) a problem
   whose device tables were never built: call pda_jacobian_pattern or any
   evaluation once before.  All calls on one device share the array, so
   calls that overlap on one device (other threads, other streams) race;
   calls ordered on one stream do not. */
#define PDA_NOALLOC_MAX_NNZ 8388608
int pda_apply_jacobian_noalloc(pda_problem p, const double *d_U, double t,
                               const double *d_B, int64_t ncols, int layout,
                               double *d_AJ, pda_stream stream);

/* ---------------------------------------------------- explicit stepping */

/* Explicit time integration of dU/dt = velocity(U) on the device with
   pressio's stage formulas
   (pressio/ode/impl/ode_explicit_stepper_without_mass_matrix.hpp: forward
   Euler :168-187, SSPRK3 :230-281, RK4 :284-352), as used by the
   reference's tests and by pressiodemoapps.advanceRK4/advanceSSP3
   (pressiodemoapps/__init__.py:117-185).  Full meshes only; a stepper
   must not outlive its problem. */
enum { PDA_FORWARD_EULER = 0, PDA_SSPRK3 = 1, PDA_RK4 = 2,
       /* the Python module's integrators with its own stage algebra, rounded
          as numpy rounds it: advanceRK2 (pressiodemoapps/__init__.py:90-114),
          advanceRK4 (:117-152), advanceSSP3 (:155-185) */
       PDA_RK2 = 3, PDA_MODULE_RK4 = 4, PDA_MODULE_SSP3 = 5 };
typedef struct pda_stepper_s *pda_stepper;
int pda_stepper_create(pda_problem p, int scheme, pda_stepper *out);
/* nsteps steps of size dt from time t0, in place on d_U.  use_graph != 0
   records one step as a hipGraph (re-recorded when d_U or dt change) and
   replays it per step on the stream.  SSPRK3 and RK4 on a 1D Euler problem
   over a natural-order full mesh run several steps per launch instead (the
   stepper's work buffer holds every other launch's state), whatever
   use_graph says. */
int pda_stepper_advance(pda_stepper st, double *d_U, double t0, double dt,
                        int64_t nsteps, int use_graph, pda_stream stream);
void pda_stepper_destroy(pda_stepper st);

/* ------------------------------------------------------------ batching */

/* Many independent instances (no counterpart in the reference, which runs
   one problem per object): 2D structured problems of one family and scheme
   whose boundaries are periodic or mirror ghosts (so the velocity kernel
   covers the whole grid), each with its own grid, parameters and state
   slice.  The batch state is the concatenation of the instances' states in
   instance order (offsets from pda_batch_sizes); one launch evaluates every
   instance's velocity.  A batch must not outlive its problems. */
typedef struct pda_batch_s *pda_batch;
int pda_batch_create(const pda_problem *probs, int count, pda_batch *out);
void pda_batch_destroy(pda_batch b);
/* total state doubles and, if offsets is non-NULL, count+1 offsets */
int pda_batch_sizes(pda_batch b, int64_t *total, int64_t *offsets);
int pda_batch_initial_condition(pda_batch b, double *d_U, pda_stream stream);
int pda_batch_velocity(pda_batch b, const double *d_U, double t, double *d_V, pda_stream stream);
/* times the candidate strip heights (rows per workgroup) of the batched
   velocity kernel on d_U and keeps the fastest; synchronizes the stream */
int pda_batch_autotune(pda_batch b, const double *d_U, pda_stream stream, int *rows_per_strip,
                       double *us);
/* fixes the batched velocity kernel's strip height (1..64, else
   PDA_E_INVALID_ARG), for callers that time heights in their own stepping
   loop (no counterpart in the reference); synchronizes the device (the
   tables are rebuilt) and drops a recorded graph; values do not depend on
   the height */
int pda_batch_set_velocity_strip_height(pda_batch b, int rows_per_strip);
/* nsteps explicit steps (PDA_FORWARD_EULER / SSPRK3 / RK4 / RK2) of the whole
   batch; steps_per_graph > 0 records that many steps as one hipGraph
   (re-recorded when d_U, dt, the scheme or the count change) and replays it */
int pda_batch_advance(pda_batch b, int scheme, double *d_U, double t0, double dt, int64_t nsteps,
                      int steps_per_graph, pda_stream stream);

/* ---------------------------------------------------- implicit stepping */

/* Implicit time stepping y_n -> y_{n+1} with pressio's discrete residuals
   and Jacobians (pressio/ode/impl/ode_implicit_discrete_residual.hpp:112-240,
   ode_implicit_discrete_jacobian.hpp:60-121):
     BDF1: R = y - y_n - dt f(y),                        J = I - dt df/dy
     BDF2: R = y - 4/3 y_n + 1/3 y_{n-1} - 2/3 dt f(y),   J = I - 2/3 dt df/dy
     CN:   R = y - y_n - dt/2 (f(y) + f(y_n)),            J = I - dt/2 df/dy
   solved by Newton on the exact Jacobian with a device BiCGSTAB
   (block-Jacobi preconditioner).  Full meshes only; an implicit stepper
   must not outlive its problem. */
enum { PDA_BDF1 = 1, PDA_BDF2 = 2, PDA_CRANK_NICOLSON = 3 };
/* PDA_STOP_CORRECTION: pressio's default rule, stop when ||dx||_2 < tol
   without applying that correction
   (pressio/solvers_nonlinear/impl/nonlinear_least_squares.hpp:70-84);
   PDA_STOP_RESIDUAL: stop when ||R(y)||_2 <= tol. */
enum { PDA_STOP_CORRECTION = 0, PDA_STOP_RESIDUAL = 1 };
typedef struct {
  int newton_iters;   /* Newton iterations (Jacobian evaluations with a solve) */
  int linear_iters;   /* BiCGSTAB iterations summed over the step */
  int converged;
  double residual_norm;            /* ||R||_2 at the last evaluation */
  double correction_norm;          /* ||dx||_2 of the last solve */
  double last_linear_rel_residual; /* ||b - A dx|| / ||b|| of the last solve */
  double initial_residual_norm;    /* ||R||_2 at the initial iterate (the guess) */
} pda_implicit_stats;
typedef struct pda_implicit_s *pda_implicit;
int pda_implicit_create(pda_problem p, int scheme, pda_implicit *out);
/* defaults: PDA_STOP_CORRECTION, tol 1e-6, 100 Newton iterations,
   linear relative tolerance 1e-14, linear iteration cap 2*unknowns (0) */
int pda_implicit_set_options(pda_implicit ip, int stop_rule, double tol, int max_newton,
                             double lin_rtol, int lin_max);
/* One step from time t: d_U holds y_n on entry and y_{n+1} on return.
   d_Uprev = y_{n-1} for BDF2 (NULL: this step is BDF1, as pressio's first
   BDF2 step).  d_guess: Newton's initial iterate (NULL: y_n, pressio's
   choice).  Synchronizes the stream (the Newton loop reads norms).
   Failures return PDA_E_STATE: a non-finite residual or Newton correction
   leaves d_U holding y_n (pressio's NonlinearSolveFailure, after which its
   stepper restores y_n: ode_implicit_stepper_standard.hpp:213-225); reaching
   max_newton without meeting tol leaves the last iterate in d_U with
   stats->converged = 0 (pressio stops there without an error,
   solvers_nonlinear/impl/root_finder.cpp:57-104; here the return code
   reports it). */
int pda_implicit_step(pda_implicit ip, double *d_U, const double *d_Uprev,
                      const double *d_guess, double t, double dt, pda_stream stream,
                      pda_implicit_stats *stats);
void pda_implicit_destroy(pda_implicit ip);

/* ------------------------------------------------- learned Newton guess */

/* A 4-layer 3x3 convolutional network (32 channels, BF16 weights and
   activations, FP32 accumulation on MFMA) that picks, per BDF2 step of a
   periodic nx x ny two-component field, the extrapolation order t in [1, 4]
   of an FP64 predictor of y_{n+1}: the BDF2 relation with f(y_{n+1})
   extrapolated in time with order k from the velocities f_n, ..., f_{n-k},
     P_k = (4/3) y_n - (1/3) y_{n-1} + (2/3) dt E_k,
     E_1 = 2 f_n - f_{n-1},  E_2 = 3 f_n - 3 f_{n-1} + f_{n-2},
     E_3 = 4 f_n - 6 f_{n-1} + 4 f_{n-2} - f_{n-3},
     E_4 = 5 f_n - 10 f_{n-1} + 10 f_{n-2} - 5 f_{n-3} + f_{n-4},
   guess = P(t), linear between P_k and P_{k+1} for k = floor(t).  The
   means m_j over the grid of the network's PDA_GUESSNET_NORDERS output
   channels estimate what the step costs from P(T_j), T_j = 1 + j/2 (j =
   0..6), relative to the mean over the orders; t = sum_j p_j T_j with p =
   softmax(-m / 0.25), a soft arg-min.  Its inputs are the time differences dt D^k f /
   s_k (k = 1, 2, 3), |dt D^4 f| / s_4 and (nu - nu_mean) / nu_scale.  No
   counterpart in the reference: it feeds pda_implicit_step's d_guess, and
   Newton on the exact Jacobian still decides the converged state.  States
   are device arrays of nx*ny*2 doubles in cell order j*nx + i; d_f is a
   host array of five device pointers {f_n, f_{n-1}, f_{n-2}, f_{n-3},
   f_{n-4}} (the problem's velocity at y_n, ..., y_{n-4}).  nx*ny must be a
   multiple of 512 and at most 2^30. */
#define PDA_GUESSNET_NORDERS 7
typedef struct pda_guessnet_s *pda_guessnet;
int pda_guessnet_create(int nx, int ny, pda_guessnet *out);
void pda_guessnet_destroy(pda_guessnet net);
int pda_guessnet_num_params(pda_guessnet net, int64_t *n);
/* FP32 master parameters, per layer: W[tap][cout][cin] then bias[cout];
   layers (cin, cout) = (8, 32), (32, 32), (32, 32), (32, 7).  Setting them
   resets the optimizer state. */
int pda_guessnet_set_params(pda_guessnet net, const float *w, int64_t n);
int pda_guessnet_get_params(pda_guessnet net, float *w, int64_t n);
/* dt: the BDF2 time step; spreads[4]: of dt D^k f, k = 1..4; all positive */
int pda_guessnet_set_scales(pda_guessnet net, double dt, const double *spreads, double nu_mean,
                            double nu_scale);
int pda_guessnet_predict(pda_guessnet net, const double *d_un, const double *d_unm1,
                         const double *const *d_f, double nu, double *d_guess, pda_stream stream);
/* order t and (nullable, PDA_GUESSNET_NORDERS values each) the weights p_j
   and the cost estimates m_j of the last predict / gradients / train_step
   call (synchronizes) */
int pda_guessnet_policy(pda_guessnet net, double *order, double *probs, double *estimates);
/* Training loss of one sample: sum_j (m_j - (costs[j] - mean costs))^2 / 7
   for what the step cost from each P(T_j) (host, PDA_GUESSNET_NORDERS
   finite values, e.g. Newton + Krylov iterations), and its gradient w.r.t.
   the parameters (host arrays, nullable; out = the net's raw output,
   nx*ny*PDA_GUESSNET_NORDERS floats). */
int pda_guessnet_gradients(pda_guessnet net, const double *const *d_f, double nu,
                           const double *costs, float *grad, float *loss, float *out,
                           pda_stream stream);
/* one Adam step (beta 0.9 / 0.999) on that loss; loss (nullable) is
   reported before the update */
int pda_guessnet_train_step(pda_guessnet net, const double *const *d_f, double nu,
                            const double *costs, float lr, float *loss, pda_stream stream);

/* Block view of the Jacobian (m x m blocks).  For each block-row r and block
   slot s < nblk[r], d_bcol[r*maxb+s] is the column cell and the block values
   start at d_Jvals[rowptr[r*m] + s*m] with row stride nblk[r]*m. */
int pda_jacobian_blocks(pda_problem p, int *maxb, const int32_t **d_nblk,
                        const int32_t **d_bcol);

#ifdef __cplusplus
}
#endif
#endif
