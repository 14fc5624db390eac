/*
 * philip_dg.h -- C ABI of the MI355X-native explicit strong-form DG path.
 *
 * The reference (PHiLiP) has no foreign-function interface: its callers are
 * C++ (ODE solvers, flow solver, tests).  The entry points below are what a
 * C/C++/ctypes/cgo caller binds to replace that path; each cites the
 * reference interface it stands in for.  No torch or HIP types cross this
 * boundary: plain pointers, sizes, doubles.  Every entry point returns 0 on
 * success or a negative DGX_ERR_* code; dgx_last_error() gives the message.
 *
 * Vector layout everywhere is the reference's DoF order without
 * renumbering: v[(cell*nstate + state)*n_nodes + node], node lexicographic
 * with x fastest (deal.II FE_DGQ), cells in deal.II active-cell order (Z order
 * after refine_global for 2^k cells per direction, lexicographic otherwise).
 */
#ifndef PHILIP_DG_H
#define PHILIP_DG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DGX_OK 0
#define DGX_ERR_INVALID (-1)      /* invalid configuration or argument        */
#define DGX_ERR_HIP (-2)          /* HIP runtime error                        */
#define DGX_ERR_NONPHYSICAL (-3)  /* non-physical state detected              */
#define DGX_ERR_UNSUPPORTED (-4)  /* configuration the device path lacks      */
#define DGX_ERR_STATE (-5)        /* call out of order (e.g. no ODE system)   */

/* enumerations mirror Parameters::AllParameters (src/parameters/all_parameters.cpp:73-399) */
enum dgx_flux_nodes { DGX_GLL = 0, DGX_GL = 1 };
enum dgx_two_point { DGX_KG = 0, DGX_IR = 1, DGX_CH = 2, DGX_RA = 3 };
enum dgx_conv_flux {
    DGX_LAX_FRIEDRICHS = 0, DGX_ROE = 1, DGX_CENTRAL = 2, DGX_TWO_POINT = 3,
    DGX_TWO_POINT_LXF = 4, DGX_TWO_POINT_ROE = 5
};
enum dgx_pde { DGX_EULER = 0, DGX_NAVIER_STOKES = 1 };
enum dgx_diss_flux { DGX_SIPG = 0, DGX_BR2 = 1 };
enum dgx_cell_order { DGX_ORDER_DEALII = 0, DGX_ORDER_LEXICOGRAPHIC = 1 };
enum dgx_precision { DGX_FP64 = 0, DGX_MIXED = 1 };
/* ODE methods (Parameters::ODESolverParam, parameters_ode_solver.cpp:111-296) */
enum dgx_ode_method {
    DGX_EULER_EX = 0, DGX_HEUN_EX = 1, DGX_SSPRK3_EX = 2, DGX_RK4_EX = 3,
    DGX_RK4_3_5_3SSTAR = 10, DGX_RK3_2_5F_3SSTARPLUS = 11,
    DGX_RK4_3_9F_3SSTARPLUS = 12, DGX_RK5_4_10F_3SSTARPLUS = 13
};

/* Scheme + mesh description: the subset of AllParameters this path reads
   (all_parameters.cpp:73-155,325-351; parameters_navier_stokes.cpp:17-32;
   flow solver grid subsection) plus the device precision. */
typedef struct {
    int dim;                    /* 2 or 3                                          */
    int poly_degree;            /* p, 1..7                                         */
    int flux_nodes;             /* DGX_GLL (device path); DGX_GL rejected           */
    int use_split_form;
    int two_point_flux;         /* two_point_num_flux_type                         */
    int conv_flux;              /* conv_num_flux                                   */
    int pde;                    /* pde_type                                        */
    int diss_flux;              /* diss_num_flux                                   */
    int use_constant_viscosity;
    double gamma_gas;
    double mach_inf;
    double reynolds_inf;
    double prandtl;
    double constant_viscosity;
    double temperature_inf;
    double sipg_penalty_factor;
    int n_cells_1d;             /* straight_periodic_cube cells per direction      */
    double domain_left;
    double domain_right;
    int cell_order;
    int precision;              /* DGX_FP64 or DGX_MIXED (FP32 fluxes, FP64 state)  */
    /* Boundary condition of each face of the box (deal.II face numbering: 2k = -e_k, 2k+1 =
       +e_k): 0 = periodic (both faces of a direction), else a type of
       Euler::boundary_face_values (euler.cpp:1511-1567) applied in assemble_boundary_term_strong
       (strong_dg.cpp:1583-2364): 1001 wall, 1002 pressure outflow, 1003 inflow, 1004 Riemann
       farfield, 1005 farfield, 1006 slip wall, 1007 extrapolation; Navier-Stokes: 1001 no-slip
       wall and 1006 slip wall (boundary_face_values_viscous_flux, navier_stokes.cpp:1354-1430). */
    int boundary_type[6];
    double angle_of_attack;     /* radians: free stream (cos a, sin a[, 0]) of the boundary states */
    /* 0: the affine box; 1: the box's nodes through the reference's curved periodic grid
       (NonsymmetricCurvedGridManifold, nonsymmetric_curved_periodic_grid.cpp:66-87, beta = 0.1,
       made for [-1,1]^2 and [0,2pi]^3) with curvilinear metric terms (Euler, and Navier-Stokes in 3D). */
    int grid_warp;
    /* Navier-Stokes no-slip walls (1001): 0 adiabatic (the reference's default), 1 isothermal at
       wall_temperature (thermal_boundary_condition_type, nondimensionalized_isothermal_wall_temperature,
       parameters_navier_stokes.cpp:21-50).  Navier-Stokes takes the wall types 1001 and 1006 only. */
    int thermal_bc;
    double wall_temperature;
} dgx_params;

typedef struct dgx_ctx dgx_ctx;

/* Fills the reference defaults (all_parameters.cpp declare_parameters). */
void dgx_default_params(dgx_params *p);

/* DGFactory::create_discontinuous_galerkin + allocate_system
   (flow_solver.cpp:44, dg_factory.cpp, dg_base.hpp:184): builds operators,
   connectivity and device buffers on the current HIP device. */
int dgx_create(const dgx_params *p, dgx_ctx **out);
int dgx_destroy(dgx_ctx *ctx);
const char *dgx_last_error(const dgx_ctx *ctx);

int64_t dgx_n_dofs(const dgx_ctx *ctx);
int dgx_n_cells(const dgx_ctx *ctx);

/* SetInitialCondition::interpolate_initial_condition (set_initial_condition.cpp:45-51)
   of InitialConditionFunction_TaylorGreenVortex[_Isothermal]
   (initial_condition_function.cpp:261-336); resets current_time/iteration. */
int dgx_set_initial_condition_tgv(dgx_ctx *ctx, int isothermal);
/* Isentropic vortex (initial_condition_function.cpp:643-690) with strength and advection
   velocity as parameters: beta = sqrt(2/gamma) sqrt(2) strength/(4 pi) e^(1/2); strength 5
   and (u,v) = sqrt(2/gamma)(cos pi/4, sin pi/4) is the reference's hard-coded case. */
int dgx_set_initial_condition_vortex(dgx_ctx *ctx, double strength, double u_inf, double v_inf);

/* DGBase::solution (dg_base.hpp:409): host <-> device copies. */
int dgx_set_solution(dgx_ctx *ctx, const double *host_u);
int dgx_get_solution(const dgx_ctx *ctx, double *host_u);
/* DGBase::right_hand_side (dg_base.hpp:396): the un-lifted residual. */
int dgx_get_right_hand_side(const dgx_ctx *ctx, double *host_r);
/* DGBase::auxiliary_solution[idim] (dg_base.hpp:419), Navier-Stokes only.  FP64 contexts hold q
   in double; mixed-precision contexts hold it in float (their fluxes are float) and return it
   widened to double. */
int dgx_get_auxiliary_solution(const dgx_ctx *ctx, int idim, double *host_q);
/* integer connectivity used by the kernels: nbr[cell][2*dim] (deal.II face numbering). */
int dgx_get_neighbors(const dgx_ctx *ctx, int32_t *host_nbr);
/* Curvilinear grid: the mapping support points of every cell at its GLL nodes (grid degree = p),
   host array [cell][node][3] in the reference's cell order -- what HighOrderGrid holds
   (high_order_grid.cpp) for grid_degree = poly_degree -- from which the device builds the metric
   terms of metric_operators (operators.cpp:2520-2884: Jacobian, det J, and in 3D the conservative
   curl form of the cofactor) and each cell's SIPG penalty from its extent along each direction
   (evaluate_penalty_scaling, dg_base.cpp:397-413).  The box topology (periodic or boundary faces) is
   kept.  The batched line kernel does not take curvilinear contexts. */
int dgx_set_grid_nodes(dgx_ctx *ctx, const double *host_xyz);
/* node coordinates in use [cell][node][3] (the affine box if none were set) */
int dgx_get_grid_nodes(dgx_ctx *ctx, double *host_xyz);
/* metric terms [cell][dim*dim + 1][node]: cofactor C[m][n] = J dxi_n/dx_m in slot m*dim + n, det J
   in slot dim*dim; DGX_ERR_STATE on the affine box */
int dgx_get_metric_terms(dgx_ctx *ctx, double *host_met);

/* DGBase::assemble_residual(false,false,false,0.0) (dg_base.hpp:571,
   dg_base.cpp:2599-2914): right_hand_side = R(solution). */
int dgx_assemble_residual(dgx_ctx *ctx);
/* DGBase::apply_inverse_global_mass_matrix(right_hand_side, out) (dg_base.hpp:267,
   dg_base.cpp:4089-4260); out is a host array of n_dofs. */
int dgx_apply_inverse_mass_to_rhs(dgx_ctx *ctx, double *host_out);

/* Device-pointer variants (inputs/outputs already resident in HBM, same layout).
   u_dev, r_dev: n_dofs doubles on the context's device. */
int dgx_assemble_residual_device(dgx_ctx *ctx, const double *u_dev, double *r_dev);
int dgx_rhs_device(dgx_ctx *ctx, const double *u_dev, double *dudt_dev);
/* DGBase::apply_inverse_global_mass_matrix(in, out) (dg_base.hpp:267, dg_base.cpp:4089-4260)
   on device buffers of n_dofs doubles; in_dev == out_dev is allowed. */
int dgx_apply_inverse_mass_device(dgx_ctx *ctx, const double *in_dev, double *out_dev);

/* ODESolverFactory::create_ODESolver + allocate_ode_system (ode_solver_factory.cpp:20-35,
   ode_solver_base.h:79): allocates the RK registers for `method`. */
int dgx_allocate_ode_system(dgx_ctx *ctx, int method);
/* ODESolverBase::step_in_time(dt, pseudotime=false) called n_steps times
   (ode_solver_base.h:70, runge_kutta_base.cpp:21-38): advances solution,
   current_time += dt, current_iteration += 1 per step. */
int dgx_step_in_time(dgx_ctx *ctx, double dt, int n_steps);
double dgx_current_time(const dgx_ctx *ctx);
int64_t dgx_current_iteration(const dgx_ctx *ctx);
/* Restores the solver clock when restarting from saved data (FlowSolver::run with
   restart_computation_from_file, flow_solver.cpp:463-481: time and iteration continue). */
int dgx_set_time(dgx_ctx *ctx, double time, int64_t iteration);
/* 3S*+ methods: also track the embedded register (extra FSAL residual per step,
   low_storage_runge_kutta_ode_solver.cpp:63-69); off by default because it does
   not change the solution at a fixed time step. */
int dgx_set_track_error_estimate(dgx_ctx *ctx, int enable);

/* Error-adaptive low-storage RK (flow_solver_param.error_adaptive_time_step, flow_solver.cpp:505-507,
   603-607; LowStorageRungeKuttaODESolver, low_storage_runge_kutta_ode_solver.cpp:89-178).
   dgx_set_error_adaptive_parameters: the ODE parameters atol, rtol, beta1..3 (defaults 1e-3, 1e-3,
   0.70, -0.23, 0.0, parameters_ode_solver.cpp:159-178); resets the controller history
   epsilon = {1,1,1} and makes 3S*+ steps evaluate their FSAL residual (the embedded register).
   dgx_get_automatic_initial_step_size: get_automatic_initial_step_size (:119-178) from two
   right-hand sides of the current solution (needs a low-storage ODE system).
   dgx_get_error_adaptive_step_size: get_automatic_error_adaptive_step_size (:89-118) after a
   step of size dt: w = RMS of (S1 - E) / (atol + rtol max(|S1|, |E|)) with E the embedded
   solution (S4, or the S2 of sum_stages for RK4(3)5[3S*]), epsilon shifted in, and
   dt_next = eps0^(beta1/k) eps1^(beta2/k) eps2^(beta3/k) dt (k the method's order). */
int dgx_set_error_adaptive_parameters(dgx_ctx *ctx, double atol, double rtol, double beta1, double beta2,
                                      double beta3);
int dgx_get_automatic_initial_step_size(dgx_ctx *ctx, double *dt);
int dgx_get_error_adaptive_step_size(dgx_ctx *ctx, double dt, double *dt_next);

/* Relaxation Runge-Kutta for the explicit Butcher methods (ode_solver_param.use_relaxation_runge_kutta
   and relaxation_runge_kutta_root_tolerance, default 5e-10, parameters_ode_solver.cpp:144-151;
   RungeKuttaODESolver::adjust_time_step, runge_kutta_ode_solver.cpp:101-109; root-finding RRK,
   root_finding_rrk_ode_solver.cpp:16-390): each step's update u_n + gamma dt sum b_i k_i and time
   advance gamma dt, gamma the root of S(u_n + gamma d) - S(u_n) = gamma dt sum b_i v(u^(i)) . M k_i
   (S the integrated numerical entropy, v the entropy variables).  Enable after
   dgx_allocate_ode_system; DGX_ERR_NONPHYSICAL from a step whose root finding fails (the reference
   aborts).  dgx_relaxation_parameter: gamma of the last step (1 before the first). */
int dgx_set_relaxation_runge_kutta(dgx_ctx *ctx, int enable, double root_tolerance);
double dgx_relaxation_parameter(const dgx_ctx *ctx);

/* PeriodicTurbulence::get_constant_time_step (periodic_turbulence.cpp:115-125). */
double dgx_constant_time_step(const dgx_ctx *ctx, double cfl);
/* PeriodicTurbulence::compute_and_update_integrated_quantities (periodic_turbulence.cpp:439-671)
   with QGauss(p+1+overint): out[0]=kinetic energy, [1]=enstrophy, [2]=incompressible KE,
   [3]=vorticity-based dissipation rate; reductions in FP64 on the device.  On a curvilinear grid
   the metric terms are built at the QGauss points as the reference's metric operators do
   (periodic_turbulence.cpp:487-492). */
int dgx_integrated_quantities(dgx_ctx *ctx, int overint, double *out4);
/* Allocation-free variant of dgx_integrated_quantities.  No counterpart in the reference: it is
   for callers that cannot allocate device memory (a time loop captured in a HIP graph that
   records the kinetic-energy history on the device, a serving loop that must not hipMalloc).
   Same kernels and same results, bit for bit; it allocates nothing (no hipMalloc/hipFree, no
   host synchronisation) and only enqueues work on the context's stream: u_dev (n_dofs doubles
   on the device; NULL = the context's solution) in, out4_dev (4 doubles on the device, in the
   order of dgx_integrated_quantities) out.  Its scratch (per-cell partial sums, the three sums,
   the QGauss table) is one static 6.3 MB device array per GPU: at most 262144 cells and
   p + 1 + overint <= 24 (DGX_ERR_INVALID beyond).  Calls that overlap on one GPU (different
   streams) race on that array. */
int dgx_integrated_quantities_static(dgx_ctx *ctx, const double *u_dev, int overint, double *out4_dev);
/* PeriodicTurbulence::compute_current_integrated_numerical_entropy (periodic_turbulence.cpp:765-868):
   sum over cells and GLL volume nodes of W det J (-rho ln(p rho^-gamma)), FP64 on the device. */
int dgx_numerical_entropy(dgx_ctx *ctx, double *out);
/* PeriodicTurbulence::update_maximum_local_wave_speed (periodic_turbulence.cpp:371-422):
   max of |u| + c over the QGauss(p+1+overint) points of every cell. */
int dgx_max_wave_speed(dgx_ctx *ctx, int overint, double *out);
/* PeriodicTurbulence::get_adaptive_time_step (periodic_turbulence.cpp:359-367):
   dt = cfl * L / (n_dofs/nstate)^(1/dim) / max wave speed (overint 10, as the reference).
   DGX_ERR_NONPHYSICAL if the maximum wave speed is not a positive finite number. */
int dgx_adaptive_time_step(dgx_ctx *ctx, double cfl, double *dt);

/* Blocks until the context's stream has finished all launched work. */
int dgx_synchronize(dgx_ctx *ctx);

/* Non-physical states (Parameters::AllParameters::non_physical_behavior, all_parameters.h:304,
   all_parameters.cpp:351-356; PhysicsBase::handle_non_physical_result, physics.cpp:298-317;
   Euler::check_positive_quantity, euler.cpp:155-173).  The stage kernels flag every stage state
   with a negative or NaN density or pressure.  The flag is read (and cleared) by
   dgx_assemble_residual, dgx_step_in_time (once per call, after its steps), dgx_batch_synchronize,
   a served request's retirement and dgx_check_state, and handled by the context's policy:
     DGX_RETURN_BIG_NUMBER  (reference default) counted in dgx_nonphysical_events, status DGX_OK;
                            the device does not substitute 1e100 for the offending pressure, so the
                            affected residual entries are NaN where the reference's are ~1e100;
     DGX_ABORT_RUN          the call returns DGX_ERR_NONPHYSICAL (the reference std::abort()s; the
                            caller decides); dgx_assemble_residual then fills right_hand_side with
                            ones, the reference's invalid-assembly fallback (dg_base.cpp:2850-2875);
     DGX_PRINT_WARNING      a warning on stderr, status DGX_OK.
   A served request that met a non-physical state makes dgx_server_result return
   DGX_ERR_NONPHYSICAL (its outputs are still filled). */
enum dgx_non_physical_behavior { DGX_RETURN_BIG_NUMBER = 0, DGX_ABORT_RUN = 1, DGX_PRINT_WARNING = 2 };
int dgx_set_non_physical_behavior(dgx_ctx *ctx, int policy);
/* Waits for the context stream, reads and clears the flag, applies the policy (for the
   device-pointer entry points, which do not wait). */
int dgx_check_state(dgx_ctx *ctx);
/* Number of checks so far that found a non-physical state (all policies), plus adaptive time
   steps refused for a non-positive or non-finite maximum wave speed. */
int64_t dgx_nonphysical_events(const dgx_ctx *ctx);
/* HIP stream the context launches on (as an opaque pointer); 0 = default stream. */
int dgx_set_stream(dgx_ctx *ctx, void *hip_stream);
/* Average duration (ms) of the last `dgx_step_in_time` stage kernels, measured with
   HIP events on the context stream when timing is enabled. */
int dgx_enable_timing(dgx_ctx *ctx, int enable);
int dgx_kernel_timing(const dgx_ctx *ctx, double *avg_stage_ms, int64_t *n_stage_launches);

/* ---------------------------------------------------------------------------
   Batched independent solves.  No counterpart in the reference, which advances
   independent solves one FlowSolver at a time: these entry points replace a loop
   over instances of ODESolverBase::step_in_time (ode_solver_base.h:70) with one
   HIP graph launch per RK step that advances every instance.  Instances of the
   split-form Euler scheme that share (dim, p, precision) are advanced by one
   batched kernel launch per stage; other schemes by one launch per instance
   inside the same graph.  Each instance's arithmetic is the one of its single
   solve, so its result equals the single solve's.
   --------------------------------------------------------------------------- */
typedef struct dgx_batch dgx_batch;
/* n instances, params[i] describing instance i (dim, p, cells, scheme, precision may differ). */
int dgx_batch_create(const dgx_params *params, int n, dgx_batch **out);
int dgx_batch_destroy(dgx_batch *b);
const char *dgx_batch_last_error(const dgx_batch *b);
int dgx_batch_size(const dgx_batch *b);
/* Context of instance i (set/get solution, initial conditions, residual, integrated quantities);
   owned by the batch. */
dgx_ctx *dgx_batch_instance(dgx_batch *b, int i);
/* Allocates every instance's RK registers for an explicit Butcher method with at least two
   stages (Heun, SSPRK3, RK4) and the time step dt[i] of each instance, picks the tile of each
   (dim, p, precision) group (elements per block x waves/SIMD bound x threads per block; 12
   candidates in 2D, 2 in 3D)
   by timing every candidate on the group's own instances (cached per process), and captures one
   RK step of all instances in a HIP graph.  No tile changes an element's arithmetic: each
   instance's result is its single solve's, bit for bit. */
int dgx_batch_allocate(dgx_batch *b, int method, const double *dt);
/* Advances every instance by n_steps steps (one graph launch per step); per-instance
   current_time / current_iteration advance as in dgx_step_in_time. */
int dgx_batch_step_in_time(dgx_batch *b, int n_steps);
int dgx_batch_synchronize(dgx_batch *b);
/* Forces candidate tile `tile` (0..11; 2D shapes take all twelve, 3D the first two) for every
   group of the next dgx_batch_allocate whose shape has it, instead of timing the candidates;
   -1 restores the autotuner.  For reproducible launch shapes and for tests: every tile gives
   the same results bit for bit. */
int dgx_batch_set_tile(dgx_batch *b, int tile);
/* Tile chosen for instance i's group, as 1 + its candidate index; 0 if the instance is launched
   on its own (schemes other than split-form Euler). */
int dgx_batch_variant(const dgx_batch *b, int i);
/* The same tile as (elements per block, waves/SIMD bound with 0 = the compiler's allocation,
   threads per block); all 0 for an instance launched on its own. */
int dgx_batch_tile(const dgx_batch *b, int i, int *elements_per_block, int *waves_bound, int *threads_per_block);

/* ---------------------------------------------------------------------------
   Serving: a request loop over independent parametric solves.  No counterpart
   in the reference (each PHiLiP run is one FlowSolver::run, flow_solver.cpp:
   552-700).  Requests queue on submit; each scheduler step admits queued
   requests while their HBM footprint fits the budget and a slot is free
   (first come, first served), advances every active request by one RK step
   with the stages of all active requests batched together (one launch per
   (dim, p, precision) group per stage for the split-form Euler scheme), and
   retires requests that reached their final time (solution copied to the
   host, latency = completion - submission).  The time loop of each request is
   FlowSolver::run's: constant dt, last step clipped to the final time
   (flow_solver.cpp:554-561).
   --------------------------------------------------------------------------- */
enum { DGX_IC_TGV = 0, DGX_IC_VORTEX = 1 };
typedef struct {
    dgx_params params;
    int initial_condition;                   /* DGX_IC_TGV or DGX_IC_VORTEX */
    int tgv_isothermal;
    double vortex_strength, vortex_u, vortex_v;
    double final_time;
    /* dt = cfl * dx / wave_speed with dx = L / (n_dofs / nstate)^(1/dim)
       (periodic_turbulence.cpp:115-125 with wave_speed = 1, periodic_entropy_tests.cpp:22-36) */
    double cfl, wave_speed;
} dgx_request;

typedef struct dgx_server dgx_server;
/* hbm_budget_bytes: device memory the active requests may hold; max_active: slots;
   method: explicit Butcher method with at least two stages. */
int dgx_server_create(int64_t hbm_budget_bytes, int max_active, int method, dgx_server **out);
int dgx_server_destroy(dgx_server *s);
const char *dgx_server_last_error(const dgx_server *s);
/* HBM bytes a request holds while active (solution, residual, RK registers, connectivity). */
int64_t dgx_request_bytes(const dgx_request *r, int method);
int64_t dgx_request_n_dofs(const dgx_request *r);
/* Queues a request (returns at once); *id identifies it.  DGX_ERR_INVALID if it can never
   fit the budget or its parameters are rejected by dgx_create. */
int dgx_server_submit(dgx_server *s, const dgx_request *r, int64_t *id);
/* One scheduler step (admit, advance all active requests by one RK step, retire); blocks
   until the step's device work is done.  Counts after the step are returned. */
int dgx_server_step(dgx_server *s, int *n_active, int *n_queued, int *n_done);
/* Result of a finished request: solution (n_dofs doubles), seconds from submit to completion,
   RK steps taken, DOF-updates performed.  DGX_ERR_STATE while not finished. */
int dgx_server_result(dgx_server *s, int64_t id, double *host_u, double *latency_s, int64_t *steps,
                      double *dof_updates);
/* Drops a finished request's stored result. */
int dgx_server_release(dgx_server *s, int64_t id);
/* HBM bytes held by the active requests and the budget. */
int dgx_server_memory(const dgx_server *s, int64_t *in_use, int64_t *budget);

#ifdef __cplusplus
}
#endif
#endif
