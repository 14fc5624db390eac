/*
 * dg_oracle.h -- CPU FP64 restatement of PHiLiP's explicit strong-form DG
 * residual and explicit Runge-Kutta step, used ONLY as a test checker and as
 * the reported CPU baseline.  Nothing in the product library links this.
 *
 * Reference behaviour restated (file:line in PHiLiP):
 *   residual assembly        src/dg/dg_base.cpp:2599-2914, :453-870 (cell loop,
 *                            face owner rule :417-449, penalty :397-413)
 *   volume term (strong)     src/dg/strong_dg.cpp:1017-1579
 *   face term (strong)       src/dg/strong_dg.cpp:2368-3576
 *   auxiliary equation       src/dg/strong_dg.cpp:498-1008
 *   inverse mass             src/dg/dg_base.cpp:4089-4260
 *   1D operators / sum fact. src/operators/operators.cpp:294-649, :931-1196, :1217-2005
 *   Euler physics            src/physics/euler.cpp:177-937, :1033-1062
 *   boundary faces (Euler)   src/physics/euler.cpp:1097-1567 (boundary_face_values),
 *                            src/dg/strong_dg.cpp:1583-2364 (assemble_boundary_term_strong)
 *   Navier-Stokes viscous    src/physics/navier_stokes.cpp:240-385, :700-830, :1197-1350
 *   numerical fluxes         src/numerical_flux/convective_numerical_flux.cpp:34-480,
 *                            src/numerical_flux/viscous_numerical_flux.cpp:144-328
 *   RK / low-storage RK      src/ode_solver/runge_kutta_ode_solver.cpp:15-110,
 *                            src/ode_solver/low_storage_runge_kutta_ode_solver.cpp:24-222,
 *                            src/ode_solver/runge_kutta_methods/ (all tableaus)
 *   TGV / vortex / dt / KE   src/physics/initial_conditions/initial_condition_function.cpp:261-336,:643-690,
 *                            src/flow_solver/flow_solver_cases/periodic_turbulence.cpp:115-125,:439-671,
 *                            src/flow_solver/flow_solver.cpp:554-610
 *
 * Vector layout is the reference's DoF order without renumbering:
 *   u[(cell*nstate + state)*Np + node], node lexicographic (x fastest),
 *   cells in deal.II active-cell order (Z/Morton order for power-of-two
 *   cells per direction after refine_global, lexicographic otherwise).
 */
#ifndef DG_ORACLE_H
#define DG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define DGO_MAXN 16

enum { DGO_GLL = 0, DGO_GL = 1 };
enum { DGO_KG = 0, DGO_IR = 1, DGO_CH = 2, DGO_RA = 3 };
enum {
    DGO_LXF = 0, DGO_ROE = 1, DGO_CENTRAL = 2, DGO_TWO_POINT = 3,
    DGO_TWO_POINT_LXF = 4, DGO_TWO_POINT_ROE = 5
};
enum { DGO_EULER = 0, DGO_NAVIER_STOKES = 1 };
enum { DGO_SIPG = 0, DGO_BR2 = 1 };
enum { DGO_ORDER_DEALII = 0, DGO_ORDER_LEXICOGRAPHIC = 1 };
/* planted mistakes (tests of the tests only): LxF wave speed min instead of max, LxF or Roe
   dissipation without its factor 1/2, Roe without Harten's entropy fix, Roe without the shear
   (tangential velocity jump) wave, the curvilinear two-point flux without its metric average, the 3D
   cofactor without the conservative curl form, the QGauss-point metric of the integrated quantities
   transposed or replaced by the box's det J */
enum {
    DGO_MUT_NONE = 0, DGO_MUT_LXF_MIN = 1, DGO_MUT_LXF_NO_HALF = 2, DGO_MUT_ROE_NO_HALF = 3,
    DGO_MUT_ROE_NO_HARTEN = 4, DGO_MUT_ROE_NO_SHEAR = 5,
    DGO_MUT_NO_METRIC_AVERAGE = 6,  /* curvilinear split form: the node's own cofactor in place of the
                                       two-point average */
    DGO_MUT_CROSS_PRODUCT_METRIC = 7, /* 3D cofactor det(J) J^-T pointwise instead of the curl form */
    DGO_MUT_TRANSPOSED_METRIC = 8,    /* integrated quantities on curved cells: cofactor transposed */
    DGO_MUT_CELL_CONSTANT_DETJ = 9    /* integrated quantities on curved cells: the box's det J */
};

typedef struct {
    int dim;                 /* 2 or 3 */
    int poly_degree;         /* p */
    int flux_nodes;          /* DGO_GLL / DGO_GL */
    int use_split_form;
    int two_point_flux;      /* DGO_KG.. */
    int conv_flux;           /* DGO_LXF.. */
    int pde;                 /* DGO_EULER / DGO_NAVIER_STOKES */
    int diss_flux;           /* DGO_SIPG / DGO_BR2 */
    int use_constant_viscosity;
    double gamma_gas;
    double mach_inf;
    double reynolds_inf;
    double prandtl;
    double constant_viscosity;
    double temperature_inf;
    double sipg_penalty_factor;
    int n_cells_1d;
    double domain_left;
    double domain_right;
    int cell_order;          /* DGO_ORDER_DEALII / DGO_ORDER_LEXICOGRAPHIC */
    /* boundary condition of each face of the box (deal.II face numbering: 2k = -e_k, 2k+1 = +e_k):
       0 = periodic (both faces of a direction), else a boundary type of
       Euler::boundary_face_values (euler.cpp:1511-1567): 1001 wall, 1002 pressure outflow,
       1003 inflow, 1004 Riemann farfield, 1005 farfield, 1006 slip wall, 1007 extrapolation */
    int boundary_type[6];
    double angle_of_attack;  /* radians; free stream (cos a, sin a[, 0]) */
    int grid_warp;           /* 0: the affine box; 1: the reference's curved periodic grid
                                (NonsymmetricCurvedGridManifold, nonsymmetric_curved_periodic_grid.cpp:66-87,
                                beta = 0.1) applied to the box's nodes; GLL, Euler */
    int thermal_bc;          /* Navier-Stokes no-slip walls (1001): 0 adiabatic, 1 isothermal
                                (thermal_boundary_condition_type, parameters_navier_stokes.cpp:24-50) */
    double wall_temperature; /* nondimensionalized_isothermal_wall_temperature (default 1) */
} dgo_params;

typedef struct dgo_ctx dgo_ctx;

void dgo_default_params(dgo_params *p);
/* Euler::boundary_face_values (euler.cpp:1511-1567) for the boundary types above: the exterior
   state of a boundary face node from the interior state and the outward unit normal. */
void dgo_boundary_state(const dgo_ctx *c, int type, const double *u_int, const double *normal, double *u_bc);
dgo_ctx *dgo_create(const dgo_params *p);
void dgo_destroy(dgo_ctx *c);

long dgo_n_dofs(const dgo_ctx *c);
int dgo_n_cells(const dgo_ctx *c);

/* 1D nodes and weights on [0,1] */
void dgo_gauss_legendre(int n, double *x, double *w);
void dgo_gauss_lobatto(int n, double *x, double *w);
/* Lagrange basis on support points s (n), values/derivatives at x: out[j] */
void dgo_lagrange_values(int n, const double *s, double x, double *val, double *der);

/* operators used by the context (row-major, [rows][DGO_MAXN] packed as rows*n) */
void dgo_get_1d_operators(const dgo_ctx *c, double *nodes, double *weights,
                          double *chi, double *chi_grad, double *flux_grad,
                          double *skew, double *proj, double *minv,
                          double *chi_surf, double *flux_surf);

/* sum-factorized apply y = (Az x Ay x Ax) x ; dim 2 ignores Az */
void dgo_tensor_apply(int dim, const double *Ax, int rx, int cx,
                      const double *Ay, int ry, int cy,
                      const double *Az, int rz, int cz,
                      const double *in, double *out, int adding, double factor);

/* Hadamard sparsity patterns (integers): rows/cols sized n^dim*n*dim */
void dgo_hadamard_volume_pattern(int dim, int n, unsigned *rows, unsigned *cols);
void dgo_hadamard_surface_pattern(int dim, int n, int dim_not_zero, unsigned *rows, unsigned *cols);

/* mesh */
void dgo_cell_ijk(const dgo_ctx *c, int *ijk);            /* [ncells][3] */
void dgo_neighbors(const dgo_ctx *c, int *nbr);            /* [ncells][2*dim] */
void dgo_face_list(const dgo_ctx *c, int *faces);          /* [nfaces][4]: owner, iface, nbr, nbr_iface */
int  dgo_n_faces(const dgo_ctx *c);
void dgo_node_coords(const dgo_ctx *c, double *xyz);
/* curvilinear grid from the caller's node coordinates [cell][node][3] (mapping support points of
   grid degree p at the GLL nodes); 0, or -1 if the discretisation is not GLL Euler */
int dgo_set_grid_nodes(dgo_ctx *c, const double *xyz);
/* metric cofactor C[m][n] = J dxi_n/dx_m [cell][node][3][3] and det J [cell][node] */
void dgo_metric_terms(const dgo_ctx *c, double *C, double *J);       /* [ncells][Np][3] */

/* point physics (nstate = dim+2) */
double dgo_pressure(const dgo_ctx *c, const double *u);
void dgo_convective_flux(const dgo_ctx *c, const double *u, double *f /*[ns][dim]*/);
void dgo_entropy_variables(const dgo_ctx *c, const double *u, double *v);
void dgo_conservative_from_entropy(const dgo_ctx *c, const double *v, double *u);
void dgo_two_point_flux(const dgo_ctx *c, int which, const double *u1, const double *u2, double *f);
double dgo_log_mean(double a, double b);
void dgo_numerical_flux(const dgo_ctx *c, int which, const double *ui, const double *ue,
                        const double *n, double *f);
void dgo_dissipative_flux(const dgo_ctx *c, const double *u, const double *grad /*[ns][dim]*/,
                          double *f /*[ns][dim]*/);
double dgo_max_convective_eigenvalue(const dgo_ctx *c, const double *u);

/* residual, inverse mass */
void dgo_assemble_residual(dgo_ctx *c, const double *u, double *rhs);
void dgo_apply_inverse_mass(const dgo_ctx *c, const double *in, double *out, int aux);
void dgo_get_auxiliary(const dgo_ctx *c, int idim, double *out);

/* initial conditions */
void dgo_ic_tgv(const dgo_ctx *c, int isothermal, double *u);
void dgo_ic_vortex(const dgo_ctx *c, double strength, double u_inf, double v_inf, double *u);

/* time integration: method ids 0 euler,1 heun,2 ssprk3,3 rk4,
   10 RK4(3)5[3S*], 11 RK3(2)5F[3S*+], 12 RK4(3)9F[3S*+], 13 RK5(4)10F[3S*+] */
int dgo_n_stages(int method);
void dgo_step(dgo_ctx *c, int method, double dt, double *u);
/* one low-storage step (methods 10..13) that also returns the error norm of its embedded pair
   (low_storage_runge_kutta_ode_solver.cpp:89-118, before the epsilon/PID update) */
double dgo_step_error_norm(dgo_ctx *c, int method, double dt, double *u, double atol, double rtol);
/* relaxation RK (root finding on the integrated numerical entropy) for a Butcher method:
   *gamma in = previous relaxation parameter, out = this step's; returns gamma*dt (-1 on failure) */
double dgo_step_rrk(dgo_ctx *c, int method, double dt, double *u, double *gamma, double tol);
/* get_automatic_initial_step_size (low_storage_runge_kutta_ode_solver.cpp:119-178) */
double dgo_initial_step_size(dgo_ctx *c, const double *u);
double dgo_constant_time_step(const dgo_ctx *c, double cfl);
double dgo_vortex_time_step(const dgo_ctx *c, double cfl, double speed);

/* integrated quantities with QGauss(p+1+overint) : out[0]=KE, [1]=enstrophy,
   [2]=incompressible KE, [3]=vorticity-based dissipation rate */
void dgo_integrated_quantities(const dgo_ctx *c, const double *u, int overint, double *out);
/* numerical entropy  sum_i w_i J_i (-rho s)/(gamma-1) at volume nodes (mass-lumped, reference
   periodic_entropy_tests style uses M-weighted v.u; see oracle source) */
double dgo_entropy_rate(dgo_ctx *c, const double *u);
/* integrated numerical entropy  sum w_q detJ (-rho ln(p rho^-gamma))  (periodic_turbulence.cpp:765-868) */
double dgo_numerical_entropy(const dgo_ctx *c, const double *u);
/* max |u| + c over QGauss(p+1+overint) points (periodic_turbulence.cpp:371-422) and the
   adaptive step cfl * dx / max speed (periodic_turbulence.cpp:359-367) */
double dgo_max_wave_speed(const dgo_ctx *c, const double *u, int overint);
double dgo_adaptive_time_step(const dgo_ctx *c, const double *u, double cfl);

/* full flow-solver loop; returns number of steps; ke_hist gets KE after every step (may be NULL) */
int dgo_run(dgo_ctx *c, int method, double dt_const, double final_time, double *u,
            double *ke_hist, int max_hist);

void dgo_set_threads(int n);
/* planted mistake for the tests of the oracle's checks (DGO_MUT_*; DGO_MUT_NONE restores) */
void dgo_set_mutation(int m);
/* 3S*+ methods: whether dgo_step evaluates the FSAL residual of the embedded estimate
   (default 1, as the reference); it does not change the solution */
void dgo_set_fsal(dgo_ctx *c, int on);
/* per-DoF rounding scale of the residual: sum of |terms| summed into each DoF (GLL only) */
void dgo_residual_magnitude(dgo_ctx *c, const double *u, double *mag);

#ifdef __cplusplus
}
#endif
#endif
