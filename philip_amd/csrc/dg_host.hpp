// Host-side pieces of the device path: 1D Gauss-Lobatto operators, periodic
// Cartesian mesh numbering/connectivity, and the context that owns device
// buffers.  Mirrors what DGBase builds at construction (dg_base.cpp:142-261
// FE/quadrature collections, straight_periodic_cube.cpp:14-67 grid).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/philip_dg.h"
#include "dg_kernels.hpp"
#include "dg_line.hpp"

namespace dgx {

// Gauss-Lobatto nodes/weights on [0,1] (the support points of deal.II FE_DGQ(p)
// and QGaussLobatto(p+1)); Newton on P'_N with P''_N from the Legendre ODE.
void gauss_lobatto(int n, double *x, double *w);
// Gauss-Legendre nodes/weights on [0,1] (QGauss), used for over-integration.
void gauss_legendre(int n, double *x, double *w);
// Lagrange basis on support points s evaluated at x: values and derivatives.
void lagrange(int n, const double *s, double x, double *val, double *der);
// Operators of a collocated GLL element of n nodes.
Ops1D build_ops(int n);


// First-fit sub-allocator over one device allocation (a server's HBM budget): request
// buffers are carved from it and returned to it, so admitting and retiring requests calls
// neither hipMalloc nor hipFree.  Free ranges are kept coalesced.  Not thread-safe.
class Arena {
public:
    int init(int64_t bytes);            // one hipMalloc; DGX_OK or DGX_ERR_HIP
    void release();                     // hipFree
    void *alloc(int64_t bytes);         // 256-byte aligned; nullptr when no free range fits
    void free(void *p);
    int64_t in_use() const { return used_bytes_; }
    int64_t capacity() const { return size_; }

private:
    char *base_ = nullptr;
    int64_t size_ = 0, used_bytes_ = 0;
    std::map<int64_t, int64_t> free_;   // offset -> length
    std::map<int64_t, int64_t> used_;   // offset -> length
};

struct Timing {
    bool enabled = false;
    hipEvent_t a = nullptr, b = nullptr;
    double sum_ms = 0.0;
    int64_t n = 0;
};

} // namespace dgx

struct dgx_ctx {
    dgx_params P{};
    int dim = 3, p = 0, n = 0, Np = 0, ns = 0, ncells = 0;
    int64_t ndofs = 0;
    double h = 0.0;
    dgx::Ops1D ops{};
    dgx::PhysConst pc{};
    int64_t nbr_count = 0;         // ncells * 2 * dim
    bool has_boundary = false;     // some box face is not periodic (nbr < 0 there: -boundary type)
    bool curv = false;             // curvilinear grid: metric terms per node in met_dev
    double *xn_dev = nullptr;      // node coordinates [cell][node][3] (allocated on demand)
    double *met_dev = nullptr;     // [cell][dim*dim + 1][node]: cofactor C[m][n] (slot m*dim+n), det J
    double *pen_dev = nullptr;     // [cell][dim]: SIPG penalty of the cell's faces normal to each direction
    int *nbr_dev = nullptr;
    double *u_dev = nullptr;       // DGBase::solution
    double *r_dev = nullptr;       // DGBase::right_hand_side
    double *aux_dev[3] = {nullptr, nullptr, nullptr};   // lifted gradient, FP64 path
    float *auxf_dev[3] = {nullptr, nullptr, nullptr};   // lifted gradient, mixed path
    double *minv_dev = nullptr;    // lifted right-hand side (dgx_apply_inverse_mass_to_rhs)
    int *status_dev = nullptr;     // non-physical state flag written by the stage kernels
    int nonphys_policy = 0;        // DGX_RETURN_BIG_NUMBER / DGX_ABORT_RUN / DGX_PRINT_WARNING
    int64_t nonphys_events = 0;    // checks that found a non-physical state
    // ODE system
    int method = -1;
    bool track_error = false;
    std::vector<double *> regs;    // RK registers
    double current_time = 0.0;
    int64_t iteration = 0;
    int lsrk_cur = 0;              // which S1 buffer holds the solution
    hipStream_t stream = nullptr;
    std::string err;
    dgx::Timing timing;
    int device = 0;
    int n_cu = 0;                  // compute units of the device
    dgx::Arena *arena = nullptr;   // set: solution, residual, connectivity, gradient and RK
                                   // registers come from this arena (serving), else hipMalloc
    bool batch_locked = false;     // RK registers captured in a batch graph: not re-allocatable
    // integrated-quantity operators (QGauss(p+1+overint))
    int iq_overint = -1;
    double *iq_ops = nullptr;      // B [nx][n], G [nx][n], w [nx]
    double *iq_partial = nullptr;  // per-block partial sums
    int iq_blocks = 0;
    // error-adaptive low-storage RK (low_storage_runge_kutta_ode_solver.cpp:89-118)
    double *vr_partial = nullptr;  // per-block partials of the vector reductions
    double atol = 1e-3, rtol = 1e-3, beta_pid[3] = {0.70, -0.23, 0.0};
    double epsilon[3] = {1.0, 1.0, 1.0};
    // relaxation RK (rrk_ode_solver_base.cpp, root_finding_rrk_ode_solver.cpp)
    bool rrk = false;
    double rrk_gamma = 1.0;        // relaxation parameter of the last step (1 before the first)
    double rrk_tol = 5e-10;        // relaxation_runge_kutta_root_tolerance
    double *rrk_partial = nullptr; // per-stage block partials of the entropy-change estimate
};

// kernel launch helpers (dg_launch.hip)
namespace dgx {
int launch_stage(dgx_ctx *c, const StageArgs &a);
int launch_inverse_mass(dgx_ctx *c, const double *in, double *out);
int launch_fill(dgx_ctx *c, double *out, int64_t n, double value);
int launch_mesh(dgx_ctx *c);
// node coordinates of the box's cells into xn [cell][node][3]; warp = 1: through the reference's
// curved periodic grid mapping
int launch_coords(dgx_ctx *c, double *xn, int warp);
// metric terms of c->xn_dev into c->met_dev, and the cells' penalties into c->pen_dev
int launch_metrics(dgx_ctx *c);
int launch_ic(dgx_ctx *c, int kind, int isothermal, double strength, double u_inf, double v_inf);
int launch_maxabs(dgx_ctx *c, const double *a, const double *b, int64_t n, double *out);
int launch_axpy(dgx_ctx *c, const double *x, double h, const double *f, double *out, int64_t n);
int launch_lsrk_error(dgx_ctx *c, const double *s1, const double *e, const double *s3, int plus, double d5, double d6,
                      double sum_delta, double atol, double rtol, double *sum_out);
// Batched split-form line kernel over the element groups `blocks` of several solves of one
// (dim, n, precision); tile = index of the candidate tile (batch_tile: elements per block x
// waves/SIMD bound x threads per block).  With blocks == nullptr only reports the tile's elements
// per block in *epb and threads per block in *threads (either may be null).  DGX_ERR_UNSUPPORTED
// if the tile does not exist for this shape.
int launch_line_batch(int dim, int n, int precision, int tile, const BatchBlock *blocks, int nblocks,
                      const StageArgs *args, const PhysConst *pcs, const Ops1D &ops, hipStream_t st, int *epb,
                      int *threads = nullptr);
int launch_aux(dgx_ctx *c, const double *u);
int launch_integrated(dgx_ctx *c, const double *u, int overint, double *host_out4);
int launch_integrated_static(dgx_ctx *c, const double *u, int overint, double *out4_dev);
int launch_entropy(dgx_ctx *c, const double *u, double *host_out);
// integrated numerical entropy of u + gamma dir (dir == nullptr: of u)
int launch_entropy_combo(dgx_ctx *c, const double *u, const double *dir, double gamma, double *host_out);
int launch_rrk_estimate(dgx_ctx *c, const double *u, const double *k, double *partial_dev);
int fetch_partial_sum(dgx_ctx *c, const double *partial_dev, double *out);
int vr_blocks();
int launch_wave_speed(dgx_ctx *c, const double *u, int overint, double *host_out);
}
