// Host-side pieces of the device path: 1D Gauss-Lobatto operators, periodic
// Cartesian mesh numbering/connectivity, and the context that owns device
// buffers.  Mirrors what DGBase builds at construction (dg_base.cpp:142-261
// FE/quadrature collections, straight_periodic_cube.cpp:14-67 grid).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
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

// Cell order of straight_periodic_cube: Z order (refine_global) when n_cells_1d
// is a power of two, lexicographic (subdivided_hyper_cube) otherwise.
void cell_ijk(int dim, int n1d, int order, std::vector<int> &ijk);
void neighbors(int dim, int n1d, const std::vector<int> &ijk, std::vector<int> &nbr);

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
    std::vector<int> nbr_host;
    std::vector<int> cell_lat;     // [cell][3] lattice coordinates
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
    bool batch_locked = false;     // RK registers captured in a batch graph: not re-allocatable
    // integrated-quantity operators (QGauss(p+1+overint))
    int iq_overint = -1;
    double *iq_ops = nullptr;      // B [nx][n], G [nx][n], w [nx]
    double *iq_partial = nullptr;  // per-block partial sums
    int iq_blocks = 0;
};

// kernel launch helpers (dg_launch.hip)
namespace dgx {
int launch_stage(dgx_ctx *c, const StageArgs &a);
int launch_inverse_mass(dgx_ctx *c, const double *in, double *out);
int launch_fill(dgx_ctx *c, double *out, int64_t n, double value);
// Batched split-form line kernel over the element groups `blocks` of several solves of one
// (dim, n, precision); es = elements-per-block variant (1 or 2).  With blocks == nullptr only
// reports the variant's elements per block in *epb.  DGX_ERR_UNSUPPORTED if the variant does
// not exist for this shape.
int launch_line_batch(int dim, int n, int precision, int es, const BatchBlock *blocks, int nblocks,
                      const StageArgs *args, const PhysConst *pcs, const Ops1D &ops, hipStream_t st, int *epb);
int launch_aux(dgx_ctx *c, const double *u);
int launch_integrated(dgx_ctx *c, const double *u, int overint, double *host_out4);
int launch_entropy(dgx_ctx *c, const double *u, double *host_out);
int launch_wave_speed(dgx_ctx *c, const double *u, int overint, double *host_out);
}
