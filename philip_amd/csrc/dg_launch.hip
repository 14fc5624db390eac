// Dispatch of the DG kernels over (dim, nodes per direction) to their per-N instantiations
// (dg_inst_n<N>.hip), launch bookkeeping, the inverse mass, and the integrated-quantity
// reductions (kinetic energy, enstrophy, numerical entropy, maximum wave speed).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "dg_dispatch.hpp"

namespace dgx {

DGX_KERNELS_N(extern, 2)
DGX_KERNELS_N(extern, 3)
DGX_KERNELS_N(extern, 4)
DGX_KERNELS_N(extern, 5)
DGX_KERNELS_N(extern, 6)
DGX_KERNELS_N(extern, 7)

#define DISPATCH_N(FN, ...)                                                     \
    switch (c->n) {                                                             \
    case 2: return (c->dim == 3) ? FN<3, 2>(__VA_ARGS__) : FN<2, 2>(__VA_ARGS__); \
    case 3: return (c->dim == 3) ? FN<3, 3>(__VA_ARGS__) : FN<2, 3>(__VA_ARGS__); \
    case 4: return (c->dim == 3) ? FN<3, 4>(__VA_ARGS__) : FN<2, 4>(__VA_ARGS__); \
    case 5: return (c->dim == 3) ? FN<3, 5>(__VA_ARGS__) : FN<2, 5>(__VA_ARGS__); \
    case 6: return (c->dim == 3) ? FN<3, 6>(__VA_ARGS__) : FN<2, 6>(__VA_ARGS__); \
    case 7: return (c->dim == 3) ? FN<3, 7>(__VA_ARGS__) : FN<2, 7>(__VA_ARGS__); \
    default: c->err = "unsupported polynomial degree"; return DGX_ERR_UNSUPPORTED; \
    }

int launch_stage(dgx_ctx *c, const StageArgs &a) { DISPATCH_N(launch_stage_dn, c, a) }

// ---------------------------------------------------------------------------
// Batched line kernel: one launch over the element groups of many solves of one
// (dim, p, precision); es selects the elements-per-block variant.
// ---------------------------------------------------------------------------
int launch_line_batch(int dim, int n, int precision, int tile, const BatchBlock *blocks, int nblocks,
                      const StageArgs *args, const PhysConst *pcs, const Ops1D &ops, hipStream_t st, int *epb,
                      int *threads)
{
#define LB_CASE(NN)                                                                                            \
    case NN:                                                                                                   \
        return (dim == 3) ? line_batch_dn<3, NN>(precision, tile, blocks, nblocks, args, pcs, ops, st, epb, threads) \
                          : line_batch_dn<2, NN>(precision, tile, blocks, nblocks, args, pcs, ops, st, epb, threads);
    switch (n) {
        LB_CASE(2)
        LB_CASE(3)
        LB_CASE(4)
        LB_CASE(5)
        LB_CASE(6)
        LB_CASE(7)
    default: return DGX_ERR_UNSUPPORTED;
    }
#undef LB_CASE
}

// det J of node i of cell e: the scalar of the affine box, or slot dim*dim of the curvilinear
// metric [cell][dim*dim + 1][node]
__device__ __forceinline__ double node_jac(const double *met, double detJ, int dim, int Np, long e, int i)
{
    return met ? met[(e * (dim * dim + 1) + dim * dim) * Np + i] : detJ;
}

// Inverse of the collocated GLL mass matrix (diagonal): out_q = in_q / (W_q det J),
// W_q the tensor-product GLL weight of the node (dg_base.cpp:4089-4260; on a curvilinear grid the
// weight-adjusted inverse, dg_base.cpp:4220-4247, is the same diagonal for collocated nodes).
__global__ void __launch_bounds__(256)
inverse_mass_kernel(const double *__restrict__ in, double *__restrict__ out, long ndofs, int n, int dim,
                    double detJ, Ops1D ops, const double *__restrict__ met)
{
    const int Np = (dim == 3) ? n * n * n : n * n, ns = dim + 2;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < ndofs; q += (long)gridDim.x * blockDim.x) {
        const int node = (int)(q % Np);
        const int a = node % n, b = (node / n) % n, c = node / (n * n);
        double W = ops.w[a] * ops.w[b];
        if (dim == 3) W *= ops.w[c];
        out[q] = in[q] / (W * node_jac(met, detJ, dim, Np, q / ((long)ns * Np), node));
    }
}

__global__ void __launch_bounds__(256) fill_kernel(double *__restrict__ out, long n, double value)
{
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) out[q] = value;
}

int launch_fill(dgx_ctx *c, double *out, int64_t n, double value)
{
    const long blocks = std::min<long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, c->stream, out, (long)n, value);
    return check_launch(c, "fill_kernel");
}

// ---------------------------------------------------------------------------
// Vector reductions of the error-adaptive low-storage RK step
// (low_storage_runge_kutta_ode_solver.cpp:89-178): per-block partials over a fixed grid, summed
// (or maxed) in a fixed order on the host, so results are deterministic.
// ---------------------------------------------------------------------------
constexpr int VR_BLOCKS = 1024, VR_THREADS = 256;

// partial[b] = max over block b's share of |a - b_| (b_ == nullptr: |a|)
__global__ void __launch_bounds__(VR_THREADS)
maxabs_kernel(const double *__restrict__ a, const double *__restrict__ b, long n, double *__restrict__ partial)
{
    __shared__ double red[VR_THREADS];
    double m = 0.0;
    for (long q = blockIdx.x * (long)VR_THREADS + threadIdx.x; q < n; q += (long)gridDim.x * VR_THREADS) {
        const double x = b ? fabs(a[q] - b[q]) : fabs(a[q]);
        m = x > m ? x : m;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int sh = VR_THREADS / 2; sh > 0; sh >>= 1) {
        if ((int)threadIdx.x < sh) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + sh]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// out = x + h f  (deal.II Vector::add(h, f) on a copy of x)
__global__ void __launch_bounds__(256) axpy_kernel(const double *__restrict__ x, double h, const double *__restrict__ f,
                                                   double *__restrict__ out, long n)
{
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
        out[q] = x[q] + h * f[q];
}

// partial[b] = block b's share of sum (err / (atol + rtol max(|S1|, |E|)))^2, err = S1 - E, with E the
// 3S*+ register S4, or (plus == 0) the embedded solution (S2 + d5 S1 + d6 S3) / sum_delta
__global__ void __launch_bounds__(VR_THREADS)
lsrk_error_kernel(const double *__restrict__ s1, const double *__restrict__ e, const double *__restrict__ s3, long n,
                  int plus, double d5, double d6, double sum_delta, double atol, double rtol,
                  double *__restrict__ partial)
{
    __shared__ double red[VR_THREADS];
    double acc = 0.0;
    for (long q = blockIdx.x * (long)VR_THREADS + threadIdx.x; q < n; q += (long)gridDim.x * VR_THREADS) {
        const double a = s1[q];
        double b = e[q];
        if (!plus) {
            b += d5 * a;
            b += d6 * s3[q];
            b /= sum_delta;
        }
        const double x = (a - b) / (atol + rtol * fmax(fabs(a), fabs(b)));
        acc += x * x;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int sh = VR_THREADS / 2; sh > 0; sh >>= 1) {
        if ((int)threadIdx.x < sh) red[threadIdx.x] += red[threadIdx.x + sh];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

static int vr_partials(dgx_ctx *c)
{
    if (!c->vr_partial && hipMalloc(&c->vr_partial, sizeof(double) * VR_BLOCKS) != hipSuccess) {
        c->err = "hipMalloc failed";
        return DGX_ERR_HIP;
    }
    return DGX_OK;
}

static int vr_fetch(dgx_ctx *c, double *host)
{
    if (hipMemcpyAsync(host, c->vr_partial, sizeof(double) * VR_BLOCKS, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        c->err = "reduction copy failed";
        return DGX_ERR_HIP;
    }
    return DGX_OK;
}

int launch_maxabs(dgx_ctx *c, const double *a, const double *b, int64_t n, double *out)
{
    int rc = vr_partials(c);
    if (rc) return rc;
    hipLaunchKernelGGL(maxabs_kernel, dim3(VR_BLOCKS), dim3(VR_THREADS), 0, c->stream, a, b, (long)n, c->vr_partial);
    rc = check_launch(c, "maxabs_kernel");
    if (rc) return rc;
    double h[VR_BLOCKS];
    rc = vr_fetch(c, h);
    if (rc) return rc;
    double m = 0.0;
    for (int k = 0; k < VR_BLOCKS; ++k) m = h[k] > m ? h[k] : m;
    *out = m;
    return DGX_OK;
}

int launch_axpy(dgx_ctx *c, const double *x, double h, const double *f, double *out, int64_t n)
{
    const long blocks = std::min<long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(axpy_kernel, dim3((unsigned)blocks), dim3(256), 0, c->stream, x, h, f, out, (long)n);
    return check_launch(c, "axpy_kernel");
}

int launch_lsrk_error(dgx_ctx *c, const double *s1, const double *e, const double *s3, int plus, double d5, double d6,
                      double sum_delta, double atol, double rtol, double *sum_out)
{
    int rc = vr_partials(c);
    if (rc) return rc;
    hipLaunchKernelGGL(lsrk_error_kernel, dim3(VR_BLOCKS), dim3(VR_THREADS), 0, c->stream, s1, e, s3, (long)c->ndofs,
                       plus, d5, d6, sum_delta, atol, rtol, c->vr_partial);
    rc = check_launch(c, "lsrk_error_kernel");
    if (rc) return rc;
    double h[VR_BLOCKS];
    rc = vr_fetch(c, h);
    if (rc) return rc;
    double acc = 0.0;
    for (int k = 0; k < VR_BLOCKS; ++k) acc += h[k];
    *sum_out = acc;
    return DGX_OK;
}

// ---------------------------------------------------------------------------
// Mesh connectivity and initial data on the device (straight_periodic_cube.cpp:14-67,
// set_initial_condition.cpp:45-51, initial_condition_function.cpp:261-336, 643-690), so that a
// context is set up without host-side loops or host-to-device copies of mesh-sized arrays.
// ---------------------------------------------------------------------------
struct MeshDesc {
    int dim, N, morton, levels;
    int bc[6];  // boundary type of each box face (0: periodic)
};

// lattice coordinates of cell e: Z order (refine_global, children x-fastest, coarsest level in
// the high bits) for 2^k cells per direction in deal.II order, lexicographic otherwise
__device__ __forceinline__ void cell_coords(const MeshDesc &m, int e, int *c)
{
    c[0] = c[1] = c[2] = 0;
    if (m.morton) {
        for (int bit = 0; bit < m.levels; ++bit)
            for (int k = 0; k < m.dim; ++k) c[k] |= ((e >> (bit * m.dim + k)) & 1) << bit;
    } else {
        c[0] = e % m.N;
        c[1] = (e / m.N) % m.N;
        if (m.dim == 3) c[2] = e / (m.N * m.N);
    }
}

__device__ __forceinline__ int cell_index(const MeshDesc &m, const int *c)
{
    if (!m.morton) return c[0] + m.N * (c[1] + (m.dim == 3 ? m.N * c[2] : 0));
    int e = 0;
    for (int bit = 0; bit < m.levels; ++bit)
        for (int k = 0; k < m.dim; ++k) e |= ((c[k] >> bit) & 1) << (bit * m.dim + k);
    return e;
}

// nbr[e][2k + side]: the periodic neighbour across face 2k (-e_k) / 2k+1 (+e_k)
__global__ void __launch_bounds__(256) mesh_kernel(int *__restrict__ nbr, int ncells, MeshDesc m)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ncells) return;
    int c[3];
    cell_coords(m, e, c);
    for (int f = 0; f < 2 * m.dim; ++f) {
        int q[3] = {c[0], c[1], c[2]};
        const int k = f / 2;
        if (m.bc[f] && q[k] == ((f & 1) ? m.N - 1 : 0)) {  // box boundary: -(boundary type)
            nbr[(long)e * 2 * m.dim + f] = -m.bc[f];
            continue;
        }
        q[k] = (q[k] + ((f & 1) ? 1 : m.N - 1)) % m.N;
        nbr[(long)e * 2 * m.dim + f] = cell_index(m, q);
    }
}

struct IcDesc {
    int kind;          // 0 TGV, 1 isentropic vortex
    int isothermal;
    double left, h, gam, mach2;
    double beta, u_inf, v_inf;
    double xs[MAXN1D]; // GLL support points on [0, 1]
};

// initial data interpolated at the GLL support points of each cell, one thread per node
// xn: node coordinates of a curvilinear grid [cell][node][3] (null: the affine box)
__global__ void __launch_bounds__(256) ic_kernel(double *__restrict__ u, int ncells, int n, MeshDesc m, IcDesc ic,
                                                 const double *__restrict__ xn)
{
    const int Np = (m.dim == 3) ? n * n * n : n * n, ns = m.dim + 2;
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= (long)ncells * Np) return;
    const int e = (int)(t / Np), i = (int)(t % Np);
    int c[3];
    cell_coords(m, e, c);
    const int id[3] = {i % n, (i / n) % n, i / (n * n)};
    double x[3];
    for (int k = 0; k < 3; ++k) x[k] = (k < m.dim) ? ic.left + (c[k] + ic.xs[id[k]]) * ic.h : 0.0;
    if (xn)
        for (int k = 0; k < 3; ++k) x[k] = xn[t * 3 + k];
    double cons[5];
    if (ic.kind == 0) {
        // InitialConditionFunction_TaylorGreenVortex[_Isothermal] (initial_condition_function.cpp:261-336)
        const double p = 1.0 / (ic.gam * ic.mach2) + (1.0 / 16.0) * (cos(2.0 * x[0]) + cos(2.0 * x[1])) *
                                                        (cos(2.0 * x[2]) + 2.0);
        const double rho = ic.isothermal ? p * ic.gam * ic.mach2 : 1.0;
        const double vx = sin(x[0]) * cos(x[1]) * cos(x[2]);
        const double vy = -cos(x[0]) * sin(x[1]) * cos(x[2]);
        const double vz = 0.0;
        cons[0] = rho;
        cons[1] = rho * vx;
        cons[2] = rho * vy;
        cons[3] = rho * vz;
        cons[4] = p / (ic.gam - 1.0) + 0.5 * rho * (vx * vx + vy * vy + vz * vz);
    } else {
        // isentropic vortex (initial_condition_function.cpp:643-690), strength and advection as parameters
        const double gam = 1.4;
        const double Om = ic.beta * exp(-0.5 * (x[0] * x[0] + x[1] * x[1]));
        const double dT = -(gam - 1.0) / 2.0 * Om * Om;
        const double rho = pow(1 + dT, 1.0 / (gam - 1.0));
        const double vx = ic.u_inf - x[1] * Om, vy = ic.v_inf + x[0] * Om;
        const double p = 1.0 / gam * pow(1 + dT, gam / (gam - 1.0));
        cons[0] = rho;
        cons[1] = rho * vx;
        cons[2] = rho * vy;
        cons[3] = p / (gam - 1.0) + 0.5 * rho * (vx * vx + vy * vy);
    }
    for (int s = 0; s < ns; ++s) u[((long)e * ns + s) * Np + i] = cons[s];
}

static MeshDesc mesh_desc(const dgx_ctx *c)
{
    MeshDesc m;
    m.dim = c->dim;
    m.N = c->P.n_cells_1d;
    m.morton = (c->P.cell_order == DGX_ORDER_DEALII) && m.N > 0 && (m.N & (m.N - 1)) == 0;
    m.levels = 0;
    while ((1 << m.levels) < m.N) ++m.levels;
    for (int f = 0; f < 6; ++f) m.bc[f] = (f < 2 * c->dim) ? c->P.boundary_type[f] : 0;
    return m;
}

int launch_mesh(dgx_ctx *c)
{
    hipLaunchKernelGGL(mesh_kernel, dim3((c->ncells + 255) / 256), dim3(256), 0, c->stream, c->nbr_dev, c->ncells,
                       mesh_desc(c));
    return check_launch(c, "mesh_kernel");
}

int launch_ic(dgx_ctx *c, int kind, int isothermal, double strength, double u_inf, double v_inf)
{
    IcDesc ic{};
    ic.kind = kind;
    ic.isothermal = isothermal;
    ic.left = c->P.domain_left;
    ic.h = c->h;
    ic.gam = c->P.gamma_gas;
    ic.mach2 = c->P.mach_inf * c->P.mach_inf;
    const double M = std::sqrt(2 / 1.4);
    ic.beta = M * strength * std::sqrt(2.0) / 4.0 / M_PI * std::exp(1.0 / 2.0);
    ic.u_inf = u_inf;
    ic.v_inf = v_inf;
    double ws[MAXN1D];
    gauss_lobatto(c->n, ic.xs, ws);
    const long n = (long)c->ncells * c->Np;
    hipLaunchKernelGGL(ic_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->u_dev, c->ncells,
                       c->n, mesh_desc(c), ic, c->curv ? c->xn_dev : nullptr);
    return check_launch(c, "ic_kernel");
}

// Node coordinates of the box's cells [cell][node][3] (the affine map of the GLL support points),
// optionally through the reference's curved periodic grid NonsymmetricCurvedGridManifold::mapping
// (nonsymmetric_curved_periodic_grid.cpp:66-87, beta = 0.1).
__global__ void __launch_bounds__(256) coords_kernel(double *__restrict__ xn, int ncells, int n, MeshDesc m,
                                                     IcDesc ic, int warp)
{
    const int Np = (m.dim == 3) ? n * n * n : n * n;
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= (long)ncells * Np) return;
    const int e = (int)(t / Np), i = (int)(t % Np);
    int c[3];
    cell_coords(m, e, c);
    const int id[3] = {i % n, (i / n) % n, i / (n * n)};
    double x[3];
    for (int k = 0; k < 3; ++k) x[k] = (k < m.dim) ? ic.left + (c[k] + ic.xs[id[k]]) * ic.h : 0.0;
    if (warp) {
        const double beta = 1.0 / 10.0, pi = 3.14159265358979323846;
        const double p0 = x[0], p1 = x[1], p2 = x[2];
        if (m.dim == 2) {
            x[0] = p0 + beta * cos(pi / 2.0 * p0) * cos(3.0 * pi / 2.0 * p1);
            x[1] = p1 + beta * sin(2.0 * pi * p0) * cos(pi / 2.0 * p1);
        } else {
            x[0] = p0 + beta * sin(p0) * sin(p1) * sin(2.0 * p2);
            x[1] = p1 + beta * sin(4.0 * p0) * sin(p1) * sin(3.0 * p2);
            x[2] = p2 + beta * sin(2.0 * p0) * sin(5.0 * p1) * sin(p2);
        }
    }
    for (int k = 0; k < 3; ++k) xn[t * 3 + k] = x[k];
}

// Metric terms at the GLL nodes from the node coordinates (mapping support points of grid degree
// p at the same nodes): Jacobian dX_a/dxi_b by the derivative operator along the node's lines,
// det J, and the cofactor C[m][n] = J dxi_n/dx_m -- in 2D from the Jacobian, in 3D by the
// conservative curl form of metric_operators::compute_local_3D_cofactor (operators.cpp:2737-2884):
// C[r][0] = d_zeta(P_eta) - d_eta(P_zeta), C[r][1] = d_xi(P_zeta) - d_zeta(P_xi),
// C[r][2] = d_eta(P_xi) - d_xi(P_eta) with P_j = X_l dX_m/dxi_j for (l, m) = (z, y), (x, z), (y, x),
// which satisfies the discrete metric identities.  Output [cell][dim*dim + 1][node].
template <int D, int N>
__global__ void __launch_bounds__(512) metric_kernel(const double *__restrict__ xn, double *__restrict__ met,
                                                     int ncells, Ops1D ops)
{
    constexpr int NP = (D == 3) ? N * N * N : N * N;
    constexpr int EPB = (NP >= 256) ? 1 : (256 / NP);
    constexpr int MS = D * D + 1;
    __shared__ double sX[EPB][D][NP];
    __shared__ double sPr[EPB][D == 3 ? 9 : 1][NP];
    const int le = threadIdx.x / NP, i = threadIdx.x % NP;
    const long e = (long)blockIdx.x * EPB + le;
    const bool active = le < EPB && e < ncells;
    if (active)
        for (int k = 0; k < D; ++k) sX[le][k][i] = xn[(e * NP + i) * 3 + k];
    __syncthreads();
    int c[3] = {0, 0, 0};
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    if (active) {
        node_coords<D, N>(i, c);
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b) {
                double acc = 0.0;
                for (int l = 0; l < N; ++l) acc += ops.D[c[b]][l] * sX[le][a][node_with<D, N>(i, b, l)];
                J[a][b] = acc;
            }
    }
    double C[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    if constexpr (D == 2) {
        C[0][0] = J[1][1];
        C[1][0] = -J[0][1];
        C[0][1] = -J[1][0];
        C[1][1] = J[0][0];
    } else {
        constexpr int LM[3][2] = {{2, 1}, {0, 2}, {1, 0}};
        if (active)
            for (int r = 0; r < 3; ++r)
                for (int j = 0; j < 3; ++j) sPr[le][r * 3 + j][i] = sX[le][LM[r][0]][i] * J[LM[r][1]][j];
        __syncthreads();
        if (active) {
            auto dref = [&](int slot, int dir) {
                double acc = 0.0;
                for (int l = 0; l < N; ++l) acc += ops.D[c[dir]][l] * sPr[le][slot][node_with<D, N>(i, dir, l)];
                return acc;
            };
            for (int r = 0; r < 3; ++r) {
                C[r][0] = -dref(r * 3 + 2, 1) + dref(r * 3 + 1, 2);
                C[r][1] = dref(r * 3 + 2, 0) - dref(r * 3 + 0, 2);
                C[r][2] = -dref(r * 3 + 1, 0) + dref(r * 3 + 0, 1);
            }
        }
    }
    if (!active) return;
    double det;
    if (D == 2) det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    else det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
               J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    for (int m = 0; m < D; ++m)
        for (int nn = 0; nn < D; ++nn) met[(e * MS + m * D + nn) * NP + i] = C[m][nn];
    met[(e * MS + D * D) * NP + i] = det;
}

int launch_coords(dgx_ctx *c, double *xn, int warp)
{
    IcDesc ic{};
    ic.left = c->P.domain_left;
    ic.h = c->h;
    double ws[MAXN1D];
    gauss_lobatto(c->n, ic.xs, ws);
    const long n = (long)c->ncells * c->Np;
    hipLaunchKernelGGL(coords_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, xn, c->ncells, c->n,
                       mesh_desc(c), ic, warp);
    return check_launch(c, "coords_kernel");
}

template <int D, int N> static void metric_dn(dgx_ctx *c)
{
    constexpr int NP = (D == 3) ? N * N * N : N * N;
    constexpr int EPB = (NP >= 256) ? 1 : (256 / NP);
    hipLaunchKernelGGL((metric_kernel<D, N>), dim3((c->ncells + EPB - 1) / EPB), dim3(EPB * NP <= 256 ? 256 : EPB * NP),
                       0, c->stream, c->xn_dev, c->met_dev, c->ncells, c->ops);
}

// SIPG penalty of a cell's faces normal to direction k on a curvilinear grid, as
// evaluate_penalty_scaling (dg_base.cpp:397-413): deg (deg + 1) / extent_in_direction(k) times the
// penalty factor, the extent being the longest of the cell's edges along k (between corner nodes).
__global__ void penalty_kernel(const double *__restrict__ xn, double *__restrict__ pen, int ncells, int n, int dim,
                               double scale)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)ncells * dim) return;
    const int cell = (int)(t / dim), k = (int)(t % dim);
    const long np = (dim == 3) ? (long)n * n * n : (long)n * n;
    double ext = 0.0;
    for (int a = 0; a < (1 << (dim - 1)); ++a) {
        int lo[3] = {0, 0, 0}, bit = 0;
        for (int m = 0; m < dim; ++m) {
            if (m == k) continue;
            lo[m] = ((a >> bit) & 1) ? n - 1 : 0;
            ++bit;
        }
        int hi[3] = {lo[0], lo[1], lo[2]};
        hi[k] = n - 1;
        const long q0 = lo[0] + n * (lo[1] + n * lo[2]), q1 = hi[0] + n * (hi[1] + n * hi[2]);
        const double *x0 = xn + ((long)cell * np + q0) * 3, *x1 = xn + ((long)cell * np + q1) * 3;
        double l2 = 0.0;
        for (int m = 0; m < 3; ++m) l2 += (x1[m] - x0[m]) * (x1[m] - x0[m]);
        ext = fmax(ext, sqrt(l2));
    }
    pen[t] = scale / ext;
}

int launch_metrics(dgx_ctx *c)
{
    if (c->pen_dev) {
        const unsigned deg = (unsigned)c->p;
        const double scale = (double)(deg == 0 ? 1 : deg * (deg + 1)) * c->P.sipg_penalty_factor;
        const long nt = (long)c->ncells * c->dim;
        hipLaunchKernelGGL(penalty_kernel, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c->stream, c->xn_dev,
                           c->pen_dev, c->ncells, c->n, c->dim, scale);
        const int rc = check_launch(c, "penalty_kernel");
        if (rc) return rc;
    }
#define MET_CASE(NN)                                                                                        \
    case NN:                                                                                                \
        if (c->dim == 3) metric_dn<3, NN>(c);                                                               \
        else metric_dn<2, NN>(c);                                                                           \
        break;
    switch (c->n) {
        MET_CASE(2) MET_CASE(3) MET_CASE(4) MET_CASE(5) MET_CASE(6) MET_CASE(7)
    default: c->err = "no metric kernel for this degree"; return DGX_ERR_UNSUPPORTED;
    }
#undef MET_CASE
    return check_launch(c, "metric_kernel");
}

int launch_inverse_mass(dgx_ctx *c, const double *in, double *out)
{
    const double detJ = (c->dim == 3) ? c->h * c->h * c->h : c->h * c->h;
    const long blocks = std::min<long>((c->ndofs + 255) / 256, 4096);
    hipLaunchKernelGGL(inverse_mass_kernel, dim3((unsigned)blocks), dim3(256), 0, c->stream, in, out, (long)c->ndofs,
                       c->n, c->dim, detJ, c->ops, c->curv ? c->met_dev : nullptr);
    return check_launch(c, "inverse_mass_kernel");
}

int launch_aux(dgx_ctx *c, const double *u)
{
    if (c->dim != 3) {
        c->err = "Navier-Stokes device path is 3D";
        return DGX_ERR_UNSUPPORTED;
    }
    switch (c->n) {
    case 2: return launch_aux_dn<3, 2>(c, u);
    case 3: return launch_aux_dn<3, 3>(c, u);
    case 4: return launch_aux_dn<3, 4>(c, u);
    case 5: return launch_aux_dn<3, 5>(c, u);
    case 6: return launch_aux_dn<3, 6>(c, u);
    case 7: return launch_aux_dn<3, 7>(c, u);
    default: c->err = "unsupported polynomial degree"; return DGX_ERR_UNSUPPORTED;
    }
}

// ---------------------------------------------------------------------------
// Integrated quantities (periodic_turbulence.cpp:439-671): one block per cell,
// QGauss(p+1+overint)^dim points, solution and physical gradient interpolated
// from the GLL nodal values; kinetic energy, enstrophy and incompressible KE
// reduced in FP64 (block partials, then a fixed-order host sum).
// ---------------------------------------------------------------------------
constexpr int IQ_THREADS = 256;

// Sum-factorised evaluation: per cell, density and momentum and their
// reference gradients at the QGauss(nx)^dim points are built direction by
// direction (x pass in LDS, y pass per quadrature column in registers, z pass
// per point), then 1/2 rho |u|^2, 1/2 rho |omega|^2 and 1/2 |u|^2 are
// accumulated with the tensor weights times det J.
template <int D, int N>
__global__ void __launch_bounds__(IQ_THREADS)
integrated_kernel(const double *__restrict__ u, const double *__restrict__ ops, int nx, int ncells, double inv_h,
                  double detJ, double *__restrict__ partial)
{
    constexpr int NS = D + 2;
    constexpr int NV = D + 1;  // density and momentum
    constexpr int n = N;
    extern __shared__ double smem[];
    constexpr int Np = (D == 3) ? n * n * n : n * n;
    constexpr int nr = (D == 3) ? n * n : n;  // nodes per x line
    double *sU = smem;                       // [NV][Np]
    double *sB = sU + NV * Np;               // [nx][n] values
    double *sG = sB + nx * n;                // [nx][n] derivatives
    double *sW = sG + nx * n;                // [nx]
    double *sT = sW + nx;                    // [NV][2][nx][nr]: x pass with B and G
    double *red = sT + NV * 2 * nx * nr;     // [3][IQ_THREADS]
    const int e = blockIdx.x;
    for (int t = threadIdx.x; t < NV * Np; t += blockDim.x) {
        const int sv = t / Np, i = t % Np;
        sU[t] = u[((long)e * NS + sv) * Np + i];
    }
    for (int t = threadIdx.x; t < 2 * nx * n + nx; t += blockDim.x) sB[t] = ops[t];
    __syncthreads();
    // x pass: T[s][k][qx][r] = sum_a {B,G}[qx][a] u[s][a + n r]
    for (int t = threadIdx.x; t < NV * 2 * nx * nr; t += blockDim.x) {
        const int r = t % nr, qx = (t / nr) % nx, k = (t / (nr * nx)) % 2, sv = t / (nr * nx * 2);
        const double *m = (k ? sG : sB) + qx * n;
        double acc = 0.0;
        for (int a = 0; a < n; ++a) acc += m[a] * sU[sv * Np + a + n * r];
        sT[t] = acc;
    }
    __syncthreads();
    double ke = 0.0, ens = 0.0, ike = 0.0;
    for (int col = threadIdx.x; col < nx * nx; col += blockDim.x) {
        const int qx = col % nx, qy = col / nx;
        // y pass for this (qx, qy): BB, BG (B in x, G in y), GB (G in x) per state and z node
        double BB[NV][N], BG[NV][N], GB[NV][N];
        constexpr int nz = (D == 3) ? n : 1;
        for (int sv = 0; sv < NV; ++sv)
            for (int c = 0; c < nz; ++c) {
                double bb = 0.0, bg = 0.0, gb = 0.0;
                for (int b = 0; b < n; ++b) {
                    const double tb = sT[((sv * 2 + 0) * nx + qx) * nr + b + n * c];
                    const double tg = sT[((sv * 2 + 1) * nx + qx) * nr + b + n * c];
                    bb += sB[qy * n + b] * tb;
                    bg += sG[qy * n + b] * tb;
                    gb += sB[qy * n + b] * tg;
                }
                BB[sv][c] = bb;
                BG[sv][c] = bg;
                GB[sv][c] = gb;
            }
        const int nqz = (D == 3) ? nx : 1;
        for (int qz = 0; qz < nqz; ++qz) {
            double val[NV], gr[NV][D];
            for (int sv = 0; sv < NV; ++sv) {
                if (D == 3) {
                    double v = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
                    for (int c = 0; c < n; ++c) {
                        const double bz = sB[qz * n + c], dz = sG[qz * n + c];
                        v += bz * BB[sv][c];
                        gx += bz * GB[sv][c];
                        gy += bz * BG[sv][c];
                        gz += dz * BB[sv][c];
                    }
                    val[sv] = v;
                    gr[sv][0] = gx;
                    gr[sv][1] = gy;
                    gr[sv][D - 1] = gz;
                } else {
                    val[sv] = BB[sv][0];
                    gr[sv][0] = GB[sv][0];
                    gr[sv][1] = BG[sv][0];
                }
            }
            const double rho = val[0];
            const double irho = 1.0 / rho;  // one division per point, reused
            double vel[D], vel2 = 0.0;
            for (int k = 0; k < D; ++k) {
                vel[k] = val[1 + k] * irho;
                vel2 = vel2 + vel[k] * vel[k];
            }
            double vg[D][D];
            for (int a2 = 0; a2 < D; ++a2)
                for (int b2 = 0; b2 < D; ++b2)
                    vg[a2][b2] = (gr[1 + a2][b2] * inv_h - vel[a2] * (gr[0][b2] * inv_h)) * irho;
            double om2;
            if (D == 2) {
                const double oz = vg[1][0] - vg[0][1];
                om2 = oz * oz;
            } else {
                const double ox = vg[2 % D][1 % D] - vg[1 % D][2 % D];
                const double oy = vg[0][2 % D] - vg[2 % D][0];
                const double oz = vg[1][0] - vg[0][1];
                om2 = ox * ox + oy * oy + oz * oz;
            }
            double w = sW[qx] * sW[qy];
            if (D == 3) w *= sW[qz];
            w *= detJ;
            ke += 0.5 * rho * vel2 * w;
            ens += 0.5 * rho * om2 * w;
            ike += 0.5 * vel2 * w;
        }
    }
    red[threadIdx.x] = ke;
    red[IQ_THREADS + threadIdx.x] = ens;
    red[2 * IQ_THREADS + threadIdx.x] = ike;
    __syncthreads();
    for (int sh = IQ_THREADS / 2; sh > 0; sh >>= 1) {
        if ((int)threadIdx.x < sh)
            for (int f = 0; f < 3; ++f) red[f * IQ_THREADS + threadIdx.x] += red[f * IQ_THREADS + threadIdx.x + sh];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int f = 0; f < 3; ++f) partial[(long)e * 3 + f] = red[f * IQ_THREADS];
}

// Integrated quantities on a curvilinear cell (periodic_turbulence.cpp:439-671 with the metric
// operators built at the QGauss points, :487-492 / operators.cpp:2520-2884): one block per cell,
// density, momentum, node coordinates and, in 3D, the conservative curl form's grid-node products
// P[r][j] = X_l dX_m/dxi_j ((l, m) = (z, y), (x, z), (y, x)) in LDS; each thread evaluates its
// points directly from the tensor-product Lagrange basis: values and reference gradients of the
// state, the Jacobian of the interpolated mapping (det J), and the cofactor (2D: the adjugate of
// that Jacobian; 3D: C[r][0] = d_zeta P[r][1] - d_eta P[r][2], C[r][1] = d_xi P[r][2] -
// d_zeta P[r][0], C[r][2] = d_eta P[r][0] - d_xi P[r][1] of the interpolant).  Physical gradient
// C grad_ref / det J, weights W_q det J; per-cell partial sums as integrated_kernel.
template <int D, int N>
__global__ void __launch_bounds__(IQ_THREADS)
integrated_curv_kernel(const double *__restrict__ u, const double *__restrict__ tab, int nx,
                       const double *__restrict__ xn, Ops1D ops, double *__restrict__ partial)
{
    constexpr int NS = D + 2, NV = D + 1, n = N;
    constexpr int Np = (D == 3) ? n * n * n : n * n;
    constexpr int NPF = (D == 3) ? 9 : 1;
    extern __shared__ double smem[];
    double *sU = smem;               // [NV][Np]
    double *sX = sU + NV * Np;       // [D][Np]
    double *sP = sX + D * Np;        // [9][Np] (3D)
    double *sB = sP + NPF * Np;      // [nx][n] values, then [nx][n] derivatives, then [nx] weights
    double *sG = sB + nx * n;
    double *sW = sG + nx * n;
    double *red = sW + nx;           // [3][IQ_THREADS]
    const int e = blockIdx.x;
    for (int t = threadIdx.x; t < NV * Np; t += blockDim.x) {
        const int sv = t / Np, i = t % Np;
        sU[t] = u[((long)e * NS + sv) * Np + i];
    }
    for (int t = threadIdx.x; t < D * Np; t += blockDim.x) {
        const int a = t / Np, i = t % Np;
        sX[t] = xn[((long)e * Np + i) * 3 + a];
    }
    for (int t = threadIdx.x; t < 2 * nx * n + nx; t += blockDim.x) sB[t] = tab[t];
    __syncthreads();
    if constexpr (D == 3) {
        constexpr int LM[3][2] = {{2, 1}, {0, 2}, {1, 0}};
        for (int i = threadIdx.x; i < Np; i += blockDim.x) {
            int c[3];
            node_coords<D, N>(i, c);
            for (int r = 0; r < 3; ++r)
                for (int j = 0; j < 3; ++j) {
                    double dx = 0.0;
                    for (int l = 0; l < n; ++l) dx += ops.D[c[j]][l] * sX[LM[r][1] * Np + node_with<D, N>(i, j, l)];
                    sP[(r * 3 + j) * Np + i] = sX[LM[r][0] * Np + i] * dx;
                }
        }
        __syncthreads();
    }
    const int nqv = (D == 3) ? nx * nx * nx : nx * nx;
    double ke = 0.0, ens = 0.0, ike = 0.0;
    for (int q = threadIdx.x; q < nqv; q += blockDim.x) {
        const int qx = q % nx, qy = (q / nx) % nx, qz = (D == 3) ? q / (nx * nx) : 0;
        double val[NV], gr[NV][D], jac[D][D], C[D][D];
        for (int v = 0; v < NV; ++v) {
            val[v] = 0.0;
            for (int k = 0; k < D; ++k) gr[v][k] = 0.0;
        }
        for (int a = 0; a < D; ++a)
            for (int k = 0; k < D; ++k) jac[a][k] = C[a][k] = 0.0;
        constexpr int nz = (D == 3) ? n : 1;
        for (int iz = 0; iz < nz; ++iz) {
            const double bz = (D == 3) ? sB[qz * n + iz] : 1.0, gz = (D == 3) ? sG[qz * n + iz] : 0.0;
            for (int iy = 0; iy < n; ++iy) {
                const double by = sB[qy * n + iy], gy = sG[qy * n + iy];
                const double byz = by * bz, gyz = gy * bz, bgz = by * gz;
                for (int ix = 0; ix < n; ++ix) {
                    const double bx = sB[qx * n + ix], gx = sG[qx * n + ix];
                    const int i = ix + n * (iy + n * iz);
                    const double wb = bx * byz;
                    double w[3] = {gx * byz, bx * gyz, bx * bgz};
                    for (int v = 0; v < NV; ++v) {
                        const double f = sU[v * Np + i];
                        val[v] += wb * f;
                        for (int k = 0; k < D; ++k) gr[v][k] += w[k] * f;
                    }
                    for (int a = 0; a < D; ++a) {
                        const double f = sX[a * Np + i];
                        for (int k = 0; k < D; ++k) jac[a][k] += w[k] * f;
                    }
                    if constexpr (D == 3) {
                        for (int r = 0; r < 3; ++r) {
                            const double p0 = sP[(r * 3 + 0) * Np + i], p1 = sP[(r * 3 + 1) * Np + i],
                                         p2 = sP[(r * 3 + 2) * Np + i];
                            C[r][0] += w[2] * p1 - w[1] * p2;
                            C[r][1] += w[0] * p2 - w[2] * p0;
                            C[r][2] += w[1] * p0 - w[0] * p1;
                        }
                    }
                }
            }
        }
        double det;
        if constexpr (D == 2) {
            C[0][0] = jac[1][1];
            C[1][0] = -jac[0][1];
            C[0][1] = -jac[1][0];
            C[1][1] = jac[0][0];
            det = jac[0][0] * jac[1][1] - jac[0][1] * jac[1][0];
        } else {
            det = jac[0][0] * (jac[1][1] * jac[2][2] - jac[1][2] * jac[2][1]) -
                  jac[0][1] * (jac[1][0] * jac[2][2] - jac[1][2] * jac[2][0]) +
                  jac[0][2] * (jac[1][0] * jac[2][1] - jac[1][1] * jac[2][0]);
        }
        double g[NV][D];
        for (int v = 0; v < NV; ++v)
            for (int k = 0; k < D; ++k) {
                double acc = 0.0;
                for (int j = 0; j < D; ++j) acc += C[k][j] * gr[v][j] / det;
                g[v][k] = acc;
            }
        const double rho = val[0];
        double vel[D], vel2 = 0.0;
        for (int k = 0; k < D; ++k) {
            vel[k] = val[1 + k] / rho;
            vel2 = vel2 + vel[k] * vel[k];
        }
        double vg[D][D];
        for (int a2 = 0; a2 < D; ++a2)
            for (int b2 = 0; b2 < D; ++b2) vg[a2][b2] = (g[1 + a2][b2] - vel[a2] * g[0][b2]) / rho;
        double om2;
        if constexpr (D == 2) {
            const double oz = vg[1][0] - vg[0][1];
            om2 = oz * oz;
        } else {
            const double ox = vg[2][1] - vg[1][2];
            const double oy = vg[0][2] - vg[2][0];
            const double oz = vg[1][0] - vg[0][1];
            om2 = ox * ox + oy * oy + oz * oz;
        }
        double wq = sW[qx] * sW[qy];
        if (D == 3) wq *= sW[qz];
        wq *= det;
        ke += 0.5 * rho * vel2 * wq;
        ens += 0.5 * rho * om2 * wq;
        ike += 0.5 * vel2 * wq;
    }
    red[threadIdx.x] = ke;
    red[IQ_THREADS + threadIdx.x] = ens;
    red[2 * IQ_THREADS + threadIdx.x] = ike;
    __syncthreads();
    for (int sh = IQ_THREADS / 2; sh > 0; sh >>= 1) {
        if ((int)threadIdx.x < sh)
            for (int f = 0; f < 3; ++f) red[f * IQ_THREADS + threadIdx.x] += red[f * IQ_THREADS + threadIdx.x + sh];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int f = 0; f < 3; ++f) partial[(long)e * 3 + f] = red[f * IQ_THREADS];
}

// Launches the per-cell integrated-quantity kernel (affine box or curvilinear cells) with the
// QGauss table `tab` (values, derivatives, weights) into partial[cell][3].
static int launch_iq_cells(dgx_ctx *c, const double *u, const double *tab, int nx, double *partial)
{
    const int n = c->n;
    if (c->curv) {
        const int npf = (c->dim == 3) ? 9 : 1;
        const size_t shmem = sizeof(double) * ((size_t)(c->dim + 1 + c->dim + npf) * c->Np + 2 * nx * n + nx +
                                               3 * IQ_THREADS);
        switch (c->dim * 16 + n) {
#define IQC_CASE(DD, NN)                                                                                          \
    case DD * 16 + NN:                                                                                             \
        hipLaunchKernelGGL((integrated_curv_kernel<DD, NN>), dim3(c->ncells), dim3(IQ_THREADS), shmem, c->stream, \
                           u, tab, nx, c->xn_dev, c->ops, partial);                                                \
        break;
            IQC_CASE(3, 2) IQC_CASE(3, 3) IQC_CASE(3, 4) IQC_CASE(3, 5) IQC_CASE(3, 6) IQC_CASE(3, 7)
            IQC_CASE(2, 2) IQC_CASE(2, 3) IQC_CASE(2, 4) IQC_CASE(2, 5) IQC_CASE(2, 6) IQC_CASE(2, 7)
#undef IQC_CASE
        default: c->err = "unsupported polynomial degree"; return DGX_ERR_UNSUPPORTED;
        }
        return check_launch(c, "integrated_curv_kernel");
    }
    const int nr = (c->dim == 3) ? n * n : n;
    const size_t shmem = sizeof(double) * ((size_t)(c->dim + 1) * c->Np + 2 * nx * n + nx +
                                           (size_t)(c->dim + 1) * 2 * nx * nr + 3 * IQ_THREADS);
    const double detJ = (c->dim == 3) ? c->h * c->h * c->h : c->h * c->h;
    switch (c->dim * 16 + n) {
#define IQ_CASE(DD, NN)                                                                                          \
    case DD * 16 + NN:                                                                                            \
        hipLaunchKernelGGL((integrated_kernel<DD, NN>), dim3(c->ncells), dim3(IQ_THREADS), shmem, c->stream, u,   \
                           tab, nx, c->ncells, 1.0 / c->h, detJ, partial);                                        \
        break;
        IQ_CASE(3, 2) IQ_CASE(3, 3) IQ_CASE(3, 4) IQ_CASE(3, 5) IQ_CASE(3, 6) IQ_CASE(3, 7)
        IQ_CASE(2, 2) IQ_CASE(2, 3) IQ_CASE(2, 4) IQ_CASE(2, 5) IQ_CASE(2, 6) IQ_CASE(2, 7)
#undef IQ_CASE
    default: c->err = "unsupported polynomial degree"; return DGX_ERR_UNSUPPORTED;
    }
    return check_launch(c, "integrated_kernel");
}

// maximum that keeps a NaN (fmax drops it): a non-physical state reaches the host
__device__ __forceinline__ double nan_max(double a, double b) { return (a != a || b != b) ? a + b : (a > b ? a : b); }

// Maximum local wave speed |u| + c over the QGauss(nx)^dim points of every cell
// (periodic_turbulence.cpp:371-422, euler.cpp:1033-1046): all states interpolated
// sum-factorised (x pass in LDS, y and z passes in registers); per-cell maxima.
template <int D, int N>
__global__ void __launch_bounds__(IQ_THREADS)
wave_speed_kernel(const double *__restrict__ u, const double *__restrict__ ops, int nx, double gam,
                  double *__restrict__ cellmax)
{
    constexpr int NS = D + 2, n = N;
    constexpr int Np = (D == 3) ? n * n * n : n * n;
    constexpr int nr = (D == 3) ? n * n : n;
    extern __shared__ double smem[];
    double *sU = smem;             // [NS][Np]
    double *sB = sU + NS * Np;     // [nx][n]
    double *sT = sB + nx * n;      // [NS][nx][nr]
    double *red = sT + NS * nx * nr;
    const int e = blockIdx.x;
    for (int t = threadIdx.x; t < NS * Np; t += blockDim.x) sU[t] = u[(long)e * NS * Np + t];
    for (int t = threadIdx.x; t < nx * n; t += blockDim.x) sB[t] = ops[t];
    __syncthreads();
    for (int t = threadIdx.x; t < NS * nx * nr; t += blockDim.x) {
        const int r = t % nr, qx = (t / nr) % nx, sv = t / (nr * nx);
        double acc = 0.0;
        for (int a = 0; a < n; ++a) acc += sB[qx * n + a] * sU[sv * Np + a + n * r];
        sT[t] = acc;
    }
    __syncthreads();
    double vmax = 0.0;
    for (int col = threadIdx.x; col < nx * nx; col += blockDim.x) {
        const int qx = col % nx, qy = col / nx;
        constexpr int nz = (D == 3) ? n : 1;
        double Y[NS][nz];
        for (int sv = 0; sv < NS; ++sv)
            for (int c = 0; c < nz; ++c) {
                double acc = 0.0;
                for (int b = 0; b < n; ++b) acc += sB[qy * n + b] * sT[(sv * nx + qx) * nr + b + n * c];
                Y[sv][c] = acc;
            }
        const int nqz = (D == 3) ? nx : 1;
        for (int qz = 0; qz < nqz; ++qz) {
            double us[NS];
            for (int sv = 0; sv < NS; ++sv) {
                if (D == 3) {
                    double acc = 0.0;
                    for (int c = 0; c < n; ++c) acc += sB[qz * n + c] * Y[sv][c];
                    us[sv] = acc;
                } else {
                    us[sv] = Y[sv][0];
                }
            }
            const double rho = us[0];
            double vel2 = 0.0;
            for (int k = 0; k < D; ++k) {
                const double v = us[1 + k] / rho;
                vel2 = vel2 + v * v;
            }
            const double p = (gam - 1.0) * (us[NS - 1] - 0.5 * rho * vel2);
            const double ev = sqrt(vel2) + sqrt(gam * p / rho);
            vmax = (ev > vmax || ev != ev) ? ev : vmax;  // a NaN (non-physical state) propagates
        }
    }
    red[threadIdx.x] = vmax;
    __syncthreads();
    for (int sh = IQ_THREADS / 2; sh > 0; sh >>= 1) {
        if ((int)threadIdx.x < sh) red[threadIdx.x] = nan_max(red[threadIdx.x], red[threadIdx.x + sh]);
        __syncthreads();
    }
    if (threadIdx.x == 0) cellmax[e] = red[0];
}

__global__ void __launch_bounds__(IQ_THREADS)
max_reduce_kernel(const double *__restrict__ v, int n, double *__restrict__ out)
{
    __shared__ double red[IQ_THREADS];
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += IQ_THREADS) m = nan_max(m, v[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int sh = IQ_THREADS / 2; sh > 0; sh >>= 1) {
        if ((int)threadIdx.x < sh) red[threadIdx.x] = nan_max(red[threadIdx.x], red[threadIdx.x + sh]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// Integrated numerical entropy on collocated GLL nodes (periodic_turbulence.cpp:765-868,
// euler.cpp:437-447): per cell sum_i W_i det J (-rho_i ln(p_i rho_i^-gamma)); the three
// partial slots of a cell hold (entropy, 0, 0) so the fixed-order reduction below applies.
__global__ void __launch_bounds__(IQ_THREADS)
entropy_kernel(const double *__restrict__ u, const double *__restrict__ dir, double gamma, int n, int dim,
               int ncells, double detJ, double gam, Ops1D ops, double *__restrict__ partial,
               const double *__restrict__ met)
{
    __shared__ double red[IQ_THREADS];
    const int e = blockIdx.x;
    const int Np = (dim == 3) ? n * n * n : n * n, ns = dim + 2;
    double acc = 0.0;
    for (int i = threadIdx.x; i < Np; i += blockDim.x) {
        double us[5];
        for (int s = 0; s < ns; ++s) {
            const long idx = ((long)e * ns + s) * Np + i;
            us[s] = dir ? u[idx] + gamma * dir[idx] : u[idx];  // u, or u + gamma dir (relaxation RK)
        }
        const double rho = us[0];
        double vel2 = 0.0;
        for (int k = 0; k < dim; ++k) {
            const double v = us[1 + k] / rho;
            vel2 = vel2 + v * v;
        }
        const double p = (gam - 1.0) * (us[ns - 1] - 0.5 * rho * vel2);
        const double ent = log(p * pow(rho, -gam));
        const int a = i % n, b = (i / n) % n, c = i / (n * n);
        double W = ops.w[a] * ops.w[b];
        if (dim == 3) W *= ops.w[c];
        acc += (-rho * ent) * W * node_jac(met, detJ, dim, Np, e, i);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int sh = IQ_THREADS / 2; sh > 0; sh >>= 1) {
        if ((int)threadIdx.x < sh) red[threadIdx.x] += red[threadIdx.x + sh];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[(long)e * 3] = red[0];
        partial[(long)e * 3 + 1] = 0.0;
        partial[(long)e * 3 + 2] = 0.0;
    }
}

// Fixed-order reduction of the per-cell partials (deterministic): thread t sums
// cells t, t+256, ... in order, then a tree over the threads.
__global__ void __launch_bounds__(IQ_THREADS)
integrated_sum_kernel(const double *__restrict__ partial, int ncells, double *__restrict__ out3)
{
    __shared__ double red[3][IQ_THREADS];
    double s[3] = {0.0, 0.0, 0.0};
    for (int e = threadIdx.x; e < ncells; e += IQ_THREADS)
        for (int f = 0; f < 3; ++f) s[f] += partial[(long)e * 3 + f];
    for (int f = 0; f < 3; ++f) red[f][threadIdx.x] = s[f];
    __syncthreads();
    for (int sh = IQ_THREADS / 2; sh > 0; sh >>= 1) {
        if ((int)threadIdx.x < sh)
            for (int f = 0; f < 3; ++f) red[f][threadIdx.x] += red[f][threadIdx.x + sh];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int f = 0; f < 3; ++f) out3[f] = red[f][0];
}

int launch_entropy(dgx_ctx *c, const double *u, double *out) { return launch_entropy_combo(c, u, nullptr, 0.0, out); }

// per-block partials of sum over nodes of W det J v(u) . k, v the entropy variables of the stage
// solution (compute_entropy_change_estimate, root_finding_rrk_ode_solver.cpp:357-390; collocated
// GLL: the projection of v is v, M = W det J)
template <int D>
__global__ void __launch_bounds__(VR_THREADS)
rrk_estimate_kernel(const double *__restrict__ u, const double *__restrict__ k, int n, int ncells, double detJ,
                    Ops1D ops, PhysConst pc, double *__restrict__ partial, const double *__restrict__ met)
{
    constexpr int NS = D + 2;
    __shared__ double red[VR_THREADS];
    const int Np = (D == 3) ? n * n * n : n * n;
    const long nodes = (long)ncells * Np;
    double acc = 0.0;
    for (long t = blockIdx.x * (long)VR_THREADS + threadIdx.x; t < nodes; t += (long)gridDim.x * VR_THREADS) {
        const long e = t / Np;
        const int i = (int)(t % Np);
        double us[NS], v[NS];
        for (int s = 0; s < NS; ++s) us[s] = u[(e * NS + s) * Np + i];
        entropy_variables<double, D>(us, pc, v);
        double W = ops.w[i % n] * ops.w[(i / n) % n];
        if (D == 3) W *= ops.w[i / (n * n)];
        const double jac = node_jac(met, detJ, D, Np, e, i);
        for (int s = 0; s < NS; ++s) acc += v[s] * (W * jac * k[(e * NS + s) * Np + i]);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int sh = VR_THREADS / 2; sh > 0; sh >>= 1) {
        if ((int)threadIdx.x < sh) red[threadIdx.x] += red[threadIdx.x + sh];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

int launch_rrk_estimate(dgx_ctx *c, const double *u, const double *k, double *partial_dev)
{
    const double detJ = (c->dim == 3) ? c->h * c->h * c->h : c->h * c->h;
    if (c->dim == 3)
        hipLaunchKernelGGL(rrk_estimate_kernel<3>, dim3(VR_BLOCKS), dim3(VR_THREADS), 0, c->stream, u, k, c->n,
                           c->ncells, detJ, c->ops, c->pc, partial_dev, c->curv ? c->met_dev : nullptr);
    else
        hipLaunchKernelGGL(rrk_estimate_kernel<2>, dim3(VR_BLOCKS), dim3(VR_THREADS), 0, c->stream, u, k, c->n,
                           c->ncells, detJ, c->ops, c->pc, partial_dev, c->curv ? c->met_dev : nullptr);
    return check_launch(c, "rrk_estimate_kernel");
}

// sum of the VR_BLOCKS partials at partial_dev, in block order, on the host
int fetch_partial_sum(dgx_ctx *c, const double *partial_dev, double *out)
{
    double h[VR_BLOCKS];
    if (hipMemcpyAsync(h, partial_dev, sizeof(double) * VR_BLOCKS, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        c->err = "reduction copy failed";
        return DGX_ERR_HIP;
    }
    double acc = 0.0;
    for (int b = 0; b < VR_BLOCKS; ++b) acc += h[b];
    *out = acc;
    return DGX_OK;
}

int vr_blocks() { return VR_BLOCKS; }

int launch_entropy_combo(dgx_ctx *c, const double *u, const double *dir, double gamma, double *out)
{
    if (!c->iq_partial && hipMalloc(&c->iq_partial, sizeof(double) * (3 * (size_t)c->ncells + 3)) != hipSuccess) {
        c->err = "hipMalloc failed";
        return DGX_ERR_HIP;
    }
    const double detJ = (c->dim == 3) ? c->h * c->h * c->h : c->h * c->h;
    hipLaunchKernelGGL(entropy_kernel, dim3(c->ncells), dim3(IQ_THREADS), 0, c->stream, u, dir, gamma, c->n, c->dim,
                       c->ncells, detJ, c->P.gamma_gas, c->ops, c->iq_partial, c->curv ? c->met_dev : nullptr);
    int rc = check_launch(c, "entropy_kernel");
    if (rc) return rc;
    double *sums = c->iq_partial + 3 * (size_t)c->ncells;
    hipLaunchKernelGGL(integrated_sum_kernel, dim3(1), dim3(IQ_THREADS), 0, c->stream, c->iq_partial, c->ncells, sums);
    rc = check_launch(c, "integrated_sum_kernel");
    if (rc) return rc;
    double s3[3];
    if (hipMemcpyAsync(s3, sums, sizeof(double) * 3, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        c->err = "numerical entropy copy failed";
        return DGX_ERR_HIP;
    }
    *out = s3[0];
    return DGX_OK;
}

// QGauss(p+1+overint) interpolation operators (values, derivatives, weights) and the per-cell
// partial buffer shared by the reduction kernels.
// QGauss(nx) table of the integrated-quantity kernels: values B[nx][n] and derivatives G[nx][n]
// of the GLL Lagrange basis at the Gauss points, then the Gauss weights w[nx] (2 nx n + nx
// doubles written to h).  nx <= 32.
static void iq_table(int n, int nx, double *h)
{
    double x[32], w[32], xs[MAXN1D], ws[MAXN1D];
    gauss_legendre(nx, x, w);
    gauss_lobatto(n, xs, ws);
    for (int q = 0; q < nx; ++q) {
        double val[MAXN1D], der[MAXN1D];
        lagrange(n, xs, x[q], val, der);
        for (int j = 0; j < n; ++j) {
            h[q * n + j] = val[j];
            h[nx * n + q * n + j] = der[j];
        }
    }
    for (int q = 0; q < nx; ++q) h[2 * nx * n + q] = w[q];
}

static int iq_operators(dgx_ctx *c, int overint)
{
    const int n = c->n, nx = c->p + 1 + overint;
    if (overint < 0 || nx > 32) {
        c->err = "over-integration too large";
        return DGX_ERR_INVALID;
    }
    if (c->iq_overint != overint) {
        std::vector<double> h((size_t)2 * nx * n + nx);
        iq_table(n, nx, h.data());
        (void)hipFree(c->iq_ops);
        c->iq_ops = nullptr;
        if (hipMalloc(&c->iq_ops, sizeof(double) * h.size()) != hipSuccess ||
            hipMemcpy(c->iq_ops, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
            c->err = "integrated-quantity operators upload failed";
            return DGX_ERR_HIP;
        }
        if (!c->iq_partial) {  // per-cell partials, then the 3 sums
            if (hipMalloc(&c->iq_partial, sizeof(double) * (3 * (size_t)c->ncells + 3)) != hipSuccess)
                return DGX_ERR_HIP;
        }
        c->iq_overint = overint;
    }
    return DGX_OK;
}

int launch_wave_speed(dgx_ctx *c, const double *u, int overint, double *out)
{
    const int n = c->n, nx = c->p + 1 + overint;
    int rc = iq_operators(c, overint);
    if (rc) return rc;
    const int nr = (c->dim == 3) ? n * n : n;
    const size_t shmem = sizeof(double) * ((size_t)c->ns * c->Np + nx * n + (size_t)c->ns * nx * nr + IQ_THREADS);
    switch (c->dim * 16 + n) {
#define WS_CASE(DD, NN)                                                                                         \
    case DD * 16 + NN:                                                                                           \
        hipLaunchKernelGGL((wave_speed_kernel<DD, NN>), dim3(c->ncells), dim3(IQ_THREADS), shmem, c->stream, u,  \
                           c->iq_ops, nx, c->P.gamma_gas, c->iq_partial);                                        \
        break;
        WS_CASE(3, 2) WS_CASE(3, 3) WS_CASE(3, 4) WS_CASE(3, 5) WS_CASE(3, 6) WS_CASE(3, 7)
        WS_CASE(2, 2) WS_CASE(2, 3) WS_CASE(2, 4) WS_CASE(2, 5) WS_CASE(2, 6) WS_CASE(2, 7)
#undef WS_CASE
    default: c->err = "unsupported polynomial degree"; return DGX_ERR_UNSUPPORTED;
    }
    rc = check_launch(c, "wave_speed_kernel");
    if (rc) return rc;
    double *m = c->iq_partial + 3 * (size_t)c->ncells;
    hipLaunchKernelGGL(max_reduce_kernel, dim3(1), dim3(IQ_THREADS), 0, c->stream, c->iq_partial, c->ncells, m);
    rc = check_launch(c, "max_reduce_kernel");
    if (rc) return rc;
    if (hipMemcpyAsync(out, m, sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        c->err = "wave speed copy failed";
        return DGX_ERR_HIP;
    }
    return DGX_OK;
}

int launch_integrated(dgx_ctx *c, const double *u, int overint, double *out4)
{
    const int nx = c->p + 1 + overint;
    int rc = iq_operators(c, overint);
    if (rc) return rc;
    rc = launch_iq_cells(c, u, c->iq_ops, nx, c->iq_partial);
    if (rc) return rc;
    double *sums = c->iq_partial + 3 * (size_t)c->ncells;
    hipLaunchKernelGGL(integrated_sum_kernel, dim3(1), dim3(IQ_THREADS), 0, c->stream, c->iq_partial, c->ncells, sums);
    rc = check_launch(c, "integrated_sum_kernel");
    if (rc) return rc;
    double s[3] = {0, 0, 0};
    if (hipMemcpyAsync(s, sums, sizeof(double) * 3, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        c->err = "integrated quantities copy failed";
        return DGX_ERR_HIP;
    }
    double vol = 1.0;
    for (int k = 0; k < c->dim; ++k) vol *= (c->P.domain_right - c->P.domain_left);
    out4[0] = s[0] / vol;
    out4[1] = s[1] / vol;
    out4[2] = s[2] / vol;
    out4[3] = 2.0 * out4[1] / c->P.reynolds_inf;
    return DGX_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free integrated quantities, for callers that cannot allocate device memory (a
// time loop captured in a HIP graph that records the kinetic-energy history on the device).
// The per-cell partial sums, the three sums and the QGauss operator table live in `scratch`, a
// file-scope device array: each GPU the library runs on holds its own copy until the process
// ends.  Its device address is looked up once per device (hipGetSymbolAddress, with that
// device current) and only that address is handed to kernels.  Every call on one GPU shares the
// array: calls that overlap on one GPU (different streams) race.  Nothing in it has to start at
// zero: every call writes the table, all the partials it reads and the sums before reading them.
// ---------------------------------------------------------------------------
constexpr int IQS_MAX_CELLS = 1 << 18;                           // 262,144 cells (64^3)
constexpr int IQS_MAX_NX = 24;                                   // QGauss points per direction
constexpr int IQS_OPS = 2 * IQS_MAX_NX * (MAXN1D - 1) + IQS_MAX_NX;  // B, G [nx][n], w[nx], n <= 7
constexpr size_t IQS_SUMS = 3 * (size_t)IQS_MAX_CELLS;           // offset of the three sums
constexpr size_t IQS_TABLE = IQS_SUMS + 8;                       // offset of the operator table
constexpr size_t IQS_DOUBLES = IQS_TABLE + IQS_OPS;              // 6.3 MB in all
static __device__ __align__(256) double scratch[IQS_DOUBLES];

struct IqTable {
    double v[IQS_OPS];  // passed by value as a kernel argument (2.9 KB)
};

__global__ void __launch_bounds__(256) iq_table_kernel(IqTable t, int count, double *__restrict__ dst)
{
    for (int i = threadIdx.x; i < count; i += blockDim.x) dst[i] = t.v[i];
}

// out4 as the host finishes dgx_integrated_quantities: sums / vol, dissipation 2 enstrophy / Re
__global__ void __launch_bounds__(64) iq_finalize_kernel(const double *__restrict__ sums, double vol,
                                                         double reynolds, double *__restrict__ out4)
{
    if (threadIdx.x == 0) {
        const double ens = sums[1] / vol;
        out4[0] = sums[0] / vol;
        out4[1] = ens;
        out4[2] = sums[2] / vol;
        out4[3] = 2.0 * ens / reynolds;
    }
}

static std::mutex g_scratch_mu;
static double *g_scratch_ptr[64];

// Device address of `scratch` on the current device (looked up once per device).
static int scratch_address(dgx_ctx *c, double **out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        c->err = "no current HIP device";
        return DGX_ERR_HIP;
    }
    if (dev != c->device) {
        c->err = "the context belongs to another device than the current one";
        return DGX_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            c->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return DGX_ERR_HIP;
        }
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return DGX_OK;
}

int launch_integrated_static(dgx_ctx *c, const double *u, int overint, double *out4_dev)
{
    const int n = c->n, nx = c->p + 1 + overint;
    if (overint < 0 || nx > IQS_MAX_NX) {
        c->err = "allocation-free integrated quantities: p + 1 + overint must be at most 24";
        return DGX_ERR_INVALID;
    }
    if (c->ncells > IQS_MAX_CELLS) {
        c->err = "allocation-free integrated quantities: at most 262144 cells";
        return DGX_ERR_INVALID;
    }
    double *sp = nullptr;
    int rc = scratch_address(c, &sp);
    if (rc) return rc;
    double *partial = sp, *sums = sp + IQS_SUMS, *table = sp + IQS_TABLE;
    IqTable t;
    const int count = 2 * nx * n + nx;
    iq_table(n, nx, t.v);
    hipLaunchKernelGGL(iq_table_kernel, dim3(1), dim3(256), 0, c->stream, t, count, table);
    rc = check_launch(c, "iq_table_kernel");
    if (rc) return rc;
    rc = launch_iq_cells(c, u, table, nx, partial);
    if (rc) return rc;
    hipLaunchKernelGGL(integrated_sum_kernel, dim3(1), dim3(IQ_THREADS), 0, c->stream, partial, c->ncells, sums);
    rc = check_launch(c, "integrated_sum_kernel");
    if (rc) return rc;
    double vol = 1.0;
    for (int k = 0; k < c->dim; ++k) vol *= (c->P.domain_right - c->P.domain_left);
    hipLaunchKernelGGL(iq_finalize_kernel, dim3(1), dim3(64), 0, c->stream, sums, vol, c->P.reynolds_inf, out4_dev);
    return check_launch(c, "iq_finalize_kernel");
}

} // namespace dgx
