// Template dispatch of the DG kernels over (dim, nodes per direction, scheme),
// launch bookkeeping (HIP-event timing of the stage kernels), and the
// integrated-quantity reduction (kinetic energy, enstrophy).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "dg_host.hpp"
#include "dg_line.hpp"
#include "dg_stage2.hpp"

namespace dgx {

static int check_launch(dgx_ctx *c, const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        c->err = std::string(what) + ": " + hipGetErrorString(e);
        return DGX_ERR_HIP;
    }
    return DGX_OK;
}

template <int D, int N, typename T, bool RT>
static void stage2_dn(dgx_ctx *c, const StageArgs &a, dim3 grid, dim3 block)
{
    const bool split = c->P.use_split_form != 0;
    const bool nsv = c->P.pde == DGX_NAVIER_STOKES;
    using LG = LineGeo<D, N, T>;
    const dim3 lgrid((c->ncells + LG::EPB - 1) / LG::EPB), lblock(LG::BLOCK);
    if constexpr (D == 3) {
        if (split && nsv)
            hipLaunchKernelGGL((stage2_kernel<D, N, true, true, RT, T>), grid, block, 0, c->stream, a, c->ops, c->pc);
        else if (!split && nsv)
            hipLaunchKernelGGL((stage2_kernel<D, N, false, true, false, T>), grid, block, 0, c->stream, a, c->ops, c->pc);
        else if (split)
            hipLaunchKernelGGL((line_kernel<D, N, RT, T>), lgrid, lblock, 0, c->stream, a, c->ops, c->pc);
        else
            hipLaunchKernelGGL((stage2_kernel<D, N, false, false, false, T>), grid, block, 0, c->stream, a, c->ops, c->pc);
    } else {
        if (split)
            hipLaunchKernelGGL((line_kernel<D, N, RT, T>), lgrid, lblock, 0, c->stream, a, c->ops, c->pc);
        else
            hipLaunchKernelGGL((stage2_kernel<D, N, false, false, false, T>), grid, block, 0, c->stream, a, c->ops, c->pc);
    }
}

template <int D, int N>
static void stage_dn(dgx_ctx *c, const StageArgs &a, dim3 grid, dim3 block)
{
    // FP64: fluxes in double with the reference's entropy-projection round trip;
    // mixed: fluxes and contractions in float, state and RK registers in double
    if (c->P.precision == DGX_MIXED) stage2_dn<D, N, float, false>(c, a, grid, block);
    else stage2_dn<D, N, double, true>(c, a, grid, block);
}

template <int D, int N>
static int launch_stage_dn(dgx_ctx *c, const StageArgs &a)
{
    using S = Geo<D, N>;
    const int blocks = (c->ncells + S::EPB - 1) / S::EPB;
    if (c->timing.enabled) hipEventRecord(c->timing.a, c->stream);
    stage_dn<D, N>(c, a, dim3(blocks), dim3(S::BLOCK));
    int rc = check_launch(c, "stage_kernel");
    if (rc) return rc;
    if (c->timing.enabled) {
        hipEventRecord(c->timing.b, c->stream);
        hipEventSynchronize(c->timing.b);
        float ms = 0.f;
        hipEventElapsedTime(&ms, c->timing.a, c->timing.b);
        c->timing.sum_ms += ms;
        c->timing.n += 1;
    }
    return DGX_OK;
}

template <int D, int N>
static int launch_aux_dn(dgx_ctx *c, const double *u)
{
    using S = Shape<D, N>;
    const int blocks = (c->ncells + S::EPB - 1) / S::EPB;
    hipLaunchKernelGGL((aux_kernel<D, N>), dim3(blocks), dim3(S::BLOCK), 0, c->stream, u, c->aux_dev[0],
                       c->aux_dev[1], c->aux_dev[2], c->nbr_dev, c->ncells, 1.0 / c->h, c->ops);
    return check_launch(c, "aux_kernel");
}

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
template <int D, int N, typename T, bool RT, int ES>
static int line_batch_es(const BatchBlock *blocks, int nblocks, const StageArgs *args, const PhysConst *pcs,
                         const Ops1D &ops, hipStream_t st, int *epb)
{
    using LG = LineGeo<D, N, T, ES>;
    if constexpr (LG::SBUF * sizeof(T) > 65536) {
        return DGX_ERR_UNSUPPORTED;  // variant would not fit the per-block LDS budget
    } else {
        if (epb) *epb = LG::EPB;
        if (!blocks) return DGX_OK;  // query only
        hipLaunchKernelGGL((line_kernel_batch<D, N, RT, T, ES>), dim3(nblocks), dim3(LG::BLOCK), 0, st, blocks, args,
                           pcs, ops);
        return hipGetLastError() == hipSuccess ? DGX_OK : DGX_ERR_HIP;
    }
}

template <int D, int N>
static int line_batch_dn(int precision, int es, const BatchBlock *blocks, int nblocks, const StageArgs *args,
                         const PhysConst *pcs, const Ops1D &ops, hipStream_t st, int *epb)
{
    if (precision == DGX_MIXED) {
        if (es == 2) return line_batch_es<D, N, float, false, 2>(blocks, nblocks, args, pcs, ops, st, epb);
        return line_batch_es<D, N, float, false, 1>(blocks, nblocks, args, pcs, ops, st, epb);
    }
    if (es == 2) return line_batch_es<D, N, double, true, 2>(blocks, nblocks, args, pcs, ops, st, epb);
    return line_batch_es<D, N, double, true, 1>(blocks, nblocks, args, pcs, ops, st, epb);
}

int launch_line_batch(int dim, int n, int precision, int es, const BatchBlock *blocks, int nblocks,
                      const StageArgs *args, const PhysConst *pcs, const Ops1D &ops, hipStream_t st, int *epb)
{
#define LB_CASE(NN)                                                                                 \
    case NN:                                                                                        \
        return (dim == 3) ? line_batch_dn<3, NN>(precision, es, blocks, nblocks, args, pcs, ops, st, epb) \
                          : line_batch_dn<2, NN>(precision, es, blocks, nblocks, args, pcs, ops, st, epb);
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

// Inverse of the collocated GLL mass matrix (diagonal): out_q = in_q / (W_q det J),
// W_q the tensor-product GLL weight of the node (dg_base.cpp:4089-4260).
__global__ void __launch_bounds__(256)
inverse_mass_kernel(const double *__restrict__ in, double *__restrict__ out, long ndofs, int n, int dim,
                    double detJ, Ops1D ops)
{
    const int Np = (dim == 3) ? n * n * n : n * n;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < ndofs; q += (long)gridDim.x * blockDim.x) {
        const int node = (int)(q % Np);
        const int a = node % n, b = (node / n) % n, c = node / (n * n);
        double W = ops.w[a] * ops.w[b];
        if (dim == 3) W *= ops.w[c];
        out[q] = in[q] / (W * detJ);
    }
}

int launch_inverse_mass(dgx_ctx *c, const double *in, double *out)
{
    const double detJ = (c->dim == 3) ? c->h * c->h * c->h : c->h * c->h;
    const long blocks = std::min<long>((c->ndofs + 255) / 256, 4096);
    hipLaunchKernelGGL(inverse_mass_kernel, dim3((unsigned)blocks), dim3(256), 0, c->stream, in, out, (long)c->ndofs,
                       c->n, c->dim, detJ, c->ops);
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

template <int D>
__global__ void __launch_bounds__(IQ_THREADS)
integrated_kernel(const double *__restrict__ u, const double *__restrict__ ops, int n, int nx, int ncells,
                  double inv_h, double detJ, double *__restrict__ partial)
{
    constexpr int NS = D + 2;
    extern __shared__ double smem[];
    const int Np = (D == 3) ? n * n * n : n * n;
    double *sU = smem;                       // [NS][Np]
    double *sB = sU + NS * Np;               // [nx][n]
    double *sG = sB + nx * n;                // [nx][n]
    double *sW = sG + nx * n;                // [nx]
    double *red = sW + nx;                   // [3][IQ_THREADS]
    const int e = blockIdx.x;
    for (int t = threadIdx.x; t < NS * Np; t += blockDim.x) {
        const int s = t / Np, i = t % Np;
        sU[t] = u[((long)e * NS + s) * Np + i];
    }
    for (int t = threadIdx.x; t < 2 * nx * n + nx; t += blockDim.x) sB[t] = ops[t];
    __syncthreads();
    const int nq = (D == 3) ? nx * nx * nx : nx * nx;
    double ke = 0.0, ens = 0.0, ike = 0.0;
    for (int q = threadIdx.x; q < nq; q += blockDim.x) {
        const int qx = q % nx, qy = (q / nx) % nx, qz = (D == 3) ? q / (nx * nx) : 0;
        double val[NS], gr[NS][D];
        for (int s = 0; s < NS; ++s) {
            val[s] = 0.0;
            for (int k = 0; k < D; ++k) gr[s][k] = 0.0;
        }
        for (int i = 0; i < Np; ++i) {
            const int a = i % n, b = (i / n) % n, cz = (D == 3) ? i / (n * n) : 0;
            const double bx = sB[qx * n + a], by = sB[qy * n + b], bz = (D == 3) ? sB[qz * n + cz] : 1.0;
            const double gx = sG[qx * n + a], gy = sG[qy * n + b], gz = (D == 3) ? sG[qz * n + cz] : 0.0;
            const double phi = bx * by * bz;
            const double dphi[3] = {gx * by * bz, bx * gy * bz, bx * by * gz};
            for (int s = 0; s < NS; ++s) {
                const double us = sU[s * Np + i];
                val[s] += phi * us;
                for (int k = 0; k < D; ++k) gr[s][k] += dphi[k] * us;
            }
        }
        const double rho = val[0];
        double vel[D], vel2 = 0.0;
        for (int k = 0; k < D; ++k) {
            vel[k] = val[1 + k] / rho;
            vel2 = vel2 + vel[k] * vel[k];
        }
        double vg[D][D];
        for (int a2 = 0; a2 < D; ++a2)
            for (int b2 = 0; b2 < D; ++b2)
                vg[a2][b2] = (gr[1 + a2][b2] * inv_h - vel[a2] * (gr[0][b2] * inv_h)) / rho;
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
    red[threadIdx.x] = ke;
    red[IQ_THREADS + threadIdx.x] = ens;
    red[2 * IQ_THREADS + threadIdx.x] = ike;
    __syncthreads();
    for (int s = IQ_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int f = 0; f < 3; ++f) red[f * IQ_THREADS + threadIdx.x] += red[f * IQ_THREADS + threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int f = 0; f < 3; ++f) partial[(long)e * 3 + f] = red[f * IQ_THREADS];
}

int launch_integrated(dgx_ctx *c, const double *u, int overint, double *out4)
{
    const int n = c->n, nx = c->p + 1 + overint;
    if (nx > 32) {
        c->err = "over-integration too large";
        return DGX_ERR_INVALID;
    }
    if (c->iq_overint != overint) {
        std::vector<double> x(nx), w(nx), xs(n), ws(n), h((size_t)2 * nx * n + nx);
        gauss_legendre(nx, x.data(), w.data());
        gauss_lobatto(n, xs.data(), ws.data());
        for (int q = 0; q < nx; ++q) {
            double val[MAXN1D], der[MAXN1D];
            lagrange(n, xs.data(), x[q], val, der);
            for (int j = 0; j < n; ++j) {
                h[q * n + j] = val[j];
                h[nx * n + q * n + j] = der[j];
            }
        }
        for (int q = 0; q < nx; ++q) h[2 * nx * n + q] = w[q];
        hipFree(c->iq_ops);
        if (hipMalloc(&c->iq_ops, sizeof(double) * h.size()) != hipSuccess) return DGX_ERR_HIP;
        hipMemcpy(c->iq_ops, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice);
        if (!c->iq_partial) {
            if (hipMalloc(&c->iq_partial, sizeof(double) * 3 * c->ncells) != hipSuccess) return DGX_ERR_HIP;
        }
        c->iq_overint = overint;
    }
    const size_t shmem = sizeof(double) * ((size_t)c->ns * c->Np + 2 * nx * n + nx + 3 * IQ_THREADS);
    const double detJ = (c->dim == 3) ? c->h * c->h * c->h : c->h * c->h;
    if (c->dim == 3)
        hipLaunchKernelGGL(integrated_kernel<3>, dim3(c->ncells), dim3(IQ_THREADS), shmem, c->stream, u, c->iq_ops, n, nx,
                           c->ncells, 1.0 / c->h, detJ, c->iq_partial);
    else
        hipLaunchKernelGGL(integrated_kernel<2>, dim3(c->ncells), dim3(IQ_THREADS), shmem, c->stream, u, c->iq_ops, n, nx,
                           c->ncells, 1.0 / c->h, detJ, c->iq_partial);
    int rc = check_launch(c, "integrated_kernel");
    if (rc) return rc;
    std::vector<double> part((size_t)3 * c->ncells);
    if (hipMemcpyAsync(part.data(), c->iq_partial, sizeof(double) * part.size(), hipMemcpyDeviceToHost, c->stream) !=
            hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        c->err = "integrated quantities copy failed";
        return DGX_ERR_HIP;
    }
    double s[3] = {0, 0, 0};
    for (int e = 0; e < c->ncells; ++e)
        for (int f = 0; f < 3; ++f) s[f] += part[(size_t)e * 3 + f];
    double vol = 1.0;
    for (int k = 0; k < c->dim; ++k) vol *= (c->P.domain_right - c->P.domain_left);
    out4[0] = s[0] / vol;
    out4[1] = s[1] / vol;
    out4[2] = s[2] / vol;
    out4[3] = 2.0 * out4[1] / c->P.reynolds_inf;
    return DGX_OK;
}

} // namespace dgx
