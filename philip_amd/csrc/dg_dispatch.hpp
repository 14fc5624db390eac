// Kernel dispatch templates over (dim, nodes per direction, scheme, precision): the stage
// kernels, the auxiliary (lifted-gradient) kernel and the batched line kernel.  Each N is
// instantiated in its own translation unit (dg_inst_n<N>.hip) so the kernels compile in
// parallel; dg_launch.hip calls the instantiations.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <type_traits>

#include "dg_host.hpp"
#include "dg_line.hpp"
#include "dg_stage2.hpp"

namespace dgx {

inline int check_launch(dgx_ctx *c, const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        c->err = std::string(what) + ": " + hipGetErrorString(e);
        return DGX_ERR_HIP;
    }
    return DGX_OK;
}

template <int D, int N, typename T>
void stage2_dn(dgx_ctx *c, const StageArgs &a, dim3 grid, dim3 block)
{
    const bool split = c->P.use_split_form != 0;
    const bool nsv = c->P.pde == DGX_NAVIER_STOKES;
    const bool bnd = c->has_boundary;
    using LG = LineGeoS<D, N, T>;
    const dim3 lgrid((c->ncells + LG::EPB - 1) / LG::EPB), lblock(LG::BLOCK);
    if (c->curv) {  // curvilinear grid: stage2_kernel with the metric terms, boundaries included
        if constexpr (D == 3) {
            if (nsv) {
                if (split)
                    hipLaunchKernelGGL((stage2_kernel<D, N, true, true, T, true, true>), grid, block, 0, c->stream, a,
                                       c->ops, c->pc);
                else
                    hipLaunchKernelGGL((stage2_kernel<D, N, false, true, T, true, true>), grid, block, 0, c->stream,
                                       a, c->ops, c->pc);
                return;
            }
        }
        if (split)
            hipLaunchKernelGGL((stage2_kernel<D, N, true, false, T, true, true>), grid, block, 0, c->stream, a, c->ops,
                               c->pc);
        else
            hipLaunchKernelGGL((stage2_kernel<D, N, false, false, T, true, true>), grid, block, 0, c->stream, a,
                               c->ops, c->pc);
        return;
    }
    if constexpr (D == 3) {
        if (nsv && bnd) {
            // Navier-Stokes with walls: stage2_kernel at both precisions
            if (split)
                hipLaunchKernelGGL((stage2_kernel<D, N, true, true, T, true>), grid, block, 0, c->stream, a, c->ops,
                                   c->pc);
            else
                hipLaunchKernelGGL((stage2_kernel<D, N, false, true, T, true>), grid, block, 0, c->stream, a,
                                   c->ops, c->pc);
        } else if (split && nsv) {
            // mixed: line kernel with the viscous terms; FP64: stage2_kernel (the FP64 line
            // variant needs ~85 KB of LDS per block at p=5 and ran slower at one block per CU)
            if constexpr (std::is_same<T, float>::value) {
                using LV = LineGeoK<D, N, T, true>;
                const dim3 vgrid((c->ncells + LV::EPB - 1) / LV::EPB), vblock(LV::BLOCK);
                if constexpr (N == 6)
                    hipLaunchKernelGGL((line_kernel_ns_w4<D, N>), vgrid, vblock, 0, c->stream, a, c->ops, c->pc);
                else
                    hipLaunchKernelGGL((line_kernel_ns<D, N>), vgrid, vblock, 0, c->stream, a, c->ops, c->pc);
            } else {
                hipLaunchKernelGGL((stage2_kernel<D, N, true, true, T>), grid, block, 0, c->stream, a, c->ops,
                                   c->pc);
            }
        }
        else if (!split && nsv)
            hipLaunchKernelGGL((stage2_kernel<D, N, false, true, T>), grid, block, 0, c->stream, a, c->ops, c->pc);
        else if (split) {
            if constexpr (line_waves<D, N, T>() > 0) {
                if (bnd)
                    hipLaunchKernelGGL((line_kernel_wb<D, N, T, line_waves<D, N, T>(), true>), lgrid, lblock, 0,
                                       c->stream, a, c->ops, c->pc);
                else
                    hipLaunchKernelGGL((line_kernel_wb<D, N, T, line_waves<D, N, T>()>), lgrid, lblock, 0,
                                       c->stream, a, c->ops, c->pc);
            } else {
                if (bnd)
                    hipLaunchKernelGGL((line_kernel<D, N, T, true>), lgrid, lblock, 0, c->stream, a, c->ops, c->pc);
                else
                    hipLaunchKernelGGL((line_kernel<D, N, T>), lgrid, lblock, 0, c->stream, a, c->ops, c->pc);
            }
        } else if (bnd) {
            hipLaunchKernelGGL((stage2_kernel<D, N, false, false, T, true>), grid, block, 0, c->stream, a, c->ops,
                               c->pc);
        } else {
            hipLaunchKernelGGL((stage2_kernel<D, N, false, false, T>), grid, block, 0, c->stream, a, c->ops, c->pc);
        }
    } else {
        if (split) {
            if (bnd)
                hipLaunchKernelGGL((line_kernel<D, N, T, true>), lgrid, lblock, 0, c->stream, a, c->ops, c->pc);
            else
                hipLaunchKernelGGL((line_kernel<D, N, T>), lgrid, lblock, 0, c->stream, a, c->ops, c->pc);
        } else if (bnd) {
            hipLaunchKernelGGL((stage2_kernel<D, N, false, false, T, true>), grid, block, 0, c->stream, a, c->ops,
                               c->pc);
        } else {
            hipLaunchKernelGGL((stage2_kernel<D, N, false, false, T>), grid, block, 0, c->stream, a, c->ops, c->pc);
        }
    }
}

template <int D, int N>
void stage_dn(dgx_ctx *c, const StageArgs &a, dim3 grid, dim3 block)
{
    // FP64: fluxes in double; mixed: fluxes and contractions in float, state and RK registers
    // in double.  The reference's entropy projection u(v(u)) is the identity for collocated GLL
    // nodes and is formed by neither.
    if (c->P.precision == DGX_MIXED) stage2_dn<D, N, float>(c, a, grid, block);
    else stage2_dn<D, N, double>(c, a, grid, block);
}

template <int D, int N>
int launch_stage_dn(dgx_ctx *c, const StageArgs &a)
{
    using S = Geo<D, N>;
    const int blocks = (c->ncells + S::EPB - 1) / S::EPB;
    if (c->timing.enabled && hipEventRecord(c->timing.a, c->stream) != hipSuccess) {
        c->err = "hipEventRecord failed";
        return DGX_ERR_HIP;
    }
    stage_dn<D, N>(c, a, dim3(blocks), dim3(S::BLOCK));
    int rc = check_launch(c, "stage_kernel");
    if (rc) return rc;
    if (c->timing.enabled) {
        float ms = 0.f;
        if (hipEventRecord(c->timing.b, c->stream) != hipSuccess || hipEventSynchronize(c->timing.b) != hipSuccess ||
            hipEventElapsedTime(&ms, c->timing.a, c->timing.b) != hipSuccess) {
            c->err = "stage kernel timing failed";
            return DGX_ERR_HIP;
        }
        c->timing.sum_ms += ms;
        c->timing.n += 1;
    }
    return DGX_OK;
}

template <int D, int N>
int launch_aux_dn(dgx_ctx *c, const double *u)
{
    using S = Shape<D, N>;
    const int blocks = (c->ncells + S::EPB - 1) / S::EPB;
    if (c->P.precision == DGX_MIXED)
        hipLaunchKernelGGL((aux_kernel<D, N, float>), dim3(blocks), dim3(S::BLOCK), 0, c->stream, u, c->auxf_dev[0],
                           c->auxf_dev[1], c->auxf_dev[2], c->nbr_dev, c->ncells, 1.0 / c->h, c->ops, c->pc,
                           c->curv ? c->met_dev : nullptr);
    else
        hipLaunchKernelGGL((aux_kernel<D, N, double>), dim3(blocks), dim3(S::BLOCK), 0, c->stream, u, c->aux_dev[0],
                           c->aux_dev[1], c->aux_dev[2], c->nbr_dev, c->ncells, 1.0 / c->h, c->ops, c->pc,
                           c->curv ? c->met_dev : nullptr);
    return check_launch(c, "aux_kernel");
}

template <int D, int N, typename T, int ES, int W, bool XJ>
int line_batch_es(const BatchBlock *blocks, int nblocks, const StageArgs *args, const PhysConst *pcs,
                  const Ops1D &ops, hipStream_t st, int *epb, int *threads)
{
    using LG = LineGeo<D, N, T, ES, false, XJ>;
    if constexpr (LG::SBUF * sizeof(T) > 160 * 1024 - 4096 || LG::BLOCK > 1024) {
        return DGX_ERR_UNSUPPORTED;  // would not fit the 160 KB of LDS or the 1024 threads of a block
    } else {
        if (epb) *epb = LG::EPB;
        if (threads) *threads = LG::BLOCK;
        if (!blocks) return DGX_OK;  // query only
        if constexpr (W > 0)
            hipLaunchKernelGGL((line_kernel_batch_wb<D, N, T, ES, W, XJ>), dim3(nblocks), dim3(LG::BLOCK), 0, st,
                               blocks, args, pcs, ops);
        else
            hipLaunchKernelGGL((line_kernel_batch<D, N, T, ES, XJ>), dim3(nblocks), dim3(LG::BLOCK), 0, st, blocks,
                               args, pcs, ops);
        return hipGetLastError() == hipSuccess ? DGX_OK : DGX_ERR_HIP;
    }
}

template <int D, int N, typename T>
int line_batch_t(int tile, const BatchBlock *blocks, int nblocks, const StageArgs *args, const PhysConst *pcs,
                 const Ops1D &ops, hipStream_t st, int *epb, int *threads)
{
    constexpr bool F = std::is_same<T, float>::value;
#define LBT_CASE(V)                                                                                              \
    case V:                                                                                                      \
        return line_batch_es<D, N, T, batch_tile(D, F, V).es, batch_tile(D, F, V).waves, batch_tile(D, F, V).xj>( \
            blocks, nblocks, args, pcs, ops, st, epb, threads);
    if constexpr (D == 2) {
        switch (tile) {
            LBT_CASE(0) LBT_CASE(1) LBT_CASE(2) LBT_CASE(3) LBT_CASE(4) LBT_CASE(5)
            LBT_CASE(6) LBT_CASE(7) LBT_CASE(8) LBT_CASE(9) LBT_CASE(10) LBT_CASE(11)
        default: return DGX_ERR_UNSUPPORTED;
        }
    } else {
        switch (tile) {
            LBT_CASE(0) LBT_CASE(1)
        default: return DGX_ERR_UNSUPPORTED;
        }
    }
#undef LBT_CASE
}

template <int D, int N>
int line_batch_dn(int precision, int tile, const BatchBlock *blocks, int nblocks, const StageArgs *args,
                  const PhysConst *pcs, const Ops1D &ops, hipStream_t st, int *epb, int *threads)
{
    if (tile < 0 || tile >= batch_tile_count(D)) return DGX_ERR_UNSUPPORTED;
    if (precision == DGX_MIXED)
        return line_batch_t<D, N, float>(tile, blocks, nblocks, args, pcs, ops, st, epb, threads);
    return line_batch_t<D, N, double>(tile, blocks, nblocks, args, pcs, ops, st, epb, threads);
}

// explicit instantiation (in dg_inst_n<NN>.hip) and its extern declaration (in dg_launch.hip)
#define DGX_KERNELS_N(PREFIX, NN)                                                                          \
    PREFIX template int launch_stage_dn<2, NN>(dgx_ctx *, const StageArgs &);                              \
    PREFIX template int launch_stage_dn<3, NN>(dgx_ctx *, const StageArgs &);                              \
    PREFIX template int launch_aux_dn<3, NN>(dgx_ctx *, const double *);                                   \
    PREFIX template int line_batch_dn<2, NN>(int, int, const BatchBlock *, int, const StageArgs *,          \
                                             const PhysConst *, const Ops1D &, hipStream_t, int *, int *);   \
    PREFIX template int line_batch_dn<3, NN>(int, int, const BatchBlock *, int, const StageArgs *,          \
                                             const PhysConst *, const Ops1D &, hipStream_t, int *, int *);

} // namespace dgx
