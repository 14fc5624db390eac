// Phase-structured fused stage kernel.
//
// Per block: EPB whole elements.  Work items change shape between phases so
// that every lane has work:
//   A  node-parallel   load the stage state (FP64, coalesced), primitive
//                      cache and (conservative/viscous) volume fluxes into
//                      LDS (the split form's entropy projection u(v(u)) is
//                      the identity for collocated GLL nodes, not formed)
//   C  face-parallel   one item per (face, face node): neighbour trace from
//                      HBM, surface numerical flux (reference owner
//                      orientation: the lower cell index is "interior"),
//                      viscous numerical flux, result into LDS per direction
//   B/D per direction  pair-parallel two-point fluxes F#(u_a,u_b), a<b, each
//                      pair evaluated once and used by both nodes (the
//                      flux is symmetric); node-parallel skew-Hadamard sum
//   E  node-parallel   divergence, face lift 1/w_0, 1/h scaling and the RK
//                      register update in FP64
// T is the flux arithmetic type: double (FP64 path) or float (mixed path,
// FP32 fluxes and contractions, FP64 state and RK registers).
#pragma once
#include <type_traits>

#include "dg_kernels.hpp"

namespace dgx {

template <int D, int N> struct Geo {
    static constexpr int NP = (D == 3) ? N * N * N : N * N;
    static constexpr int NS = D + 2;
    static constexpr int NL = (D == 3) ? N * N : N;          // lines per direction = face nodes
    static constexpr int NPAIR = N * (N - 1) / 2;
    static constexpr int EPB = (NP >= 256) ? 1 : (256 / NP);
    static constexpr int BLOCK = ((EPB * NP + 63) / 64) * 64;
};

// node index of position `pos` along direction k on line `line`
template <int D, int N>
__device__ __forceinline__ int line_node(int k, int line, int pos)
{
    if (D == 3) {
        if (k == 0) return pos + N * line;
        if (k == 1) return (line % N) + N * pos + N * N * (line / N);
        return line + N * N * pos;
    } else {
        if (k == 0) return pos + N * line;
        return line + N * pos;
    }
}

template <int D, int N>
__device__ __forceinline__ int node_line(int k, int i)
{
    if (D == 3) {
        if (k == 0) return i / N;
        if (k == 1) return (i % N) + N * (i / (N * N));
        return i % (N * N);
    } else {
        if (k == 0) return i / N;
        return i % N;
    }
}

template <int N>
__device__ __forceinline__ void pair_nodes(int p, int &a, int &b)
{
    int q = p;
    a = 0;
    while (q >= N - 1 - a) {
        q -= N - 1 - a;
        ++a;
    }
    b = a + 1 + q;
}

template <int N>
__device__ __forceinline__ int pair_index(int a, int b)
{
    return a * (2 * N - a - 1) / 2 + (b - a - 1);
}

// CURV: curvilinear grid (Euler): the metric cofactor C[m][n] and det J of every node from a.met;
// two-point fluxes contracted with the pair's average cofactor column, conservative fluxes with the
// node's, face fluxes along the owner's physical normal C n_ref scaled by its length, du/dt = acc/J.
template <int D, int N, bool SPLIT, bool NSV, typename T, bool BND = false, bool CURV = false>
__global__ void __launch_bounds__((Geo<D, N>::BLOCK))
stage2_kernel(StageArgs a, Ops1D ops, PhysConst pc)
{
    using G = Geo<D, N>;
    constexpr int NP = G::NP, NS = G::NS, NL = G::NL, NPAIR = G::NPAIR, EPB = G::EPB;
    constexpr int NFLUX = (!SPLIT || NSV) ? D : 1;
    __shared__ T sU[EPB][NS][NP];                       // state for pairs and faces
    __shared__ T sP[EPB][SPLIT ? D + 1 : 1][NP];          // velocity, pressure (split form)
    __shared__ T sF[EPB][NFLUX][NS][NP];                // volume fluxes for the D-divergence
    __shared__ T sC[EPB][CURV ? D * D : 1][NP];         // CURV: metric cofactor, slot m * D + n
    constexpr int MS = D * D + 1;                       // metric slots per node in a.met (C, det J)
    // face terms per face direction (phase C) and, once they are summed into the node's
    // registers, the pair fluxes of one direction (phase B) in the same storage
    constexpr int PFN = SPLIT ? NL * NPAIR : 1;
    constexpr int FPW = (D * NS * NP > PFN * NS) ? D * NS * NP : PFN * NS;
    __shared__ T sFP[EPB * FPW];
    T (*const sFace)[D][NS][NP] = reinterpret_cast<T (*)[D][NS][NP]>(sFP);
    T (*const sPF)[PFN][NS] = reinterpret_cast<T (*)[PFN][NS]>(sFP);

    const int tid = threadIdx.x;
    const int blk0 = blockIdx.x * EPB;
    const T gamm1 = T(pc.gamm1);

    // ---------------- phase A: node-parallel load -----------------------
    const int le = tid / NP, i = tid % NP;
    const int e = blk0 + le;
    const bool active = (le < EPB) && (e < a.ncells);
    double ud[NS];
    if (active) {
        for (int s = 0; s < NS; ++s) ud[s] = a.u[((long)e * NS + s) * NP + i];
        T ut[NS];
        for (int s = 0; s < NS; ++s) ut[s] = T(ud[s]);
        for (int s = 0; s < NS; ++s) sU[le][s][i] = ut[s];
        T Cn[CURV ? D * D : 1];
        if constexpr (CURV) {
#pragma unroll
            for (int slot = 0; slot < D * D; ++slot) {
                Cn[slot] = T(a.met[((long)e * MS + slot) * NP + i]);
                sC[le][slot][i] = Cn[slot];
            }
        }
        const T pr = pressure<T, D>(ut, gamm1);
        flag_nonphysical<T>(a, ut[0], pr);
        if (SPLIT) {
            const T rho = ut[0];
#pragma unroll
            for (int k = 0; k < D; ++k) sP[le][k][i] = ut[1 + k] / rho;
            sP[le][D][i] = pr;
        }
        if (!SPLIT || NSV) {
            T g[NS * D];
            if (NSV)
                for (int s = 0; s < NS; ++s)
                    for (int d = 0; d < D; ++d) g[s * D + d] = aux_at<T>(a, d, ((long)e * NS + s) * NP + i);
            // CURV with the viscous terms: the physical viscous flux components, contracted below
            T fvm[D][NS];
            if constexpr (CURV && NSV) viscous_flux_all<T, D>(ut, g, pc, fvm);
#pragma unroll
            for (int k = 0; k < D; ++k) {
                T f[NS];
                for (int s = 0; s < NS; ++s) f[s] = T(0);
                if (!SPLIT) {
                    T uu[NS];
                    for (int s = 0; s < NS; ++s) uu[s] = T(ud[s]);
                    if constexpr (CURV) {
                        // contravariant flux sum_m C[m][k] F_m
                        T ck[D];
#pragma unroll
                        for (int m = 0; m < D; ++m) ck[m] = Cn[m * D + k];
                        convective_flux_vec<T, D>(uu, ck, gamm1, f);
                    } else {
                        convective_flux_dir<T, D>(uu, k, gamm1, f);
                    }
                }
                if (NSV) {
                    T fv[NS];
                    if constexpr (CURV) {
                        for (int s = 0; s < NS; ++s) {
                            T acc = T(0);
#pragma unroll
                            for (int m = 0; m < D; ++m) acc += fvm[m][s] * Cn[m * D + k];
                            fv[s] = acc;
                        }
                    } else {
                        viscous_flux_dir<T, D>(ut, g, k, pc, fv);
                    }
                    for (int s = 0; s < NS; ++s) f[s] += fv[s];
                }
                for (int s = 0; s < NS; ++s) sF[le][k][s][i] = f[s];
            }
        }
#pragma unroll
        for (int k = 0; k < D; ++k)
            for (int s = 0; s < NS; ++s) sFace[le][k][s][i] = T(0);
    }
    __syncthreads();

    // ---------------- phase C: face-parallel numerical fluxes -----------
    // Items ordered direction-major (direction, element, side, face node) so that
    // the items of a wave mostly share one direction, a template parameter of the
    // item code (no register array indexed per lane).
    constexpr int FITEMS = 2 * D * NL;
    for (int t = tid; t < EPB * FITEMS; t += G::BLOCK) {
        const int k = t / (EPB * 2 * NL), r0 = t % (EPB * 2 * NL);
        const int fle = r0 / (2 * NL), side = (r0 / NL) % 2, fn = r0 % NL;
        const int fe = blk0 + fle;
        if (fe >= a.ncells) continue;
        auto item = [&](auto kc) {
            constexpr int K = decltype(kc)::value;
            const int fi = line_node<D, N>(K, fn, side ? N - 1 : 0);
            const int nb = a.nbr[fe * 2 * D + 2 * K + side];
            const int fj = line_node<D, N>(K, fn, side ? 0 : N - 1);
            const T sgn = side ? T(1) : T(-1);
            T us[NS], un[NS], unt[NS];
            for (int s = 0; s < NS; ++s) us[s] = sU[fle][s][fi];
            bool own = fe < nb;
            if constexpr (CURV) {
                // the owner's outward physical normal C n_ref at its face node (metric_oper_int):
                // this element's, or the neighbour's read from a.met
                own = own || nb < 0;
                T nv[D], area2 = T(0);
#pragma unroll
                for (int m = 0; m < D; ++m) {
                    nv[m] = own ? sgn * sC[fle][m * D + K][fi]
                                : -sgn * T(a.met[((long)(nb < 0 ? 0 : nb) * MS + m * D + K) * NP + fj]);
                    area2 += nv[m] * nv[m];
                }
                const T area = dsqrt<T>(area2);
                T nh[D];
#pragma unroll
                for (int m = 0; m < D; ++m) nh[m] = nv[m] / area;
                if (nb >= 0) {
                    for (int s = 0; s < NS; ++s) un[s] = T(a.u[((long)nb * NS + s) * NP + fj]);
                } else {
                    boundary_state_n<T, D>(-nb, us, nh, pc, un);
                }
                T ua[NS], ub[NS];
                for (int s = 0; s < NS; ++s) {
                    ua[s] = own ? us[s] : un[s];
                    ub[s] = own ? un[s] : us[s];
                }
                const Prim<T, D> pa = to_prim<T, D>(ua, gamm1);
                const Prim<T, D> pb = to_prim<T, D>(ub, gamm1);
                T fs[NS];
                numerical_flux_n<T, D>(ua, pa, ub, pb, nh, pc, fs);
                const T fsgn = own ? T(1) : T(-1);
                T ft[NS];
                for (int s = 0; s < NS; ++s) ft[s] = -((area * fs[s]) * fsgn);
                // the strong form's interior contravariant flux: convective and viscous (conservative
                // form), viscous only (split form, whose convective part is in the two-point sums)
                if (!SPLIT || NSV)
                    for (int s = 0; s < NS; ++s) ft[s] += sgn * sF[fle][K][s][fi];
                if constexpr (NSV) {
                    // viscous numerical flux along this element's outward unit normal, scaled by the
                    // face area; walls as on the affine box; the penalty of the face from the cells'
                    // extents (evaluate_penalty_scaling, dg_base.cpp:397-413, 563, 615-617)
                    T ne[D];
#pragma unroll
                    for (int m = 0; m < D; ++m) ne[m] = own ? nh[m] : -nh[m];
                    T ui[NS], gi[NS * D], gn[NS * D], uv[NS];
                    bool adia = false;
                    for (int s = 0; s < NS; ++s) {
                        ui[s] = T(a.u[((long)fe * NS + s) * NP + fi]);
                        for (int d = 0; d < D; ++d) gi[s * D + d] = aux_at<T>(a, d, ((long)fe * NS + s) * NP + fi);
                    }
                    if (nb < 0) {
                        visc_wall_state<T, D>(-nb, ui, pc, uv);
                        adia = (-nb == 1001) && pc.thermal_bc == 0;
                        for (int s = 0; s < NS; ++s)
                            for (int d = 0; d < D; ++d) gn[s * D + d] = (adia && s == NS - 1) ? T(0) : gi[s * D + d];
                    } else {
                        for (int s = 0; s < NS; ++s) {
                            uv[s] = un[s];
                            for (int d = 0; d < D; ++d) gn[s * D + d] = aux_at<T>(a, d, ((long)nb * NS + s) * NP + fj);
                        }
                    }
                    // the viscous flux is linear in the gradient: central flux and SIPG penalty
                    // 1/2 [F(u_i, g_i - pen J) + F(u_v, g_v - pen J)] . n in two evaluations, J the
                    // jump (u_i - u_v) n
                    T pen = T(0);
                    if (pc.diss_flux == 0) {
                        const double pe = a.pen[(long)fe * D + K];
                        pen = T(nb < 0 ? pe : 0.5 * (pe + a.pen[(long)nb * D + K]));
                    }
                    for (int s = 0; s < NS; ++s)
                        for (int d = 0; d < D; ++d) {
                            const T jd = pen * ((ui[s] - uv[s]) * ne[d]);
                            gi[s * D + d] -= jd;
                            gn[s * D + d] -= jd;
                        }
                    T Fa[D][NS], Fb[D][NS], fvs[NS];
                    viscous_flux_all<T, D>(ui, gi, pc, Fa, adia);
                    viscous_flux_all<T, D>(uv, gn, pc, Fb, adia);
                    for (int s = 0; s < NS; ++s) {
                        T acc = T(0);
#pragma unroll
                        for (int m = 0; m < D; ++m) acc += (Fa[m][s] + Fb[m][s]) * ne[m];
                        fvs[s] = T(0.5) * acc;
                    }
                    for (int s = 0; s < NS; ++s) ft[s] -= area * fvs[s];
                }
                for (int s = 0; s < NS; ++s) sFace[fle][K][s][fi] = ft[s];
                return;
            }
            if constexpr (BND) {
                if (nb >= 0) {
                    for (int s = 0; s < NS; ++s) un[s] = T(a.u[((long)nb * NS + s) * NP + fj]);
                } else {
                    // boundary face of type -nb (Euler): the exterior state from the interior trace
                    boundary_state<T, D>(-nb, us, K, sgn, pc, un);
                    own = true;
                }
            } else {
                for (int s = 0; s < NS; ++s) un[s] = T(a.u[((long)nb * NS + s) * NP + fj]);
            }
            for (int s = 0; s < NS; ++s) unt[s] = un[s];
            // reference orientation: the lower cell index is the interior side (a boundary face's
            // interior is its cell); selected by value so the flux evaluation itself is not under a
            // divergent branch
            T ua[NS], ub[NS];
            for (int s = 0; s < NS; ++s) {
                ua[s] = own ? us[s] : unt[s];
                ub[s] = own ? unt[s] : us[s];
            }
            const T sg = own ? sgn : -sgn;
            const Prim<T, D> pa = to_prim<T, D>(ua, gamm1);
            const Prim<T, D> pb = to_prim<T, D>(ub, gamm1);
            T fs[NS];
            numerical_flux<T, D>(ua, pa, ub, pb, K, sg, pc, fs);
            const T fsgn = own ? T(1) : T(-1);
            T ft[NS];
            for (int s = 0; s < NS; ++s) ft[s] = -(fs[s] * fsgn);
            // the strong form's interior flux: convective and viscous (conservative form), viscous
            // only (split form), with its heat flux also on an adiabatic wall
            if (!SPLIT || NSV)
                for (int s = 0; s < NS; ++s) ft[s] += sgn * sF[fle][K][s][fi];
            if (NSV) {
                T ui[NS], gi[NS * D], gn[NS * D], uv[NS];
                bool wall = false, adia = false;
                if constexpr (BND) wall = nb < 0;
                for (int s = 0; s < NS; ++s) {
                    ui[s] = T(a.u[((long)fe * NS + s) * NP + fi]);
                    for (int d = 0; d < D; ++d) gi[s * D + d] = aux_at<T>(a, d, ((long)fe * NS + s) * NP + fi);
                }
                if (wall) {
                    // a wall (1001 no-slip, 1006 slip): the viscous wall state and the interior gradient,
                    // its energy row zeroed and every flux without heat flux on an adiabatic no-slip wall
                    // (boundary_face_values_viscous_flux, navier_stokes.cpp:1354-1430)
                    visc_wall_state<T, D>(-nb, ui, pc, uv);
                    adia = (-nb == 1001) && pc.thermal_bc == 0;
                    for (int s = 0; s < NS; ++s)
                        for (int d = 0; d < D; ++d) gn[s * D + d] = (adia && s == NS - 1) ? T(0) : gi[s * D + d];
                } else {
                    for (int s = 0; s < NS; ++s) {
                        uv[s] = un[s];
                        for (int d = 0; d < D; ++d) gn[s * D + d] = aux_at<T>(a, d, ((long)nb * NS + s) * NP + fj);
                    }
                }
                // the viscous flux is linear in the gradient: central flux and SIPG penalty
                // 1/2 [F(u_i, g_i - pen J) + F(u_v, g_v - pen J)] . n in two evaluations, the jump
                // J = (u_i - u_v) n nonzero in column K only
                const T pen = (pc.diss_flux == 0) ? T(a.penalty) : T(0);
                for (int s = 0; s < NS; ++s) {
                    const T jk = pen * ((ui[s] - uv[s]) * sgn);
                    gi[s * D + K] -= jk;
                    gn[s * D + K] -= jk;
                }
                T fa[NS], fb[NS];
                viscous_flux_dir<T, D>(ui, gi, K, pc, fa, adia);
                viscous_flux_dir<T, D>(uv, gn, K, pc, fb, adia);
                for (int s = 0; s < NS; ++s) ft[s] -= T(0.5) * (fa[s] * sgn + fb[s] * sgn);
            }
            for (int s = 0; s < NS; ++s) sFace[fle][K][s][fi] = ft[s];
        };
        if (k == 0) item(std::integral_constant<int, 0>{});
        else if (k == 1) item(std::integral_constant<int, 1>{});
        else if (D == 3) item(std::integral_constant<int, (D == 3 ? 2 : 1)>{});
    }

    // ---------------- surface terms into registers ----------------------
    int c[3] = {0, 0, 0};
    T acc[NS];
    for (int s = 0; s < NS; ++s) acc[s] = T(0);
    const T inv_wb = T(ops.inv_wb);
    __syncthreads();  // phase C done
    if (active) {
        node_coords<D, N>(i, c);
#pragma unroll
        for (int k = 0; k < D; ++k)
            for (int s = 0; s < NS; ++s) acc[s] += inv_wb * sFace[le][k][s][i];
    }

    // ---------------- phases B/D: volume ---------------------------------
    if (SPLIT) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
            __syncthreads();  // face terms / previous direction's sPF consumed
            constexpr int PITEMS = NL * NPAIR;
            for (int t = tid; t < EPB * PITEMS; t += G::BLOCK) {
                const int ple = t / PITEMS, r = t % PITEMS;
                if (blk0 + ple >= a.ncells) continue;
                const int line = r / NPAIR, p = r % NPAIR;
                int pa_, pb_;
                pair_nodes<N>(p, pa_, pb_);
                const int na = line_node<D, N>(k, line, pa_), nbn = line_node<D, N>(k, line, pb_);
                T u1[NS], u2[NS];
                Prim<T, D> q1, q2;
                for (int s = 0; s < NS; ++s) {
                    u1[s] = sU[ple][s][na];
                    u2[s] = sU[ple][s][nbn];
                }
                q1.rho = u1[0];
                q2.rho = u2[0];
                for (int d = 0; d < D; ++d) {
                    q1.vel[d] = sP[ple][d][na];
                    q2.vel[d] = sP[ple][d][nbn];
                }
                q1.p = sP[ple][D][na];
                q2.p = sP[ple][D][nbn];
                T F[NS];
                if constexpr (CURV) {
                    // the pair's average cofactor column (strong_dg.cpp:1348-1372)
                    T nv[D];
#pragma unroll
                    for (int m = 0; m < D; ++m) nv[m] = T(0.5) * (sC[ple][m * D + k][na] + sC[ple][m * D + k][nbn]);
                    two_point_flux_vec<T, D>(u1, q1, u2, q2, nv, pc, F);
                } else {
                    two_point_flux_dir<T, D>(u1, q1, u2, q2, k, pc, F);
                }
                for (int s = 0; s < NS; ++s) sPF[ple][r][s] = F[s];
            }
            __syncthreads();
            if (active) {
                const int line = node_line<D, N>(k, i);
                const int ck = c[k];
                for (int l = 0; l < N; ++l) {
                    if (l == ck) continue;
                    const int p = (ck < l) ? pair_index<N>(ck, l) : pair_index<N>(l, ck);
                    const T sk = T(ops.Sk[ck][l]);
                    for (int s = 0; s < NS; ++s) acc[s] -= sk * sPF[le][line * NPAIR + p][s];
                }
            }
        }
    }
    if (!active) return;
    if (!SPLIT || NSV) {
#pragma unroll
        for (int k = 0; k < D; ++k)
            for (int l = 0; l < N; ++l) {
                const int j = node_with<D, N>(i, k, l);
                const T dk = T(ops.D[c[k]][l]);
                for (int s = 0; s < NS; ++s) acc[s] -= dk * sF[le][k][s][j];
            }
    }

    // ---------------- phase E: FP64 epilogue -----------------------------
    if constexpr (CURV)
        stage_epilogue<D, N, T>(a, ops, e, i, c, ud, acc, false, nullptr, nullptr, a.met[((long)e * MS + D * D) * NP + i]);
    else
        stage_epilogue<D, N, T>(a, ops, e, i, c, ud, acc);
}

} // namespace dgx
