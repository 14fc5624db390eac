// Fused strong-form DG right-hand-side kernels for collocated Gauss-Lobatto
// nodes on periodic Cartesian hexahedra/quadrilaterals, with the explicit RK
// stage update folded into the epilogue.
//
// One thread owns one solution node; a block holds EPB whole elements so the
// element's state is read from HBM once, staged in LDS for the line
// contractions / two-point pair loops, and the face coupling reads only the
// neighbour's face layer.  Reference path restated:
//   volume   src/dg/strong_dg.cpp:1017-1579   (conservative D-divergence or
//            Hadamard skew two-point form, Chan 2019 eq. 15)
//   faces    src/dg/strong_dg.cpp:2368-3576   (numerical flux lifted with
//            W_f/W = 1/w_0 at GLL boundary nodes; the surface-volume hybrid
//            terms cancel exactly for collocated GLL)
//   aux eq.  src/dg/strong_dg.cpp:498-1008
//   M^-1     src/dg/dg_base.cpp:4089-4260 (diagonal for collocated GLL)
//   RK       src/ode_solver/runge_kutta_ode_solver.cpp:15-99,
//            src/ode_solver/low_storage_runge_kutta_ode_solver.cpp:24-79
#pragma once
#include <hip/hip_runtime.h>

#include "dg_physics.hpp"

namespace dgx {

constexpr int MAXN1D = 8;

struct Ops1D {
    double D[MAXN1D][MAXN1D];   // GLL derivative matrix D[a][b] = phi_b'(x_a)
    double Sk[MAXN1D][MAXN1D];  // (Q - Q^T)[a][b] / w_a  with Q = W D
    double w[MAXN1D];           // GLL weights on [0,1]
    double inv_wb;              // 1 / w_0 (boundary weight)
};

enum StageMode { MODE_RESIDUAL = 0, MODE_RHS = 1, MODE_LSRK = 2, MODE_BUTCHER = 3, MODE_ACCUM = 4 };

struct StageArgs {
    const double *u;          // stage solution (read incl. neighbour face layers)
    const double *aux[3];     // lifted gradient per direction (Navier-Stokes)
    const int *nbr;           // [cell][2*dim]
    int ncells;
    double inv_h;             // 1/h
    double penalty;           // SIPG penalty (averaged over both sides)
    double resid_scale;       // h^dim (un-lifted residual = W_i h^dim du/dt)
    int mode;
    double *out;              // MODE_RESIDUAL / MODE_RHS output; MODE_BUTCHER: acc (or final)
    // low-storage RK 3S* (S1 ping-pong, S2 in place, S3 = u_n, S4 optional)
    double *s1_out, *s2, *s3, *s4;
    int lsrk_first;           // stage 0: S2 = delta*S1, S3 = S1 written
    double delta, g0, g1, g2, beta, bhat, dt;
    // Butcher explicit RK
    const double *u0;         // u_n
    double *acc;              // solution_update accumulator
    double *u_next;           // next stage solution (null at the last stage)
    double *k_store;          // where to store k_i (null if not needed later)
    const double *k_prev[4];  // stored k_j, j < i
    double a_next[4];         // a[i+1][j], j <= i
    double b_dt;              // dt * b_i
    int first_stage, last_stage, stage;
};

template <int D, int N> struct Shape {
    static constexpr int NP = (D == 3) ? N * N * N : N * N;
    static constexpr int NS = D + 2;
    static constexpr int EPB = (NP >= 256) ? 1 : (256 / NP);
    static constexpr int BLOCK = ((EPB * NP + 63) / 64) * 64;
};

template <int D, int N>
__device__ __forceinline__ void node_coords(int i, int *c)
{
    c[0] = i % N;
    c[1] = (i / N) % N;
    if (D == 3) c[2] = i / (N * N);
}

template <int D, int N>
__device__ __forceinline__ int node_with(int i, int k, int l)
{
    int c[3];
    node_coords<D, N>(i, c);
    c[k] = l;
    return (D == 3) ? c[0] + N * (c[1] + N * c[2]) : c[0] + N * c[1];
}

template <typename T, int D>
__device__ __forceinline__ void roundtrip(const T *u, const PhysConst &pc, T *ut)
{
    T v[D + 2];
    entropy_variables<T, D>(u, pc, v);
    conservative_from_entropy<T, D>(v, pc, ut);
}

// Lifted-gradient (auxiliary) equation, collocated GLL:
//   q_d = (1/h) [ D_d u + (1/w_0) (u_hat - u) n_d ]  on faces normal to d, u_hat = {{u}}.
template <int D, int N>
__global__ void __launch_bounds__((Shape<D, N>::BLOCK))
aux_kernel(const double *__restrict__ u, double *__restrict__ q0, double *__restrict__ q1,
           double *__restrict__ q2, const int *__restrict__ nbr, int ncells, double inv_h, Ops1D ops)
{
    using S = Shape<D, N>;
    constexpr int NP = S::NP, NS = S::NS, EPB = S::EPB;
    __shared__ double sU[EPB][NS][NP];
    const int tid = threadIdx.x;
    const int le = tid / NP, i = tid % NP;
    const int e = blockIdx.x * EPB + le;
    const bool active = (le < EPB) && (e < ncells);
    double ui[NS];
    if (active) {
        for (int s = 0; s < NS; ++s) {
            ui[s] = u[((long)e * NS + s) * NP + i];
            sU[le][s][i] = ui[s];
        }
    }
    __syncthreads();
    if (!active) return;
    int c[3];
    node_coords<D, N>(i, c);
    double *qout[3] = {q0, q1, q2};
    for (int d = 0; d < D; ++d) {
        double g[NS];
        for (int s = 0; s < NS; ++s) {
            double acc = 0.0;
            for (int l = 0; l < N; ++l) acc += ops.D[c[d]][l] * sU[le][s][node_with<D, N>(i, d, l)];
            g[s] = acc;
        }
        for (int side = 0; side < 2; ++side) {
            if (c[d] != (side ? N - 1 : 0)) continue;
            const int nb = nbr[e * 2 * D + 2 * d + side];
            const int j = node_with<D, N>(i, d, side ? 0 : N - 1);
            const double sgn = side ? 1.0 : -1.0;
            for (int s = 0; s < NS; ++s) {
                const double un = u[((long)nb * NS + s) * NP + j];
                const double avg = 0.5 * (ui[s] + un);
                g[s] += ops.inv_wb * ((avg - ui[s]) * sgn);
            }
        }
        for (int s = 0; s < NS; ++s) qout[d][((long)e * NS + s) * NP + i] = g[s] * inv_h;
    }
}

// Residual / stage kernel. SPLIT: two-point flux-differencing volume term;
// NSV: Navier-Stokes viscous terms; RT: entropy projection round trip (exact
// reference behaviour for the split form).
template <int D, int N, bool SPLIT, bool NSV, bool RT>
__global__ void __launch_bounds__((Shape<D, N>::BLOCK))
stage_kernel(StageArgs a, Ops1D ops, PhysConst pc)
{
    using S = Shape<D, N>;
    constexpr int NP = S::NP, NS = S::NS, EPB = S::EPB;
    __shared__ double sU[EPB][NS][NP];
    __shared__ double sF[EPB][D][NS][NP];
    const int tid = threadIdx.x;
    const int le = tid / NP, i = tid % NP;
    const int e = blockIdx.x * EPB + le;
    const bool active = (le < EPB) && (e < a.ncells);
    const double gamm1 = pc.gamm1;

    double ui[NS], ut[NS], qi[D][NS];
    if (active) {
        for (int s = 0; s < NS; ++s) ui[s] = a.u[((long)e * NS + s) * NP + i];
        if (NSV)
            for (int d = 0; d < D; ++d)
                for (int s = 0; s < NS; ++s) qi[d][s] = a.aux[d][((long)e * NS + s) * NP + i];
        if (SPLIT && RT) roundtrip<double, D>(ui, pc, ut);
        else
            for (int s = 0; s < NS; ++s) ut[s] = ui[s];
        for (int s = 0; s < NS; ++s) sU[le][s][i] = SPLIT ? ut[s] : ui[s];
        for (int k = 0; k < D; ++k) {
            double f[NS];
            for (int s = 0; s < NS; ++s) f[s] = 0.0;
            if (!SPLIT) convective_flux_dir<double, D>(ui, k, gamm1, f);
            if (NSV) {
                double g[NS * D], fv[NS];
                for (int s = 0; s < NS; ++s)
                    for (int d = 0; d < D; ++d) g[s * D + d] = qi[d][s];
                viscous_flux_dir<double, D>(SPLIT ? ut : ui, g, k, pc, fv);
                for (int s = 0; s < NS; ++s) f[s] += fv[s];
            }
            for (int s = 0; s < NS; ++s) sF[le][k][s][i] = f[s];
        }
    }
    __syncthreads();
    if (!active) return;

    int c[3];
    node_coords<D, N>(i, c);
    double acc[NS];
    for (int s = 0; s < NS; ++s) acc[s] = 0.0;

    // ---- volume -------------------------------------------------------
    if (SPLIT) {
        const Prim<double, D> qa = to_prim<double, D>(ut, gamm1);
        for (int k = 0; k < D; ++k) {
            for (int l = 0; l < N; ++l) {
                if (l == c[k]) continue;
                const int j = node_with<D, N>(i, k, l);
                double uj[NS];
                for (int s = 0; s < NS; ++s) uj[s] = sU[le][s][j];
                const Prim<double, D> qb = to_prim<double, D>(uj, gamm1);
                double F[NS];
                two_point_flux_dir<double, D>(ut, qa, uj, qb, k, pc, F);
                const double sk = ops.Sk[c[k]][l];
                for (int s = 0; s < NS; ++s) acc[s] -= sk * F[s];
            }
        }
    }
    if (!SPLIT || NSV) {
        for (int k = 0; k < D; ++k)
            for (int l = 0; l < N; ++l) {
                const int j = node_with<D, N>(i, k, l);
                const double dk = ops.D[c[k]][l];
                for (int s = 0; s < NS; ++s) acc[s] -= dk * sF[le][k][s][j];
            }
    }

    // ---- faces --------------------------------------------------------
    for (int k = 0; k < D; ++k) {
        for (int side = 0; side < 2; ++side) {
            if (c[k] != (side ? N - 1 : 0)) continue;
            const int nb = a.nbr[e * 2 * D + 2 * k + side];
            const int j = node_with<D, N>(i, k, side ? 0 : N - 1);
            const double sgn = side ? 1.0 : -1.0;
            double un[NS], unt[NS];
            for (int s = 0; s < NS; ++s) un[s] = a.u[((long)nb * NS + s) * NP + j];
            if (SPLIT && RT) roundtrip<double, D>(un, pc, unt);
            else
                for (int s = 0; s < NS; ++s) unt[s] = un[s];
            const double *sa = SPLIT ? ut : ui;
            const double *sb = SPLIT ? unt : un;
            const Prim<double, D> pa = to_prim<double, D>(sa, gamm1);
            const Prim<double, D> pb = to_prim<double, D>(sb, gamm1);
            double fs[NS];
            if (e < nb) {
                numerical_flux<double, D>(sa, pa, sb, pb, k, sgn, pc, fs);
            } else {
                numerical_flux<double, D>(sb, pb, sa, pa, k, -sgn, pc, fs);
                for (int s = 0; s < NS; ++s) fs[s] = -fs[s];
            }
            double ft[NS];
            for (int s = 0; s < NS; ++s) ft[s] = -fs[s];
            if (!SPLIT)
                for (int s = 0; s < NS; ++s) ft[s] += sgn * sF[le][k][s][i];
            if (NSV) {
                double gi[NS * D], gn[NS * D], jump[NS * D];
                for (int s = 0; s < NS; ++s)
                    for (int d = 0; d < D; ++d) {
                        gi[s * D + d] = qi[d][s];
                        gn[s * D + d] = a.aux[d][((long)nb * NS + s) * NP + j];
                        jump[s * D + d] = (d == k) ? (ui[s] - un[s]) * sgn : (ui[s] - un[s]) * 0.0;
                    }
                double fa[NS], fb[NS];
                viscous_flux_dir<double, D>(ui, gi, k, pc, fa);
                viscous_flux_dir<double, D>(un, gn, k, pc, fb);
                double fvs[NS];
                for (int s = 0; s < NS; ++s) fvs[s] = 0.5 * (fa[s] * sgn + fb[s] * sgn);
                if (pc.diss_flux == 0) {
                    double ja[NS], jb[NS];
                    viscous_flux_dir<double, D>(ui, jump, k, pc, ja);
                    viscous_flux_dir<double, D>(un, jump, k, pc, jb);
                    for (int s = 0; s < NS; ++s) fvs[s] = fvs[s] - a.penalty * (0.5 * (ja[s] * sgn + jb[s] * sgn));
                }
                for (int s = 0; s < NS; ++s) ft[s] -= fvs[s];
                if (SPLIT)
                    for (int s = 0; s < NS; ++s) ft[s] += sgn * fa[s];
            }
            for (int s = 0; s < NS; ++s) acc[s] += ops.inv_wb * ft[s];
        }
    }

    double dudt[NS];
    for (int s = 0; s < NS; ++s) dudt[s] = acc[s] * a.inv_h;

    // ---- epilogue -----------------------------------------------------
    const long base = (long)e * NS * NP + i;
    if (a.mode == MODE_RHS) {
        for (int s = 0; s < NS; ++s) a.out[base + s * NP] = dudt[s];
    } else if (a.mode == MODE_RESIDUAL) {
        double W = ops.w[c[0]] * ops.w[c[1]];
        if (D == 3) W *= ops.w[c[2]];
        for (int s = 0; s < NS; ++s) a.out[base + s * NP] = dudt[s] * (W * a.resid_scale);
    } else if (a.mode == MODE_ACCUM) {
        // register += bhat * (dt * R)   (low_storage_runge_kutta_ode_solver.cpp:66-68)
        for (int s = 0; s < NS; ++s) {
            const double r = dudt[s] * a.dt;
            a.out[base + s * NP] = a.out[base + s * NP] + a.bhat * r;
        }
    } else if (a.mode == MODE_LSRK) {
        for (int s = 0; s < NS; ++s) {
            const long idx = base + s * NP;
            const double s1 = ui[s];
            const double s2 = a.lsrk_first ? 0.0 + a.delta * s1 : a.s2[idx] + a.delta * s1;
            const double s3 = a.lsrk_first ? s1 : a.s3[idx];
            double n1 = s1 * a.g0;
            n1 += a.g1 * s2;
            n1 += a.g2 * s3;
            const double r = dudt[s] * a.dt;
            n1 += a.beta * r;
            a.s2[idx] = s2;
            if (a.lsrk_first) a.s3[idx] = s1;
            a.s1_out[idx] = n1;
            if (a.s4) a.s4[idx] = (a.lsrk_first ? s1 : a.s4[idx]) + a.bhat * r;
        }
    } else { // MODE_BUTCHER
        for (int s = 0; s < NS; ++s) {
            const long idx = base + s * NP;
            const double k = dudt[s];
            double accv = a.first_stage ? a.u0[idx] : a.acc[idx];
            accv += a.b_dt * k;
            a.acc[idx] = accv;
            if (a.k_store) a.k_store[idx] = k;
            if (a.u_next) {
                // rk_stage[i+1] = (sum_j a_{i+1,j} k_j) * dt + u_n  (runge_kutta_ode_solver.cpp:15-33)
                double st = 0.0;
                for (int j2 = 0; j2 < a.stage; ++j2)
                    if (a.a_next[j2] != 0.0) st += a.a_next[j2] * a.k_prev[j2][idx];
                if (a.a_next[a.stage] != 0.0) st += a.a_next[a.stage] * k;
                st *= a.dt;
                st += 1.0 * a.u0[idx];
                a.u_next[idx] = st;
            }
        }
    }
}

} // namespace dgx
