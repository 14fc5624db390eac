// Shared pieces of the strong-form DG kernels for collocated Gauss-Lobatto
// nodes on periodic Cartesian hexahedra/quadrilaterals: operator tables, stage
// arguments, node indexing, and the auxiliary (lifted-gradient) kernel.  The
// fused stage kernel is in dg_stage2.hpp.
// Reference path restated:
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
#include <type_traits>
#include <hip/hip_runtime.h>

#include "dg_physics.hpp"

namespace dgx {

constexpr int MAXN1D = 8;

struct Ops1D {
    double D[MAXN1D][MAXN1D];   // GLL derivative matrix D[a][b] = phi_b'(x_a)
    double Sk[MAXN1D][MAXN1D];  // (Q - Q^T)[a][b] / w_a  with Q = W D
    double w[MAXN1D];           // GLL weights on [0,1]
    double inv_wb;              // 1 / w_0 (boundary weight)
    float Skf[MAXN1D][MAXN1D];  // Sk rounded to float (mixed path)
    float Df[MAXN1D][MAXN1D];   // D rounded to float (mixed path)
    float inv_wbf;              // inv_wb rounded to float
};

enum StageMode { MODE_RESIDUAL = 0, MODE_RHS = 1, MODE_LSRK = 2, MODE_BUTCHER = 3, MODE_ACCUM = 4 };

struct StageArgs {
    const double *u;          // stage solution (read incl. neighbour face layers)
    const double *aux[3];     // lifted gradient per direction (Navier-Stokes, FP64 path)
    const float *auxf[3];     // the same stored in float (Navier-Stokes, mixed path)
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
    double *acc;              // solution_update accumulator (read)
    double *acc_out;          // where the updated accumulator is written (acc, or u_n at the last stage)
    double *u_next;           // next stage solution (null at the last stage)
    double *k_store;          // where to store k_i (null if not needed later)
    const double *k_prev[4];  // stored k_j, j < i
    double a_next[4];         // a[i+1][j], j <= i
    double b_dt;              // dt * b_i
    int first_stage, last_stage, stage;
    int acc_zero;             // relaxation RK: the accumulator holds dt sum b_i k_i (starts at 0)
    // set to 1 by any thread that finds a non-physical stage state (density or pressure
    // negative or NaN; euler.cpp:155-173, 486-500); read and cleared by the host
    int *status;
    // curvilinear grid: [cell][dim*dim + 1][node], the metric cofactor C[m][n] = J dxi_n/dx_m in
    // slot m*dim + n and det J in slot dim*dim; null on the affine box
    const double *met;
    // curvilinear grid: SIPG penalty of each cell's faces normal to each direction [cell][dim]
    const double *pen;
};

// Non-physical state check of one node (the reference checks every density and pressure it
// computes, euler.cpp:155-173, 486-500; NaN also fails here).  A plain global store of the
// same value from every offending lane: no atomics needed.
template <typename T>
__device__ __forceinline__ void flag_nonphysical(const StageArgs &a, T rho, T p)
{
    if (!(rho >= T(0)) || !(p >= T(0))) *a.status = 1;
}

// Lifted gradient component (state s, direction d) at flat index idx of the
// layout [cell][state][node], in the precision of the path's fluxes.
template <typename T> __device__ __forceinline__ T aux_at(const StageArgs &a, int d, long idx)
{
    if constexpr (std::is_same<T, float>::value) return a.auxf[d][idx];
    else return T(a.aux[d][idx]);
}

template <int D, int N> struct Shape {
    static constexpr int NP = (D == 3) ? N * N * N : N * N;
    static constexpr int NS = D + 2;
    static constexpr int EPB = (NP >= 256) ? 1 : (256 / NP);
    static constexpr int BLOCK = ((EPB * NP + 63) / 64) * 64;
};

// XCD-aware block order.  Workgroups are dispatched to the 8 XCDs of the chip
// round-robin (hardware block b runs on XCD b % 8), and each XCD has its own
// L2.  Remapping so that XCD x processes a contiguous range of logical blocks
// gives every XCD a compact region of the (Morton- or lexicographically
// ordered) mesh, so the face neighbours an element reads were mostly loaded
// into the same L2 by nearby blocks.  A bijection for any grid size.
constexpr int NUM_XCD = 8;
__device__ __forceinline__ int xcd_block()
{
    const int nb = gridDim.x, b = blockIdx.x;
    const int q = nb / NUM_XCD, r = nb % NUM_XCD;
    const int xcd = b % NUM_XCD, slot = b / NUM_XCD;
    return (xcd < r) ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
}

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

// Streaming access to the RK registers: each value is read or written once per
// stage and the registers together exceed the caches, so the stage kernels
// mark these accesses non-temporal (STREAM_NT) to keep L2 for the neighbour
// face traces.
constexpr bool STREAM_NT = true;
__device__ __forceinline__ double ld_stream(const double *p)
{
    if constexpr (STREAM_NT) return __builtin_nontemporal_load(p);
    else return *p;
}
__device__ __forceinline__ void st_stream(double *p, double v)
{
    if constexpr (STREAM_NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// Phase E of the stage kernels, in FP64: du/dt = acc / h and the output of the
// selected mode (residual, lifted right-hand side, low-storage RK 3S* register
// update, Butcher-tableau stage update, or the FSAL error-register accumulation).
// jac > 0: the node's det J on a curvilinear grid (acc then holds the cofactor-weighted terms):
// du/dt = acc / J and the un-lifted residual W J du/dt.
template <int D, int N, typename T>
__device__ __forceinline__ void stage_epilogue(const StageArgs &a, const Ops1D &ops, int e, int i, const int *c,
                                               const double *ud, const T *acc, bool use_pf = false,
                                               const double *pf2 = nullptr, const double *pf3 = nullptr,
                                               double jac = 0.0)
{
    constexpr int NP = Shape<D, N>::NP, NS = D + 2;
    const double inv_h = jac > 0.0 ? 1.0 / jac : a.inv_h;
    const double resid_scale = jac > 0.0 ? jac : a.resid_scale;
    double dudt[NS];
    for (int s = 0; s < NS; ++s) dudt[s] = double(acc[s]) * inv_h;
    const long base = (long)e * NS * NP + i;
    if (a.mode == MODE_RHS) {
        for (int s = 0; s < NS; ++s) a.out[base + s * NP] = dudt[s];
    } else if (a.mode == MODE_RESIDUAL) {
        double W = ops.w[c[0]] * ops.w[c[1]];
        if (D == 3) W *= ops.w[c[2]];
        for (int s = 0; s < NS; ++s) a.out[base + s * NP] = dudt[s] * (W * resid_scale);
    } else if (a.mode == MODE_ACCUM) {
        for (int s = 0; s < NS; ++s) {
            const double r = dudt[s] * a.dt;
            a.out[base + s * NP] = a.out[base + s * NP] + a.bhat * r;
        }
    } else if (a.mode == MODE_LSRK) {
        for (int s = 0; s < NS; ++s) {
            const long idx = base + s * NP;
            const double s1 = ud[s];
            // use_pf: S2, S3 already loaded by the caller into pf2 / pf3 (non-first stages)
            const double s2 = a.lsrk_first ? 0.0 + a.delta * s1 : (use_pf ? pf2[s] : ld_stream(a.s2 + idx)) + a.delta * s1;
            const double s3 = a.lsrk_first ? s1 : (use_pf ? pf3[s] : ld_stream(a.s3 + idx));
            double n1 = s1 * a.g0;
            n1 += a.g1 * s2;
            n1 += a.g2 * s3;
            const double r = dudt[s] * a.dt;
            n1 += a.beta * r;
            st_stream(a.s2 + idx, s2);
            if (a.lsrk_first) st_stream(a.s3 + idx, s1);
            st_stream(a.s1_out + idx, n1);
            if (a.s4) a.s4[idx] = (a.lsrk_first ? s1 : a.s4[idx]) + a.bhat * r;
        }
    } else {
        for (int s = 0; s < NS; ++s) {
            const long idx = base + s * NP;
            const double kk = dudt[s];
            double accv = a.first_stage ? (a.acc_zero ? 0.0 : a.u0[idx]) : a.acc[idx];
            accv += a.b_dt * kk;
            a.acc_out[idx] = accv;
            if (a.k_store) a.k_store[idx] = kk;
            if (a.u_next) {
                // stage indices unrolled so the register file is never indexed per lane
                double st = 0.0, a_cur = 0.0;
#pragma unroll
                for (int j2 = 0; j2 < 4; ++j2) {
                    if (j2 < a.stage && a.a_next[j2] != 0.0) st += a.a_next[j2] * a.k_prev[j2][idx];
                    if (j2 == a.stage) a_cur = a.a_next[j2];
                }
                if (a_cur != 0.0) st += a_cur * kk;
                st *= a.dt;
                st += 1.0 * a.u0[idx];
                a.u_next[idx] = st;
            }
        }
    }
}

// Lifted-gradient (auxiliary) equation, collocated GLL:
//   q_d = (1/h) [ D_d u + (1/w_0) (u_hat - u) n_d ]  on faces normal to d, u_hat = {{u}}.
// Q: storage type of q (double for the FP64 path, float for the mixed path, whose
// kernels evaluate the viscous fluxes in float).
// met: curvilinear grid (the metric terms [cell][dim*dim + 1][node]; null on the affine box):
//   q_k = (1/J) [ sum_j C[k][j] D_j u + (1/w_0) (u_hat - u) N_k ] with N the face's physical normal
//   C n_ref at the node of the face's owner (the lower cell index; this cell on a boundary).
template <int D, int N, typename Q>
__global__ void __launch_bounds__((Shape<D, N>::BLOCK))
aux_kernel(const double *__restrict__ u, Q *__restrict__ q0, Q *__restrict__ q1,
           Q *__restrict__ q2, const int *__restrict__ nbr, int ncells, double inv_h, Ops1D ops, PhysConst pc,
           const double *__restrict__ met)
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
    Q *qout[3] = {q0, q1, q2};
    constexpr int MS = D * D + 1;
    double qp[D][NS];  // curvilinear grid: J times the physical gradient
    if (met)
        for (int k = 0; k < D; ++k)
            for (int s = 0; s < NS; ++s) qp[k][s] = 0.0;
    for (int d = 0; d < D; ++d) {
        double g[NS];
        for (int s = 0; s < NS; ++s) {
            double acc = 0.0;
            for (int l = 0; l < N; ++l) acc += ops.D[c[d]][l] * sU[le][s][node_with<D, N>(i, d, l)];
            g[s] = acc;
        }
        if (met)
            for (int k = 0; k < D; ++k) {
                const double ckd = met[((long)e * MS + k * D + d) * NP + i];
                for (int s = 0; s < NS; ++s) qp[k][s] += ckd * g[s];
            }
        for (int side = 0; side < 2; ++side) {
            if (c[d] != (side ? N - 1 : 0)) continue;
            const int nb = nbr[e * 2 * D + 2 * d + side];
            const int j = node_with<D, N>(i, d, side ? 0 : N - 1);
            const double sgn = side ? 1.0 : -1.0;
            // a wall face (nb < 0: -type): the viscous wall state of the node
            // (assemble_boundary_term_auxiliary_equation, strong_dg.cpp:790-870)
            double ub[NS];
            if (nb < 0) visc_wall_state<double, D>(-nb, ui, pc, ub);
            double lift[NS];
            for (int s = 0; s < NS; ++s) {
                const double un = (nb < 0) ? ub[s] : u[((long)nb * NS + s) * NP + j];
                const double avg = 0.5 * (ui[s] + un);
                lift[s] = ops.inv_wb * (avg - ui[s]);
            }
            if (met) {
                // the face's physical normal is the owner's (the lower cell index, or this cell on a
                // boundary): sgn C[:, d] at the owner's face node
                const bool own = nb < 0 || e < nb;
                for (int k = 0; k < D; ++k) {
                    const double nk = sgn * (own ? met[((long)e * MS + k * D + d) * NP + i]
                                                 : met[((long)(nb < 0 ? 0 : nb) * MS + k * D + d) * NP + j]);
                    for (int s = 0; s < NS; ++s) qp[k][s] += lift[s] * nk;
                }
            } else {
                for (int s = 0; s < NS; ++s) g[s] += lift[s] * sgn;
            }
        }
        if (!met)
            for (int s = 0; s < NS; ++s) qout[d][((long)e * NS + s) * NP + i] = Q(g[s] * inv_h);
    }
    if (met) {
        const double inv_j = 1.0 / met[((long)e * MS + D * D) * NP + i];
        for (int k = 0; k < D; ++k)
            for (int s = 0; s < NS; ++s) qout[k][((long)e * NS + s) * NP + i] = Q(qp[k][s] * inv_j);
    }
}

} // namespace dgx
