// Cached-node point physics of the line kernels, direction K fixed at compile
// time, in the arithmetic type T: float on the mixed path, double on the FP64
// path.  Each function is an algebraic restatement of the point physics in
// dg_physics.hpp (same formula in exact arithmetic) arranged to use few
// divisions: per-node quantities (1/rho, pressure, beta, sqrt(rho), sound
// speed, enthalpy) are computed once per node and reused by every pair and
// face that touches the node.  In float the reciprocals, square roots and
// logarithms are the hardware v_rcp_f32 / v_sqrt_f32 / v_log_f32 instructions;
// in double the reciprocal is v_rcp_f64 with two Newton-Raphson steps, square
// root and logarithm the library's double routines.
//
//   ec_flux_fast   Chandrashekar / Ranocha two-point flux (euler.cpp:709-807):
//                  log means from f = (a-b)/(a+b), and 1/(2 beta_ln (gamma-1))
//                  and p_hat/rho_ln expressed through F_beta, F_rho and
//                  1/(beta1+beta2), so a pair costs three reciprocals.
//   roe_fast       Roe-Pike + Harten entropy fix (convective_numerical_flux.cpp:
//                  332-480) with Roe averages from sqrt(rho) of each side.
//   num_flux_fast  baseline (central or two-point) + dissipation (LxF / Roe)
//                  (convective_numerical_flux.cpp:34-54).
#pragma once
#include "dg_physics.hpp"

namespace dgx {

template <typename T> struct Fast;
#ifndef DGX_HOST_TEST
template <> struct Fast<float> {
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
    static __device__ __forceinline__ float log_(float x) { return __logf(x); }
};
template <> struct Fast<double> {
    // v_rcp_f64 refined by two Newton-Raphson steps r <- r + r (1 - x r): v_rcp_f64 alone is within
    // 4.6e-8 of 1/x, one step within 11 ulp, two steps gave the correctly rounded 1/x for all of 2^20
    // samples over [e^-7, e^7] on gfx950 -- at half the instructions of the IEEE division sequence
    // (FP64 stage kernel at p=3: 206 -> 197 us per stage; one step only would save 2 %)
    static __device__ __forceinline__ double rcp(double x)
    {
        double r = __builtin_amdgcn_rcp(x);
        double e = __builtin_fma(-x, r, 1.0);
        r = __builtin_fma(r, e, r);
        e = __builtin_fma(-x, r, 1.0);
        return __builtin_fma(r, e, r);
    }
    static __device__ __forceinline__ double sqrt_(double x) { return sqrt(x); }
    static __device__ __forceinline__ double log_(double x) { return log(x); }
};
#else
// host build of the same formulas (CPU checks of the restatement)
template <typename T> struct FastHost {
    static inline T rcp(T x) { return T(1) / x; }
    static inline T sqrt_(T x) { return std::sqrt(x); }
    static inline T log_(T x) { return std::log(x); }
};
template <> struct Fast<float> : FastHost<float> {};
template <> struct Fast<double> : FastHost<double> {};
#endif
template <typename T> __device__ __forceinline__ T frcp(T x) { return Fast<T>::rcp(x); }
template <typename T> __device__ __forceinline__ T fsqrt(T x) { return Fast<T>::sqrt_(x); }
template <typename T> __device__ __forceinline__ T flog(T x) { return Fast<T>::log_(x); }
template <typename T> __device__ __forceinline__ T fmx(T a, T b) { return a > b ? a : b; }
template <typename T> __device__ __forceinline__ T fab(T a) { return a < T(0) ? -a : a; }

// Per-node cache: conservative state and derived quantities.
template <typename T, int D> struct NodeC {
    T u[D + 2];
    T vel[D];
    T p, b, v2, sr, c, h;  // pressure, beta, |v|^2, sqrt(rho), sound speed, total enthalpy
};

template <typename T> struct FastConst {
    T gam, gamm1, inv_gamm1;
    int two_point, conv_flux;
};

template <typename T> __device__ __forceinline__ FastConst<T> fast_const(const PhysConst &pc)
{
    FastConst<T> f;
    f.gam = T(pc.gam);
    f.gamm1 = T(pc.gamm1);
    f.inv_gamm1 = T(1.0 / pc.gamm1);
    f.two_point = pc.two_point;
    f.conv_flux = pc.conv_flux;
    return f;
}

// Fills every field of q from q.u; beta is rho/(2p) (Chandrashekar) or rho/p (Ranocha).
template <typename T, int D>
__device__ __forceinline__ void node_fill(NodeC<T, D> &q, const FastConst<T> &fc)
{
    const T ir = frcp<T>(q.u[0]);
    T v2 = T(0);
    for (int k = 0; k < D; ++k) {
        q.vel[k] = q.u[1 + k] * ir;
        v2 += q.vel[k] * q.vel[k];
    }
    q.v2 = v2;
    q.p = fc.gamm1 * (q.u[D + 1] - T(0.5) * q.u[0] * v2);
    const T rp = q.u[0] * frcp<T>(q.p);
    q.b = (fc.two_point == 3) ? rp : T(0.5) * rp;
    q.sr = fsqrt<T>(q.u[0]);
    q.c = fsqrt<T>(fc.gam * q.p * ir);
    q.h = (q.u[D + 1] + q.p) * ir;
}

// F of the Ismail-Roe log mean  mean_ln(a, b) = (a + b) / (2 F),  f = (a - b) / (a + b)
// (series branch below 1e-2 as euler.cpp:629-647, Horner form with constant reciprocals).
template <typename T>
__device__ __forceinline__ T log_mean_F(T a, T b, T f)
{
    const T uu = f * f;
    if (uu < T(1.0e-2)) return T(1) + uu * (T(1.0 / 3.0) + uu * (T(1.0 / 5.0) + uu * T(1.0 / 7.0)));
    return T(0.5) * flog<T>(a * frcp<T>(b)) * frcp<T>(f);
}

template <typename T, int D, int K>
__device__ __forceinline__ void ec_flux_fast(const T rho1, const T *v1, const T p1, const T b1, const T q1v2,
                                             const T rho2, const T *v2, const T p2, const T b2, const T q2v2,
                                             const FastConst<T> &fc, T *f)
{
    const bool ra = (fc.two_point == 3);
    const T sr = rho1 + rho2, isr = frcp<T>(sr);
    const T Fr = log_mean_F<T>(rho1, rho2, (rho1 - rho2) * isr);
    const T sb = b1 + b2, isb = frcp<T>(sb);
    const T Fb = log_mean_F<T>(b1, b2, (b1 - b2) * isb);
    const T rho_ln = T(0.5) * sr * frcp<T>(Fr);
    T va[D], vsq = T(0);
    for (int k = 0; k < D; ++k) {
        va[k] = T(0.5) * (v1[k] + v2[k]);
        vsq += va[k] * va[k];
    }
    T p_hat, H;
    if (ra) {
        p_hat = T(0.5) * (p1 + p2);
        H = T(2) * Fb * isb * fc.inv_gamm1 + T(4) * p_hat * Fr * isr;
    } else {
        p_hat = T(0.5) * sr * isb;
        H = isb * (Fb * fc.inv_gamm1 + Fr);
    }
    H += vsq - T(0.25) * (q1v2 + q2v2);
    const T m = rho_ln * va[K];
    f[0] = m;
    for (int vd = 0; vd < D; ++vd) f[1 + vd] = m * va[vd];
    f[1 + K] += p_hat;
    f[D + 1] = m * H;
    if (ra) f[D + 1] -= T(0.5) * (p1 * v1[K] + p2 * v2[K]);
}

// Two-point flux of two cached nodes; Kennedy-Gruber and Ismail-Roe go through
// the generic evaluation of dg_physics.hpp.
template <typename T, int D, int K>
__device__ __forceinline__ void pair_flux_fast(const NodeC<T, D> &a, const NodeC<T, D> &b, const FastConst<T> &fc,
                                               const PhysConst &pc, T *f)
{
    if (fc.two_point >= 2) {
        ec_flux_fast<T, D, K>(a.u[0], a.vel, a.p, a.b, a.v2, b.u[0], b.vel, b.p, b.b, b.v2, fc, f);
    } else {
        Prim<T, D> qa, qb;
        qa.rho = a.u[0];
        qb.rho = b.u[0];
        for (int k = 0; k < D; ++k) {
            qa.vel[k] = a.vel[k];
            qb.vel[k] = b.vel[k];
        }
        qa.p = a.p;
        qb.p = b.p;
        two_point_flux_dir<T, D>(a.u, qa, b.u, qb, K, pc, f);
    }
}

template <typename T, int D, int K>
__device__ __forceinline__ void roe_fast(const NodeC<T, D> &L, const NodeC<T, D> &R, T sgn, const FastConst<T> &fc,
                                         T *out)
{
    const T vnL = L.vel[K] * sgn, vnR = R.vel[K] * sgn;
    const T isum = frcp<T>(L.sr + R.sr);
    const T rho_ravg = L.sr * R.sr;
    T v_ravg[D], vel2 = T(0);
    for (int k = 0; k < D; ++k) {
        v_ravg[k] = (R.sr * R.vel[k] + L.sr * L.vel[k]) * isum;
        vel2 += v_ravg[k] * v_ravg[k];
    }
    const T h_ravg = (R.sr * R.h + L.sr * L.h) * isum;
    const T vn = v_ravg[K] * sgn;
    const T c2 = fc.gamm1 * (h_ravg - T(0.5) * vel2);
    const T c = (c2 > T(0)) ? fsqrt<T>(c2) : T(1e10);
    T eig[3] = {fab<T>(vn - c), fab<T>(vn), fab<T>(vn + c)};
    const T eL[3] = {fab<T>(vnL - L.c), fab<T>(vnL), fab<T>(vnL + L.c)};
    const T eR[3] = {fab<T>(vnR - R.c), fab<T>(vnR), fab<T>(vnR + R.c)};
    for (int e = 0; e < 3; ++e) {
        const T eps = fmx<T>(fab<T>(eig[e] - eL[e]), fab<T>(eR[e] - eig[e]));
        if (eig[e] < eps) eig[e] = T(0.5) * (eig[e] * eig[e] * frcp<T>(eps) + eps);
    }
    const T dp = R.p - L.p, drho = R.u[0] - L.u[0];
    const T dVn = vnR - vnL;
    const T hic2 = T(0.5) * frcp<T>(c2);
    const T rc = rho_ravg * c * dVn;
    const T co0 = eig[0] * (dp - rc) * hic2;
    const T co1 = eig[1] * (drho - T(2) * dp * hic2);
    const T co2 = eig[1] * rho_ravg;
    const T co3 = eig[2] * (dp + rc) * hic2;
    // tangential velocity jump: the normal component cancels exactly
    T dvt_dot = T(0);
    T A[D + 2];
    A[0] = co0 + co1 + co3;
    for (int k = 0; k < D; ++k) {
        const T dvt = (k == K) ? T(0) : (R.vel[k] - L.vel[k]);
        dvt_dot += v_ravg[k] * dvt;
        const T nk = (k == K) ? sgn : T(0);
        A[1 + k] = co0 * (v_ravg[k] - c * nk) + co1 * v_ravg[k] + co2 * dvt + co3 * (v_ravg[k] + c * nk);
    }
    A[D + 1] = co0 * (h_ravg - c * vn) + co1 * vel2 * T(0.5) + co2 * dvt_dot + co3 * (h_ravg + c * vn);
    for (int s = 0; s < D + 2; ++s) out[s] = T(-0.5) * A[s];
}

// Numerical flux of a face node along the normal sgn * e_K, L the interior
// (owner) side.
template <typename T, int D, int K>
__device__ __forceinline__ void num_flux_fast(const NodeC<T, D> &L, const NodeC<T, D> &R, T sgn,
                                              const FastConst<T> &fc, const PhysConst &pc, T *out)
{
    const int cf = fc.conv_flux;
    T base[D + 2];
    if (cf >= 3) {
        pair_flux_fast<T, D, K>(L, R, fc, pc, base);
        for (int s = 0; s < D + 2; ++s) base[s] *= sgn;
    } else {
        T f1[D + 2], f2[D + 2];
        convective_flux_dir<T, D>(L.u, K, fc.gamm1, f1);
        convective_flux_dir<T, D>(R.u, K, fc.gamm1, f2);
        for (int s = 0; s < D + 2; ++s) base[s] = (T(0.5) * (f1[s] + f2[s])) * sgn;
    }
    T diss[D + 2];
    for (int s = 0; s < D + 2; ++s) diss[s] = T(0);
    if (cf == 0 || cf == 4) {
        const T lam = fmx<T>(fab<T>(L.vel[K]) + L.c, fab<T>(R.vel[K]) + R.c);
        for (int s = 0; s < D + 2; ++s) diss[s] = T(-0.5) * lam * (R.u[s] - L.u[s]);
    }
    if (cf == 1 || cf == 5) roe_fast<T, D, K>(L, R, sgn, fc, diss);
    for (int s = 0; s < D + 2; ++s) out[s] = base[s] + diss[s];
}

} // namespace dgx
