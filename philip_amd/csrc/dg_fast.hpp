// Float point physics of the mixed-precision path, direction K fixed at
// compile time.  Each function is an algebraic restatement of the FP64 point
// physics in dg_physics.hpp (same formula in exact arithmetic) arranged to
// use few divisions: per-node quantities (1/rho, pressure, beta, sqrt(rho),
// sound speed, enthalpy) are computed once per node and reused by every pair
// and face that touches the node, and the reciprocals go through the
// hardware v_rcp_f32 / v_sqrt_f32 / v_log_f32 instructions.
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

#ifndef DGX_HOST_TEST
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float flog(float x) { return __logf(x); }
#else
// host build of the same formulas (CPU checks of the restatement)
inline float frcp(float x) { return 1.0f / x; }
inline float fsqrt(float x) { return std::sqrt(x); }
inline float flog(float x) { return std::log(x); }
#endif

// Per-node cache: conservative state and derived quantities.
template <int D> struct NodeF {
    float u[D + 2];
    float vel[D];
    float p, b, v2, sr, c, h;  // pressure, beta, |v|^2, sqrt(rho), sound speed, total enthalpy
};
template <int D> constexpr int node_f_count() { return 2 * D + 8; }

struct FastConst {
    float gam, gamm1, inv_gamm1;
    int two_point, conv_flux;
};

__device__ __forceinline__ FastConst fast_const(const PhysConst &pc)
{
    FastConst f;
    f.gam = float(pc.gam);
    f.gamm1 = float(pc.gamm1);
    f.inv_gamm1 = float(1.0 / pc.gamm1);
    f.two_point = pc.two_point;
    f.conv_flux = pc.conv_flux;
    return f;
}

// Fills every field of q from q.u; beta is rho/(2p) (Chandrashekar) or rho/p (Ranocha).
template <int D>
__device__ __forceinline__ void node_fill(NodeF<D> &q, const FastConst &fc)
{
    const float ir = frcp(q.u[0]);
    float v2 = 0.f;
    for (int k = 0; k < D; ++k) {
        q.vel[k] = q.u[1 + k] * ir;
        v2 += q.vel[k] * q.vel[k];
    }
    q.v2 = v2;
    q.p = fc.gamm1 * (q.u[D + 1] - 0.5f * q.u[0] * v2);
    const float rp = q.u[0] * frcp(q.p);
    q.b = (fc.two_point == 3) ? rp : 0.5f * rp;
    q.sr = fsqrt(q.u[0]);
    q.c = fsqrt(fc.gam * q.p * ir);
    q.h = (q.u[D + 1] + q.p) * ir;
}

// F of the Ismail-Roe log mean  mean_ln(a, b) = (a + b) / (2 F),  f = (a - b) / (a + b).
__device__ __forceinline__ float log_mean_F(float a, float b, float f)
{
    const float uu = f * f;
    if (uu < 1.0e-2f) return 1.f + uu * (1.f / 3.f) + uu * uu * (1.f / 5.f) + uu * uu * uu * (1.f / 7.f);
    return 0.5f * flog(a * frcp(b)) * frcp(f);
}

template <int D, int K>
__device__ __forceinline__ void ec_flux_fast(const float rho1, const float *v1, const float p1, const float b1,
                                             const float q1v2, const float rho2, const float *v2, const float p2,
                                             const float b2, const float q2v2, const FastConst &fc, float *f)
{
    const bool ra = (fc.two_point == 3);
    const float sr = rho1 + rho2, isr = frcp(sr);
    const float Fr = log_mean_F(rho1, rho2, (rho1 - rho2) * isr);
    const float sb = b1 + b2, isb = frcp(sb);
    const float Fb = log_mean_F(b1, b2, (b1 - b2) * isb);
    const float rho_ln = 0.5f * sr * frcp(Fr);
    float va[D], vsq = 0.f;
    for (int k = 0; k < D; ++k) {
        va[k] = 0.5f * (v1[k] + v2[k]);
        vsq += va[k] * va[k];
    }
    float p_hat, H;
    if (ra) {
        p_hat = 0.5f * (p1 + p2);
        H = 2.f * Fb * isb * fc.inv_gamm1 + 4.f * p_hat * Fr * isr;
    } else {
        p_hat = 0.5f * sr * isb;
        H = isb * (Fb * fc.inv_gamm1 + Fr);
    }
    H += vsq - 0.25f * (q1v2 + q2v2);
    const float m = rho_ln * va[K];
    f[0] = m;
    for (int vd = 0; vd < D; ++vd) f[1 + vd] = m * va[vd];
    f[1 + K] += p_hat;
    f[D + 1] = m * H;
    if (ra) f[D + 1] -= 0.5f * (p1 * v1[K] + p2 * v2[K]);
}

// Two-point flux of two cached nodes; Kennedy-Gruber and Ismail-Roe go through
// the generic float evaluation.
template <int D, int K>
__device__ __forceinline__ void pair_flux_fast(const NodeF<D> &a, const NodeF<D> &b, const FastConst &fc,
                                               const PhysConst &pc, float *f)
{
    if (fc.two_point >= 2) {
        ec_flux_fast<D, K>(a.u[0], a.vel, a.p, a.b, a.v2, b.u[0], b.vel, b.p, b.b, b.v2, fc, f);
    } else {
        Prim<float, D> qa, qb;
        qa.rho = a.u[0];
        qb.rho = b.u[0];
        for (int k = 0; k < D; ++k) {
            qa.vel[k] = a.vel[k];
            qb.vel[k] = b.vel[k];
        }
        qa.p = a.p;
        qb.p = b.p;
        two_point_flux_dir<float, D>(a.u, qa, b.u, qb, K, pc, f);
    }
}

template <int D, int K>
__device__ __forceinline__ void roe_fast(const NodeF<D> &L, const NodeF<D> &R, float sgn, const FastConst &fc,
                                         float *out)
{
    const float vnL = L.vel[K] * sgn, vnR = R.vel[K] * sgn;
    const float isum = frcp(L.sr + R.sr);
    const float rho_ravg = L.sr * R.sr;
    float v_ravg[D], vel2 = 0.f;
    for (int k = 0; k < D; ++k) {
        v_ravg[k] = (R.sr * R.vel[k] + L.sr * L.vel[k]) * isum;
        vel2 += v_ravg[k] * v_ravg[k];
    }
    const float h_ravg = (R.sr * R.h + L.sr * L.h) * isum;
    const float vn = v_ravg[K] * sgn;
    const float c2 = fc.gamm1 * (h_ravg - 0.5f * vel2);
    const float c = (c2 > 0.f) ? fsqrt(c2) : 1e10f;
    float eig[3] = {fabsf(vn - c), fabsf(vn), fabsf(vn + c)};
    const float eL[3] = {fabsf(vnL - L.c), fabsf(vnL), fabsf(vnL + L.c)};
    const float eR[3] = {fabsf(vnR - R.c), fabsf(vnR), fabsf(vnR + R.c)};
    for (int e = 0; e < 3; ++e) {
        const float eps = fmaxf(fabsf(eig[e] - eL[e]), fabsf(eR[e] - eig[e]));
        if (eig[e] < eps) eig[e] = 0.5f * (eig[e] * eig[e] * frcp(eps) + eps);
    }
    const float dp = R.p - L.p, drho = R.u[0] - L.u[0];
    const float dVn = vnR - vnL;
    const float hic2 = 0.5f * frcp(c2);
    const float rc = rho_ravg * c * dVn;
    const float co0 = eig[0] * (dp - rc) * hic2;
    const float co1 = eig[1] * (drho - 2.f * dp * hic2);
    const float co2 = eig[1] * rho_ravg;
    const float co3 = eig[2] * (dp + rc) * hic2;
    // tangential velocity jump: the normal component cancels exactly
    float dvt_dot = 0.f;
    float A[D + 2];
    A[0] = co0 + co1 + co3;
    for (int k = 0; k < D; ++k) {
        const float dvt = (k == K) ? 0.f : (R.vel[k] - L.vel[k]);
        dvt_dot += v_ravg[k] * dvt;
        const float nk = (k == K) ? sgn : 0.f;
        A[1 + k] = co0 * (v_ravg[k] - c * nk) + co1 * v_ravg[k] + co2 * dvt + co3 * (v_ravg[k] + c * nk);
    }
    A[D + 1] = co0 * (h_ravg - c * vn) + co1 * vel2 * 0.5f + co2 * dvt_dot + co3 * (h_ravg + c * vn);
    for (int s = 0; s < D + 2; ++s) out[s] = -0.5f * A[s];
}

// Numerical flux of a face node along the normal sgn * e_K, L the interior
// (owner) side.
template <int D, int K>
__device__ __forceinline__ void num_flux_fast(const NodeF<D> &L, const NodeF<D> &R, float sgn, const FastConst &fc,
                                              const PhysConst &pc, float *out)
{
    const int cf = fc.conv_flux;
    float base[D + 2];
    if (cf >= 3) {
        pair_flux_fast<D, K>(L, R, fc, pc, base);
        for (int s = 0; s < D + 2; ++s) base[s] *= sgn;
    } else {
        float f1[D + 2], f2[D + 2];
        convective_flux_dir<float, D>(L.u, K, fc.gamm1, f1);
        convective_flux_dir<float, D>(R.u, K, fc.gamm1, f2);
        for (int s = 0; s < D + 2; ++s) base[s] = (0.5f * (f1[s] + f2[s])) * sgn;
    }
    float diss[D + 2];
    for (int s = 0; s < D + 2; ++s) diss[s] = 0.f;
    if (cf == 0 || cf == 4) {
        const float lam = fmaxf(fabsf(L.vel[K]) + L.c, fabsf(R.vel[K]) + R.c);
        for (int s = 0; s < D + 2; ++s) diss[s] = -0.5f * lam * (R.u[s] - L.u[s]);
    }
    if (cf == 1 || cf == 5) roe_fast<D, K>(L, R, sgn, fc, diss);
    for (int s = 0; s < D + 2; ++s) out[s] = base[s] + diss[s];
}

} // namespace dgx
