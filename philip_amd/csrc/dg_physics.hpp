// Device physics for the strong-form DG path: Euler fluxes, entropy variables,
// two-point split fluxes, surface numerical fluxes, Navier-Stokes viscous flux.
// T is the arithmetic type of the flux evaluation (double, or float in the
// mixed-precision kernels).  Formulas follow the reference:
//   src/physics/euler.cpp:177-937,1033-1062 ; src/physics/navier_stokes.cpp:240-385,700-830,1197-1350
//   src/numerical_flux/convective_numerical_flux.cpp:34-480
//   src/numerical_flux/viscous_numerical_flux.cpp:144-328
#pragma once
#ifndef DGX_HOST_TEST
#include <hip/hip_runtime.h>
#endif

namespace dgx {

struct PhysConst {
    double gam, gamm1, mach2, reynolds, prandtl, mu_const, temperature_ratio;
    double vinf[3];  // free-stream velocity (cos a, sin a, 0): farfield, Riemann and inflow states
    double wall_temperature;  // Navier-Stokes isothermal no-slip walls
    int thermal_bc;  // Navier-Stokes no-slip walls (1001): 0 adiabatic, 1 isothermal
    int use_constant_viscosity;
    int two_point;   // KG=0 IR=1 CH=2 RA=3
    int conv_flux;   // LXF=0 ROE=1 CENTRAL=2 TWO_POINT=3 TP+LXF=4 TP+ROE=5
    int diss_flux;   // SIPG=0 BR2=1
};

template <typename T> __device__ __forceinline__ T dsqrt(T x) { return sqrt(x); }
template <typename T> __device__ __forceinline__ T dlog(T x) { return log(x); }
template <typename T> __device__ __forceinline__ T dpow(T x, T y) { return pow(x, y); }
template <typename T> __device__ __forceinline__ T dabs(T x) { return fabs(x); }
template <> __device__ __forceinline__ float dpow<float>(float x, float y) { return powf(x, y); }
template <> __device__ __forceinline__ float dabs<float>(float x) { return fabsf(x); }
// Division and square root: IEEE in double (the FP64 path reproduces the
// reference's operations); hardware reciprocal / square root (about 1 ulp) in
// the mixed path's float arithmetic.
template <typename T> __device__ __forceinline__ T fdiv(T a, T b) { return a / b; }
#ifdef DGX_HOST_TEST
template <> inline float dsqrt<float>(float x) { return sqrtf(x); }
#else
template <> __device__ __forceinline__ float dlog<float>(float x) { return __logf(x); }
template <> __device__ __forceinline__ float fdiv<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <> __device__ __forceinline__ float dsqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
#endif
// An integer the optimizer may not reason about past this point (an empty asm that takes and
// returns it in a register): used where gfx950 -O3 mis-folded a chain of comparisons.
#ifdef DGX_HOST_TEST
inline void opaque_int(int &) {}
#else
__device__ __forceinline__ void opaque_int(int &x) { asm volatile("" : "+v"(x)); }
#endif

// compute_pressure_templated (euler.cpp:486-500)
template <typename T, int D>
__device__ __forceinline__ T pressure(const T *u, T gamm1)
{
    const T rho = u[0];
    T vel2 = T(0);
    for (int k = 0; k < D; ++k) { const T v = fdiv<T>(u[1 + k], rho); vel2 = vel2 + v * v; }
    return gamm1 * (u[D + 1] - T(0.5) * rho * vel2);
}

// convective_flux (euler.cpp:915-937), component along direction fd only
template <typename T, int D>
__device__ __forceinline__ void convective_flux_dir(const T *u, int fd, T gamm1, T *f)
{
    const T rho = u[0];
    const T p = pressure<T, D>(u, gamm1);
    T vel[D];
    for (int k = 0; k < D; ++k) vel[k] = fdiv<T>(u[1 + k], rho);
    const T e = fdiv<T>(u[D + 1], u[0]);
    const T H = e + fdiv<T>(p, rho);
    f[0] = u[1 + fd];
    for (int vd = 0; vd < D; ++vd) f[1 + vd] = rho * vel[fd] * vel[vd];
    f[1 + fd] += p;
    f[D + 1] = rho * vel[fd] * H;
}

// compute_entropy_variables (euler.cpp:810-829)
template <typename T, int D>
__device__ __forceinline__ void entropy_variables(const T *u, const PhysConst &pc, T *v)
{
    const T gam = T(pc.gam), gamm1 = T(pc.gamm1);
    const T rho = u[0];
    const T p = pressure<T, D>(u, gamm1);
    const T s = dlog<T>(p * dpow<T>(rho, -gam));
    const T rho_theta = p / gamm1;
    v[0] = (rho_theta * (gam + T(1) - s) - u[D + 1]) / rho_theta;
    for (int k = 0; k < D; ++k) v[k + 1] = u[k + 1] / rho_theta;
    v[D + 1] = -rho / rho_theta;
}

// compute_ismail_roe_logarithmic_mean (euler.cpp:629-647)
template <typename T>
__device__ __forceinline__ T log_mean(T a, T b)
{
    const T zeta = fdiv<T>(a, b);
    const T f = fdiv<T>(zeta - T(1), zeta + T(1));
    const T uu = f * f;
    T F;
    if (uu < T(1.0e-2)) F = T(1) + fdiv<T>(uu, T(3)) + fdiv<T>(uu * uu, T(5)) + fdiv<T>(uu * uu * uu, T(7));
    else F = fdiv<T>(fdiv<T>(dlog<T>(zeta), T(2)), f);
    return fdiv<T>(a + b, T(2) * F);
}

// Primitive cache used by the pair loops: rho, velocity, pressure.
template <typename T, int D> struct Prim { T rho, vel[D], p; };

// Euler::boundary_face_values (euler.cpp:1511-1567) on a Cartesian face whose outward normal is
// sg e_K: the exterior state ub of a boundary face node from the interior state ui.  Types:
// 1001 / 1006 slip wall (normal velocity reflected, :1171-1228), 1002 pressure outflow at 0.99 of
// the inlet static pressure (:1290-1325), 1003 inflow at the free-stream total pressure and
// temperature (:1327-1431), 1004 Riemann-invariant farfield (:1097-1169), 1005 free stream
// (:1433-1447), 1007 extrapolation (:1449-1460).  Free stream: rho = 1, p = 1/(gamma M^2), T = 1.
// Every branch sets the exterior primitive state (rho_b, vel_b, p_b) as scalars chosen by value,
// with the face direction picked by (k == K) in unrolled loops: a form that selected between two
// private arrays through a pointer and stored through a computed index (wb[1 + K]) compiled at
// -O3 for gfx950 to a slip-wall state carrying the free-stream velocity component instead of the
// reflected one (right at -O1 and on the host).
template <typename T, int D>
__device__ __forceinline__ void boundary_state(int type, const T *ui, int K, T sg, const PhysConst &pc, T *ub)
{
    constexpr int NS = D + 2;
    const T gam = T(pc.gam), gamm1 = T(pc.gamm1), m2 = T(pc.mach2);
    const T pinf = T(1) / (gam * m2);
    const T rho_i = ui[0];
    T vel_i[D], v2 = T(0), vn_i = T(0);
#pragma unroll
    for (int k = 0; k < D; ++k) {
        vel_i[k] = ui[1 + k] / rho_i;
        v2 += vel_i[k] * vel_i[k];
        if (k == K) vn_i = vel_i[k] * sg;  // velocity along the outward normal
    }
    const T p_i = gamm1 * (ui[D + 1] - T(0.5) * rho_i * v2);
    T rho_b = rho_i, p_b = p_i, vel_b[D];
#pragma unroll
    for (int k = 0; k < D; ++k) vel_b[k] = vel_i[k];
    bool copy = false;  // exterior = interior conservative state
    // 1007 is tested first, by equality, with the type opaque to the optimizer: inlined into the
    // curvilinear stage kernel, gfx950 -O3 turned this chain into range tests that sent 1007 down
    // the slip-wall branch (its code held no comparison against 1007)
    opaque_int(type);
    if (type == 1007) {
        copy = true;
    } else if (type == 1001 || type == 1006) {
#pragma unroll
        for (int k = 0; k < D; ++k)
            if (k == K) vel_b[k] = -vel_i[k];  // v - 2 (v.n) n with n = +-e_K
    } else if (type == 1002) {
        const T mach = dsqrt<T>(v2) / dsqrt<T>(gam * p_i / rho_i);
        if (mach > T(1)) {
            copy = true;
        } else {
            // the inlet static pressure from the total inlet pressure (the free stream's)
            const T rad = T(1) + T(0.5) * gamm1 * m2;
            const T p_inlet = (pinf * dpow<T>(rad, gam / gamm1)) * dpow<T>(rad, -gam / gamm1);
            p_b = (mach >= T(1)) ? p_i : T(0.99) * p_inlet;
            const T Ti = gam * m2 * (p_i / rho_i);
            rho_b = gam * m2 * (p_b / Ti);
        }
    } else if (type == 1003) {
        const T p0 = pinf * dpow<T>(T(1) + T(0.5) * gamm1 * m2, gam / gamm1);
        const T T0 = dpow<T>(p0 / pinf, gamm1 / gam);
        T speed;  // along the inward normal
        if (m2 < T(1)) {
            const T snd = dsqrt<T>(gam * p_i / rho_i);
            const T rp = -vn_i + T(2) * snd / gamm1;
            const T H = ui[D + 1] / rho_i + p_i / rho_i;
            const T A = T(1) + T(2) / gamm1, B = T(-2) * rp, C = T(0.5) * gamm1 * (rp * rp - T(2) * H);
            const T disc = dsqrt<T>(B * B - T(4) * A * C);
            const T c1 = (-B + disc) / (T(2) * A), c2 = (-B - disc) / (T(2) * A);
            const T cb = c1 > c2 ? c1 : c2;
            const T vb = rp - T(2) * cb / gamm1;
            const T mb = vb / cb;
            const T rad = T(1) + T(0.5) * gamm1 * mb * mb;
            p_b = p0 * dpow<T>(rad, -gam / gamm1);
            rho_b = gam * m2 * (p_b / (T0 / rad));
            speed = vb;
        } else {
            const T rad = T(1) + T(0.5) * gamm1 * m2;
            p_b = p0 * dpow<T>(rad, -gam / gamm1);
            rho_b = gam * m2 * (p_b / (T0 / rad));
            speed = dsqrt<T>(m2) * dsqrt<T>(gam * p_b / rho_b);
        }
#pragma unroll
        for (int k = 0; k < D; ++k) vel_b[k] = (k == K) ? -speed * sg : T(0);
    } else if (type == 1004) {
        T vn_e = T(0);
#pragma unroll
        for (int k = 0; k < D; ++k)
            if (k == K) vn_e = T(pc.vinf[k]) * sg;
        const T ci = dsqrt<T>(gam * p_i / rho_i), ce = dsqrt<T>(gam * pinf);
        const T Rout = vn_i + T(2) / gamm1 * ci, Rinc = vn_e - T(2) / gamm1 * ce;
        const T vnb = T(0.5) * (Rout + Rinc), cb = T(0.25) * gamm1 * (Rout - Rinc);
        const bool inlet = vnb < T(0);
        // the upwind side's state, by value
        const T rho_s = inlet ? T(1) : rho_i, p_s = inlet ? pinf : p_i, vn_s = inlet ? vn_e : vn_i;
        T vel_s[D];
#pragma unroll
        for (int k = 0; k < D; ++k) vel_s[k] = inlet ? T(pc.vinf[k]) : vel_i[k];
        if (dabs<T>(vnb) >= dabs<T>(cb)) {
            rho_b = rho_s;
            p_b = p_s;
#pragma unroll
            for (int k = 0; k < D; ++k) vel_b[k] = vel_s[k];
        } else {
            const T ent = p_s * dpow<T>(rho_s, -gam);
            rho_b = dpow<T>(T(1) / gam * cb * cb / ent, T(1) / gamm1);
#pragma unroll
            for (int k = 0; k < D; ++k) vel_b[k] = (k == K) ? (vel_s[k] - vn_s * sg) + vnb * sg : vel_s[k];
            p_b = T(1) / gam * cb * cb * rho_b;
        }
    } else if (type == 1005) {
        rho_b = T(1);
        p_b = pinf;
#pragma unroll
        for (int k = 0; k < D; ++k) vel_b[k] = T(pc.vinf[k]);
    } else {
        copy = true;
    }
    T w2 = T(0);
#pragma unroll
    for (int k = 0; k < D; ++k) w2 += vel_b[k] * vel_b[k];
    ub[0] = copy ? ui[0] : rho_b;
#pragma unroll
    for (int k = 0; k < D; ++k) ub[1 + k] = copy ? ui[1 + k] : rho_b * vel_b[k];
    ub[D + 1] = copy ? ui[D + 1] : p_b / gamm1 + T(0.5) * rho_b * w2;
    (void)NS;
}

// NavierStokes::boundary_face_values_viscous_flux for the wall types (navier_stokes.cpp:1354-1430): the
// exterior state of the viscous flux and of the lifted gradient.  1001 no-slip wall: momentum
// reversed, and on an isothermal wall the energy set so that the average temperature is the wall
// temperature; 1006 slip wall: the interior state.  (The exterior gradient is the interior one,
// with the energy row zeroed on an adiabatic no-slip wall.)
template <typename T, int D>
__device__ __forceinline__ void visc_wall_state(int type, const T *ui, const PhysConst &pc, T *ub)
{
    for (int s = 0; s < D + 2; ++s) ub[s] = ui[s];
    if (type != 1001) return;
#pragma unroll
    for (int k = 0; k < D; ++k) ub[1 + k] = -ui[1 + k];
    if (pc.thermal_bc == 1) {
        const T gam = T(pc.gam), gamm1 = T(pc.gamm1), m2 = T(pc.mach2);
        const T rho = ui[0];
        T v2 = T(0);
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const T v = ub[1 + k] / rho;
            v2 += v * v;
        }
        const T p_i = gamm1 * (ui[D + 1] - T(0.5) * rho * v2);
        const T T_int = gam * m2 * (p_i / rho);
        const T T_ext = T(2) * T(pc.wall_temperature) - T_int;
        const T p_e = rho * T_ext / (gam * m2);
        ub[D + 1] = p_e / gamm1 + T(0.5) * rho * v2;
    }
}

// The same boundary states on a face of a curvilinear grid, n the outward unit normal (the
// reference's general form, euler.cpp:1097-1460); written like boundary_state with scalar,
// by-value selections.
template <typename T, int D>
__device__ __forceinline__ void boundary_state_n(int type, const T *ui, const T *n, const PhysConst &pc, T *ub)
{
    const T gam = T(pc.gam), gamm1 = T(pc.gamm1), m2 = T(pc.mach2);
    const T pinf = T(1) / (gam * m2);
    const T rho_i = ui[0];
    T vel_i[D], v2 = T(0), vn_i = T(0), vn_e = T(0);
#pragma unroll
    for (int k = 0; k < D; ++k) {
        vel_i[k] = ui[1 + k] / rho_i;
        v2 += vel_i[k] * vel_i[k];
        vn_i += vel_i[k] * n[k];
        vn_e += T(pc.vinf[k]) * n[k];
    }
    const T p_i = gamm1 * (ui[D + 1] - T(0.5) * rho_i * v2);
    T rho_b = rho_i, p_b = p_i, vel_b[D];
#pragma unroll
    for (int k = 0; k < D; ++k) vel_b[k] = vel_i[k];
    bool copy = false;
    // 1007 first, by equality, the type opaque to the optimizer (see boundary_state)
    opaque_int(type);
    if (type == 1007) {
        copy = true;
    } else if (type == 1001 || type == 1006) {
        // v - 2 (v.n) n
#pragma unroll
        for (int k = 0; k < D; ++k) vel_b[k] = vel_i[k] - T(2) * vn_i * n[k];
    } else if (type == 1002) {
        const T mach = dsqrt<T>(v2) / dsqrt<T>(gam * p_i / rho_i);
        if (mach > T(1)) {
            copy = true;
        } else {
            const T rad = T(1) + T(0.5) * gamm1 * m2;
            const T p_inlet = (pinf * dpow<T>(rad, gam / gamm1)) * dpow<T>(rad, -gam / gamm1);
            p_b = (mach >= T(1)) ? p_i : T(0.99) * p_inlet;
            const T Ti = gam * m2 * (p_i / rho_i);
            rho_b = gam * m2 * (p_b / Ti);
        }
    } else if (type == 1003) {
        const T p0 = pinf * dpow<T>(T(1) + T(0.5) * gamm1 * m2, gam / gamm1);
        const T T0 = dpow<T>(p0 / pinf, gamm1 / gam);
        T speed;
        if (m2 < T(1)) {
            const T snd = dsqrt<T>(gam * p_i / rho_i);
            const T rp = -vn_i + T(2) * snd / gamm1;
            const T H = ui[D + 1] / rho_i + p_i / rho_i;
            const T A = T(1) + T(2) / gamm1, B = T(-2) * rp, C = T(0.5) * gamm1 * (rp * rp - T(2) * H);
            const T disc = dsqrt<T>(B * B - T(4) * A * C);
            const T c1 = (-B + disc) / (T(2) * A), c2 = (-B - disc) / (T(2) * A);
            const T cb = c1 > c2 ? c1 : c2;
            const T vb = rp - T(2) * cb / gamm1;
            const T mb = vb / cb;
            const T rad = T(1) + T(0.5) * gamm1 * mb * mb;
            p_b = p0 * dpow<T>(rad, -gam / gamm1);
            rho_b = gam * m2 * (p_b / (T0 / rad));
            speed = vb;
        } else {
            const T rad = T(1) + T(0.5) * gamm1 * m2;
            p_b = p0 * dpow<T>(rad, -gam / gamm1);
            rho_b = gam * m2 * (p_b / (T0 / rad));
            speed = dsqrt<T>(m2) * dsqrt<T>(gam * p_b / rho_b);
        }
#pragma unroll
        for (int k = 0; k < D; ++k) vel_b[k] = speed * (-n[k]);
    } else if (type == 1004) {
        const T ci = dsqrt<T>(gam * p_i / rho_i), ce = dsqrt<T>(gam * pinf);
        const T Rout = vn_i + T(2) / gamm1 * ci, Rinc = vn_e - T(2) / gamm1 * ce;
        const T vnb = T(0.5) * (Rout + Rinc), cb = T(0.25) * gamm1 * (Rout - Rinc);
        const bool inlet = vnb < T(0);
        const T rho_s = inlet ? T(1) : rho_i, p_s = inlet ? pinf : p_i, vn_s = inlet ? vn_e : vn_i;
        T vel_s[D];
#pragma unroll
        for (int k = 0; k < D; ++k) vel_s[k] = inlet ? T(pc.vinf[k]) : vel_i[k];
        if (dabs<T>(vnb) >= dabs<T>(cb)) {
            rho_b = rho_s;
            p_b = p_s;
#pragma unroll
            for (int k = 0; k < D; ++k) vel_b[k] = vel_s[k];
        } else {
            const T ent = p_s * dpow<T>(rho_s, -gam);
            rho_b = dpow<T>(T(1) / gam * cb * cb / ent, T(1) / gamm1);
#pragma unroll
            for (int k = 0; k < D; ++k) vel_b[k] = (vel_s[k] - vn_s * n[k]) + vnb * n[k];
            p_b = T(1) / gam * cb * cb * rho_b;
        }
    } else if (type == 1005) {
        rho_b = T(1);
        p_b = pinf;
#pragma unroll
        for (int k = 0; k < D; ++k) vel_b[k] = T(pc.vinf[k]);
    } else {
        copy = true;
    }
    T w2 = T(0);
#pragma unroll
    for (int k = 0; k < D; ++k) w2 += vel_b[k] * vel_b[k];
    ub[0] = copy ? ui[0] : rho_b;
#pragma unroll
    for (int k = 0; k < D; ++k) ub[1 + k] = copy ? ui[1 + k] : rho_b * vel_b[k];
    ub[D + 1] = copy ? ui[D + 1] : p_b / gamm1 + T(0.5) * rho_b * w2;
}

template <typename T, int D>
__device__ __forceinline__ Prim<T, D> to_prim(const T *u, T gamm1)
{
    Prim<T, D> q;
    q.rho = u[0];
    for (int k = 0; k < D; ++k) q.vel[k] = fdiv<T>(u[1 + k], u[0]);
    q.p = pressure<T, D>(u, gamm1);
    return q;
}

// Two-point split fluxes, direction fd component (euler.cpp:587-807).
// u1/u2 are the conservative states, q1/q2 their primitive caches.
template <typename T, int D>
__device__ __forceinline__ void two_point_flux_dir(const T *u1, const Prim<T, D> &q1, const T *u2,
                                                   const Prim<T, D> &q2, int fd, const PhysConst &pc, T *f)
{
    const T gam = T(pc.gam), gamm1 = T(pc.gamm1);
    if (pc.two_point == 2 || pc.two_point == 3) {
        const bool ranocha = (pc.two_point == 3);
        const T rho_log = log_mean<T>(q1.rho, q2.rho);
        const T b1 = ranocha ? fdiv<T>(q1.rho, q1.p) : fdiv<T>(q1.rho, T(2) * q1.p);
        const T b2 = ranocha ? fdiv<T>(q2.rho, q2.p) : fdiv<T>(q2.rho, T(2) * q2.p);
        const T beta_log = log_mean<T>(b1, b2);
        const T p_hat = ranocha ? T(0.5) * (q1.p + q2.p)
                                : fdiv<T>(T(0.5) * (q1.rho + q2.rho), T(2) * T(0.5) * (b1 + b2));
        T va[D];
        T vsq = T(0);
        for (int k = 0; k < D; ++k) {
            va[k] = T(0.5) * (q1.vel[k] + q2.vel[k]);
            vsq += (T(0.5) * (q1.vel[k] + q2.vel[k])) * (T(0.5) * (q1.vel[k] + q2.vel[k]));
        }
        T H = ranocha ? fdiv<T>(T(1), beta_log * gamm1) + vsq + fdiv<T>(T(2) * p_hat, rho_log)
                      : fdiv<T>(T(1), T(2) * beta_log * gamm1) + vsq + fdiv<T>(p_hat, rho_log);
        for (int k = 0; k < D; ++k) H -= T(0.5) * (T(0.5) * (q1.vel[k] * q1.vel[k] + q2.vel[k] * q2.vel[k]));
        f[0] = rho_log * va[fd];
        for (int vd = 0; vd < D; ++vd) f[1 + vd] = rho_log * va[fd] * va[vd];
        f[1 + fd] += p_hat;
        f[D + 1] = rho_log * va[fd] * H;
        if (ranocha) f[D + 1] -= (T(0.5) * (q1.p * q1.vel[fd] + q2.p * q2.vel[fd]));
    } else if (pc.two_point == 1) {
        // Ismail-Roe (euler.cpp:614-707)
        T z1[D + 2], z2[D + 2];
        z1[0] = dsqrt<T>(fdiv<T>(q1.rho, q1.p));
        z2[0] = dsqrt<T>(fdiv<T>(q2.rho, q2.p));
        for (int k = 0; k < D; ++k) { z1[1 + k] = z1[0] * q1.vel[k]; z2[1 + k] = z2[0] * q2.vel[k]; }
        z1[D + 1] = dsqrt<T>(q1.rho * q1.p);
        z2[D + 1] = dsqrt<T>(q2.rho * q2.p);
        T avg[D + 2];
        for (int s = 0; s < D + 2; ++s) avg[s] = T(0.5) * (z1[s] + z2[s]);
        const T lm0 = log_mean<T>(z1[0], z2[0]);
        const T lmE = log_mean<T>(z1[D + 1], z2[D + 1]);
        const T mrho = avg[0] * lmE;
        T mv[D];
        for (int k = 0; k < D; ++k) mv[k] = fdiv<T>(avg[1 + k], avg[0]);
        const T mp = fdiv<T>(avg[D + 1], avg[0]);
        T mh = (gam + T(1)) * fdiv<T>(lmE, lm0) + gamm1 * mp;
        mh = fdiv<T>(mh, T(2) * gam);
        mh *= fdiv<T>(gam, mrho * gamm1);
        T vsq = T(0);
        for (int k = 0; k < D; ++k) vsq += mv[k] * mv[k];
        mh += T(0.5) * vsq;
        f[0] = mrho * mv[fd];
        for (int vd = 0; vd < D; ++vd) f[1 + vd] = mrho * mv[fd] * mv[vd];
        f[1 + fd] += mp;
        f[D + 1] = mrho * mv[fd] * mh;
    } else {
        // Kennedy-Gruber (euler.cpp:587-611)
        const T mrho = fdiv<T>(q1.rho + q2.rho, T(2));
        const T mp = fdiv<T>(q1.p + q2.p, T(2));
        T mv[D];
        for (int k = 0; k < D; ++k) mv[k] = T(0.5) * (q1.vel[k] + q2.vel[k]);
        const T me = fdiv<T>(fdiv<T>(u1[D + 1], u1[0]) + fdiv<T>(u2[D + 1], u2[0]), T(2));
        f[0] = mrho * mv[fd];
        for (int vd = 0; vd < D; ++vd) f[1 + vd] = mrho * mv[fd] * mv[vd];
        f[1 + fd] += mp;
        f[D + 1] = mrho * mv[fd] * me + mp * mv[fd];
    }
}

// Roe-Pike dissipation with Harten's entropy fix along the unit normal n
// (convective_numerical_flux.cpp:216-232, :332-480)
template <typename T, int D>
__device__ __forceinline__ void roe_dissipation_n(const T *ui, const Prim<T, D> &L, const T *ue,
                                                  const Prim<T, D> &R, const T *n, const PhysConst &pc, T *out)
{
    const T gam = T(pc.gam), gamm1 = T(pc.gamm1);
    T vnL = T(0), vnR = T(0);
    for (int k = 0; k < D; ++k) vnL += L.vel[k] * n[k];
    for (int k = 0; k < D; ++k) vnR += R.vel[k] * n[k];
    const T hL = fdiv<T>(ui[D + 1] + L.p, L.rho);
    const T hR = fdiv<T>(ue[D + 1] + R.p, R.rho);
    const T r = dsqrt<T>(fdiv<T>(R.rho, L.rho));
    const T rp1 = r + T(1);
    const T rho_ravg = r * L.rho;
    T v_ravg[D];
    for (int k = 0; k < D; ++k) v_ravg[k] = fdiv<T>(r * R.vel[k] + L.vel[k], rp1);
    const T h_ravg = fdiv<T>(r * hR + hL, rp1);
    T vel2_ravg = T(0);
    for (int k = 0; k < D; ++k) vel2_ravg = vel2_ravg + v_ravg[k] * v_ravg[k];
    T vn_ravg = T(0);
    for (int k = 0; k < D; ++k) vn_ravg += v_ravg[k] * n[k];
    const T c2 = gamm1 * (h_ravg - T(0.5) * vel2_ravg);
    T c_ravg = T(1e10);
    if (c2 > T(0)) c_ravg = dsqrt<T>(c2);
    T eig[3] = {dabs<T>(vn_ravg - c_ravg), dabs<T>(vn_ravg), dabs<T>(vn_ravg + c_ravg)};
    const T cL = dsqrt<T>(fdiv<T>(L.p * gam, L.rho)), cR = dsqrt<T>(fdiv<T>(R.p * gam, R.rho));
    const T eL[3] = {dabs<T>(vnL - cL), dabs<T>(vnL), dabs<T>(vnL + cL)};
    const T eR[3] = {dabs<T>(vnR - cR), dabs<T>(vnR), dabs<T>(vnR + cR)};
    const T dp = R.p - L.p, drho = R.rho - L.rho;
    const T dVn = vnR - vnL;
    T dVt[D];
    for (int k = 0; k < D; ++k) dVt[k] = (R.vel[k] - L.vel[k]) - dVn * n[k];
    for (int e = 0; e < 3; ++e) {
        const T a1 = dabs<T>(eig[e] - eL[e]), a2 = dabs<T>(eR[e] - eig[e]);
        const T eps = a1 > a2 ? a1 : a2;
        if (eig[e] < eps) eig[e] = T(0.5) * (fdiv<T>(eig[e] * eig[e], eps) + eps);
    }
    T co0 = fdiv<T>(eig[0] * (dp - rho_ravg * c_ravg * dVn), T(2) * c2);
    T co1 = eig[1] * (drho - fdiv<T>(dp, c2));
    T co2 = eig[1] * rho_ravg;
    T co3 = fdiv<T>(eig[2] * (dp + rho_ravg * c_ravg * dVn), T(2) * c2);
    T A[D + 2];
    A[0] = co0;
    for (int k = 0; k < D; ++k) A[1 + k] = co0 * (v_ravg[k] - c_ravg * n[k]);
    A[D + 1] = co0 * (h_ravg - c_ravg * vn_ravg);
    A[0] += co1;
    for (int k = 0; k < D; ++k) A[1 + k] += co1 * v_ravg[k];
    A[D + 1] += co1 * vel2_ravg * T(0.5);
    T dvt_dot = T(0);
    for (int k = 0; k < D; ++k) { A[1 + k] += co2 * dVt[k]; dvt_dot += v_ravg[k] * dVt[k]; }
    A[D + 1] += co2 * dvt_dot;
    A[0] += co3;
    for (int k = 0; k < D; ++k) A[1 + k] += co3 * (v_ravg[k] + c_ravg * n[k]);
    A[D + 1] += co3 * (h_ravg + c_ravg * vn_ravg);
    for (int s = 0; s < D + 2; ++s) out[s] = T(-0.5) * A[s];
}

// the same along the unit normal sgn * e_fd
template <typename T, int D>
__device__ __forceinline__ void roe_dissipation(const T *ui, const Prim<T, D> &L, const T *ue,
                                                const Prim<T, D> &R, int fd, T sgn, const PhysConst &pc,
                                                T *out)
{
    T n[D];
    for (int k = 0; k < D; ++k) n[k] = (k == fd) ? sgn : T(0);
    roe_dissipation_n<T, D>(ui, L, ue, R, n, pc, out);
}

// NumericalFluxConvective::evaluate_flux = baseline + dissipation (convective_numerical_flux.cpp:34-54)
// with the unit normal sgn * e_fd.
template <typename T, int D>
__device__ __forceinline__ void numerical_flux(const T *ui, const Prim<T, D> &L, const T *ue,
                                               const Prim<T, D> &R, int fd, T sgn, const PhysConst &pc,
                                               T *out)
{
    const T gamm1 = T(pc.gamm1);
    const int cf = pc.conv_flux;
    T base[D + 2];
    if (cf >= 3) {
        T f[D + 2];
        two_point_flux_dir<T, D>(ui, L, ue, R, fd, pc, f);
        for (int s = 0; s < D + 2; ++s) base[s] = f[s] * sgn;
    } else {
        T f1[D + 2], f2[D + 2];
        convective_flux_dir<T, D>(ui, fd, gamm1, f1);
        convective_flux_dir<T, D>(ue, fd, gamm1, f2);
        for (int s = 0; s < D + 2; ++s) base[s] = (T(0.5) * (f1[s] + f2[s])) * sgn;
    }
    // baseline + Riemann-solver dissipation; the central / entropy-conserving fluxes add a zero
    // dissipation vector, as ZeroRiemannSolverDissipation does in the reference
    T diss[D + 2];
    for (int s = 0; s < D + 2; ++s) diss[s] = T(0);
    if (cf == 0 || cf == 4) {
        const T gam = T(pc.gam);
        const T ei = dabs<T>(L.vel[fd] * sgn) + dsqrt<T>(fdiv<T>(L.p * gam, L.rho));
        const T ee = dabs<T>(R.vel[fd] * sgn) + dsqrt<T>(fdiv<T>(R.p * gam, R.rho));
        const T lam = (ei > ee) ? ei : ee;
        for (int s = 0; s < D + 2; ++s) diss[s] = T(-0.5) * lam * (ue[s] - ui[s]);
    }
    if (cf == 1 || cf == 5) roe_dissipation<T, D>(ui, L, ue, R, fd, sgn, pc, diss);
    for (int s = 0; s < D + 2; ++s) out[s] = base[s] + diss[s];
}

// Curvilinear grids: the two-point flux contracted with a direction vector (a metric-cofactor
// column or its two-point average, not necessarily unit): sum_m nv[m] F#_m (the contraction of
// transform_physical_to_reference, operators.cpp:2354-2365).
template <typename T, int D>
__device__ __forceinline__ void two_point_flux_vec(const T *u1, const Prim<T, D> &q1, const T *u2,
                                                   const Prim<T, D> &q2, const T *nv, const PhysConst &pc, T *f)
{
    if (pc.two_point == 2 || pc.two_point == 3) {
        // Chandrashekar / Ranocha: rho_ln (vbar.n) (1, vbar, H) + p_hat (0, n, 0) [- Ranocha's
        // pressure-work term], the components of two_point_flux_dir contracted in one evaluation
        const T gamm1 = T(pc.gamm1);
        const bool ranocha = (pc.two_point == 3);
        const T rho_log = log_mean<T>(q1.rho, q2.rho);
        const T b1 = ranocha ? fdiv<T>(q1.rho, q1.p) : fdiv<T>(q1.rho, T(2) * q1.p);
        const T b2 = ranocha ? fdiv<T>(q2.rho, q2.p) : fdiv<T>(q2.rho, T(2) * q2.p);
        const T beta_log = log_mean<T>(b1, b2);
        const T p_hat = ranocha ? T(0.5) * (q1.p + q2.p)
                                : fdiv<T>(T(0.5) * (q1.rho + q2.rho), T(2) * T(0.5) * (b1 + b2));
        T va[D], vsq = T(0), vn = T(0), w1 = T(0), w2 = T(0);
#pragma unroll
        for (int k = 0; k < D; ++k) {
            va[k] = T(0.5) * (q1.vel[k] + q2.vel[k]);
            vsq += va[k] * va[k];
            vn += va[k] * nv[k];
            w1 += q1.vel[k] * nv[k];
            w2 += q2.vel[k] * nv[k];
        }
        T H = ranocha ? fdiv<T>(T(1), beta_log * gamm1) + vsq + fdiv<T>(T(2) * p_hat, rho_log)
                      : fdiv<T>(T(1), T(2) * beta_log * gamm1) + vsq + fdiv<T>(p_hat, rho_log);
        for (int k = 0; k < D; ++k) H -= T(0.5) * (T(0.5) * (q1.vel[k] * q1.vel[k] + q2.vel[k] * q2.vel[k]));
        const T m = rho_log * vn;
        f[0] = m;
#pragma unroll
        for (int vd = 0; vd < D; ++vd) f[1 + vd] = m * va[vd] + p_hat * nv[vd];
        f[D + 1] = m * H;
        if (ranocha) f[D + 1] -= T(0.5) * (q1.p * w1 + q2.p * w2);
        return;
    }
    for (int s = 0; s < D + 2; ++s) f[s] = T(0);
#pragma unroll
    for (int m = 0; m < D; ++m) {
        T fm[D + 2];
        two_point_flux_dir<T, D>(u1, q1, u2, q2, m, pc, fm);
        for (int s = 0; s < D + 2; ++s) f[s] += fm[s] * nv[m];
    }
}

// sum_m nv[m] F_m of the convective flux
template <typename T, int D>
__device__ __forceinline__ void convective_flux_vec(const T *u, const T *nv, T gamm1, T *f)
{
    for (int s = 0; s < D + 2; ++s) f[s] = T(0);
#pragma unroll
    for (int m = 0; m < D; ++m) {
        T fm[D + 2];
        convective_flux_dir<T, D>(u, m, gamm1, fm);
        for (int s = 0; s < D + 2; ++s) f[s] += fm[s] * nv[m];
    }
}

// NumericalFluxConvective::evaluate_flux along a general unit normal n (curvilinear faces)
template <typename T, int D>
__device__ __forceinline__ void numerical_flux_n(const T *ui, const Prim<T, D> &L, const T *ue,
                                                 const Prim<T, D> &R, const T *n, const PhysConst &pc, T *out)
{
    const T gamm1 = T(pc.gamm1);
    const int cf = pc.conv_flux;
    T base[D + 2];
    if (cf >= 3) {
        two_point_flux_vec<T, D>(ui, L, ue, R, n, pc, base);
    } else {
        T f1[D + 2], f2[D + 2];
        for (int s = 0; s < D + 2; ++s) base[s] = T(0);
#pragma unroll
        for (int m = 0; m < D; ++m) {
            convective_flux_dir<T, D>(ui, m, gamm1, f1);
            convective_flux_dir<T, D>(ue, m, gamm1, f2);
            for (int s = 0; s < D + 2; ++s) base[s] += (T(0.5) * (f1[s] + f2[s])) * n[m];
        }
    }
    T diss[D + 2];
    for (int s = 0; s < D + 2; ++s) diss[s] = T(0);
    if (cf == 0 || cf == 4) {
        const T gam = T(pc.gam);
        T vnL = T(0), vnR = T(0);
        for (int k = 0; k < D; ++k) {
            vnL += L.vel[k] * n[k];
            vnR += R.vel[k] * n[k];
        }
        const T ei = dabs<T>(vnL) + dsqrt<T>(fdiv<T>(L.p * gam, L.rho));
        const T ee = dabs<T>(vnR) + dsqrt<T>(fdiv<T>(R.p * gam, R.rho));
        const T lam = (ei > ee) ? ei : ee;
        for (int s = 0; s < D + 2; ++s) diss[s] = T(-0.5) * lam * (ue[s] - ui[s]);
    }
    if (cf == 1 || cf == 5) roe_dissipation_n<T, D>(ui, L, ue, R, n, pc, diss);
    for (int s = 0; s < D + 2; ++s) out[s] = base[s] + diss[s];
}

// dissipative_flux_templated (navier_stokes.cpp:1197-1220), direction fd component.
// g is the conservative gradient g[s*D + k].
template <typename T, int D>
__device__ __forceinline__ void viscous_flux_dir(const T *u, const T *g, int fd, const PhysConst &pc, T *f,
                                                 bool adiabatic = false)
{
    const T gam = T(pc.gam), gamm1 = T(pc.gamm1), mach2 = T(pc.mach2);
    const T rho = u[0];
    T vel[D];
    for (int k = 0; k < D; ++k) vel[k] = u[1 + k] / rho;
    const T p = pressure<T, D>(u, gamm1);
    T pg[D + 2][D];
    for (int k = 0; k < D; ++k) pg[0][k] = g[0 * D + k];
    for (int a = 0; a < D; ++a)
        for (int b = 0; b < D; ++b) pg[1 + a][b] = (g[(1 + a) * D + b] - vel[a] * g[0 * D + b]) / rho;
    for (int a = 0; a < D; ++a) {
        pg[D + 1][a] = g[(D + 1) * D + a];
        for (int b = 0; b < D; ++b) pg[D + 1][a] -= T(0.5) * (vel[b] * g[(1 + b) * D + a] + u[1 + b] * pg[1 + b][a]);
        pg[D + 1][a] *= gamm1;
    }
    const T T_ = gam * mach2 * (p / rho);
    T mu;
    if (pc.use_constant_viscosity) mu = T(1) * T(pc.mu_const);
    else {
        const T tr = T(pc.temperature_ratio);
        mu = ((T(1) + tr) / (T_ + tr)) * dpow<T>(T_, T(1.5));
    }
    const T mu_hat = mu / T(pc.reynolds);
    T divv = T(0);
    for (int k = 0; k < D; ++k) divv += T(0.5) * (pg[1 + k][k] + pg[1 + k][k]);
    const T mu2 = (T(-2) / T(3)) * mu_hat;
    // tau[a][fd] and tau[fd][a] (symmetric)
    T tau_fd[D];
    for (int a = 0; a < D; ++a) {
        tau_fd[a] = T(2) * mu_hat * (T(0.5) * (pg[1 + a][fd] + pg[1 + fd][a]));
        if (a == fd) tau_fd[a] += mu2 * divv;
    }
    const T kappa = mu_hat / (gamm1 * mach2 * T(pc.prandtl));
    const T dT = (gam * mach2 * pg[D + 1][fd] - T_ * pg[0][fd]) / rho;
    // adiabatic: an adiabatic no-slip wall drops the heat flux
    // (dissipative_flux_dot_normal_on_adiabatic_boundary, navier_stokes.cpp:1266-1310)
    const T q = adiabatic ? T(0) : -kappa * dT;
    f[0] = T(0);
    for (int a = 0; a < D; ++a) f[1 + a] = -tau_fd[a];
    T fe = T(0);
    for (int a = 0; a < D; ++a) fe -= vel[a] * tau_fd[a];
    f[D + 1] = fe + q;
}

// Dissipative flux in every direction at once, f[k] = F_v(u, grad u) . e_k: the primitive
// gradient, viscosity and divergence are shared between the directions; each direction's
// arithmetic is that of viscous_flux_dir (adiabatic: without the heat flux, as there).
template <typename T, int D>
__device__ __forceinline__ void viscous_flux_all(const T *u, const T *g, const PhysConst &pc, T (&f)[D][D + 2],
                                                 bool adiabatic = false)
{
    const T gam = T(pc.gam), gamm1 = T(pc.gamm1), mach2 = T(pc.mach2);
    const T rho = u[0];
    T vel[D];
    for (int k = 0; k < D; ++k) vel[k] = u[1 + k] / rho;
    const T p = pressure<T, D>(u, gamm1);
    T pg[D + 2][D];
    for (int k = 0; k < D; ++k) pg[0][k] = g[0 * D + k];
    for (int a = 0; a < D; ++a)
        for (int b = 0; b < D; ++b) pg[1 + a][b] = (g[(1 + a) * D + b] - vel[a] * g[0 * D + b]) / rho;
    for (int a = 0; a < D; ++a) {
        pg[D + 1][a] = g[(D + 1) * D + a];
        for (int b = 0; b < D; ++b) pg[D + 1][a] -= T(0.5) * (vel[b] * g[(1 + b) * D + a] + u[1 + b] * pg[1 + b][a]);
        pg[D + 1][a] *= gamm1;
    }
    const T T_ = gam * mach2 * (p / rho);
    T mu;
    if (pc.use_constant_viscosity) mu = T(1) * T(pc.mu_const);
    else {
        const T tr = T(pc.temperature_ratio);
        mu = ((T(1) + tr) / (T_ + tr)) * dpow<T>(T_, T(1.5));
    }
    const T mu_hat = mu / T(pc.reynolds);
    T divv = T(0);
    for (int k = 0; k < D; ++k) divv += T(0.5) * (pg[1 + k][k] + pg[1 + k][k]);
    const T mu2 = (T(-2) / T(3)) * mu_hat;
    const T kappa = mu_hat / (gamm1 * mach2 * T(pc.prandtl));
#pragma unroll
    for (int fd = 0; fd < D; ++fd) {
        T tau_fd[D];
        for (int a = 0; a < D; ++a) {
            tau_fd[a] = T(2) * mu_hat * (T(0.5) * (pg[1 + a][fd] + pg[1 + fd][a]));
            if (a == fd) tau_fd[a] += mu2 * divv;
        }
        const T dT = (gam * mach2 * pg[D + 1][fd] - T_ * pg[0][fd]) / rho;
        const T q = adiabatic ? T(0) : -kappa * dT;
        f[fd][0] = T(0);
        for (int a = 0; a < D; ++a) f[fd][1 + a] = -tau_fd[a];
        T fe = T(0);
        for (int a = 0; a < D; ++a) fe -= vel[a] * tau_fd[a];
        f[fd][D + 1] = fe + q;
    }
}

} // namespace dgx
