// Host build of the mixed path's float point physics (philip_amd/csrc/dg_fast.hpp)
// for CPU checks against the oracle.  Reads from stdin:
//   n two_point conv_flux gamma mach   then n lines of ui[5] ue[5]
// and prints, per state pair and direction K = 0, 1, 2: the two-point flux
// F#(ui, ue) along e_K (5 values) and the surface numerical flux along +e_K
// with ui interior (5 values).
#include <cmath>
#include <cstdio>
#define __device__
#define __forceinline__ inline
#define DGX_HOST_TEST
#include "../../philip_amd/csrc/dg_fast.hpp"
using namespace dgx;

template <int K> static void eval(const NodeF<3> &a, const NodeF<3> &b, const FastConst &fc, const PhysConst &pc)
{
    float F[5], G[5];
    pair_flux_fast<3, K>(a, b, fc, pc, F);
    num_flux_fast<3, K>(a, b, 1.f, fc, pc, G);
    for (int s = 0; s < 5; ++s) printf("%.9e ", F[s]);
    for (int s = 0; s < 5; ++s) printf("%.9e ", G[s]);
}

int main()
{
    int n, tp, cf;
    double gam, mach;
    if (scanf("%d %d %d %lf %lf", &n, &tp, &cf, &gam, &mach) != 5) return 1;
    PhysConst pc{};
    pc.gam = gam;
    pc.gamm1 = gam - 1.0;
    pc.mach2 = mach * mach;
    pc.two_point = tp;
    pc.conv_flux = cf;
    const FastConst fc = fast_const(pc);
    for (int i = 0; i < n; ++i) {
        double x[10];
        for (int j = 0; j < 10; ++j)
            if (scanf("%lf", &x[j]) != 1) return 1;
        NodeF<3> a, b;
        for (int s = 0; s < 5; ++s) {
            a.u[s] = float(x[s]);
            b.u[s] = float(x[5 + s]);
        }
        node_fill<3>(a, fc);
        node_fill<3>(b, fc);
        eval<0>(a, b, fc, pc);
        eval<1>(a, b, fc, pc);
        eval<2>(a, b, fc, pc);
        printf("\n");
    }
    return 0;
}
