// Host build of the line kernels' cached-node point physics (philip_amd/csrc/dg_fast.hpp) in
// float (mixed path) or double (FP64 path) for CPU checks against the oracle.  Reads from stdin:
//   n two_point conv_flux gamma mach precision(0 float, 1 double)   then n lines of ui[5] ue[5]
// and prints, per state pair and direction K = 0, 1, 2: the two-point flux F#(ui, ue) along e_K
// (5 values) and the surface numerical flux along +e_K with ui interior (5 values).
#include <cmath>
#include <cstdio>
#define __device__
#define __forceinline__ inline
#define DGX_HOST_TEST
#include "../../philip_amd/csrc/dg_fast.hpp"
using namespace dgx;

template <typename T, int K>
static void eval(const NodeC<T, 3> &a, const NodeC<T, 3> &b, const FastConst<T> &fc, const PhysConst &pc)
{
    T F[5], G[5];
    pair_flux_fast<T, 3, K>(a, b, fc, pc, F);
    num_flux_fast<T, 3, K>(a, b, T(1), fc, pc, G);
    for (int s = 0; s < 5; ++s) printf("%.17e ", (double)F[s]);
    for (int s = 0; s < 5; ++s) printf("%.17e ", (double)G[s]);
}

template <typename T> static int run(int n, const PhysConst &pc)
{
    const FastConst<T> fc = fast_const<T>(pc);
    for (int i = 0; i < n; ++i) {
        double x[10];
        for (int j = 0; j < 10; ++j)
            if (scanf("%lf", &x[j]) != 1) return 1;
        NodeC<T, 3> a, b;
        for (int s = 0; s < 5; ++s) {
            a.u[s] = T(x[s]);
            b.u[s] = T(x[5 + s]);
        }
        node_fill<T, 3>(a, fc);
        node_fill<T, 3>(b, fc);
        eval<T, 0>(a, b, fc, pc);
        eval<T, 1>(a, b, fc, pc);
        eval<T, 2>(a, b, fc, pc);
        printf("\n");
    }
    return 0;
}

int main()
{
    int n, tp, cf, prec;
    double gam, mach;
    if (scanf("%d %d %d %lf %lf %d", &n, &tp, &cf, &gam, &mach, &prec) != 6) return 1;
    PhysConst pc{};
    pc.gam = gam;
    pc.gamm1 = gam - 1.0;
    pc.mach2 = mach * mach;
    pc.two_point = tp;
    pc.conv_flux = cf;
    return prec ? run<double>(n, pc) : run<float>(n, pc);
}
