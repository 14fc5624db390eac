// Host build of the device boundary states (philip_amd/csrc/dg_physics.hpp boundary_state) for
// CPU checks against the oracle.  Reads from stdin:
//   n dim mach angle_of_attack precision(0 float, 1 double)   then n lines of: type K sg u[dim+2]
// and prints the exterior state of each line.
#include <cmath>
#include <cstdio>
#define __device__
#define __forceinline__ inline
#define DGX_HOST_TEST
#include "../../philip_amd/csrc/dg_fast.hpp"
using namespace dgx;

template <typename T, int D> static int run(int n, const PhysConst &pc)
{
    for (int i = 0; i < n; ++i) {
        int type, K;
        double sg, x[5];
        if (scanf("%d %d %lf", &type, &K, &sg) != 3) return 1;
        for (int s = 0; s < D + 2; ++s)
            if (scanf("%lf", &x[s]) != 1) return 1;
        T ui[5], ub[5];
        for (int s = 0; s < D + 2; ++s) ui[s] = T(x[s]);
        boundary_state<T, D>(type, ui, K, T(sg), pc, ub);
        for (int s = 0; s < D + 2; ++s) printf("%.17e ", (double)ub[s]);
        printf("\n");
    }
    return 0;
}

int main()
{
    int n, dim, prec;
    double mach, aoa;
    if (scanf("%d %d %lf %lf %d", &n, &dim, &mach, &aoa, &prec) != 5) return 1;
    PhysConst pc{};
    pc.gam = 1.4;
    pc.gamm1 = 0.4;
    pc.mach2 = mach * mach;
    pc.vinf[0] = std::cos(aoa);
    pc.vinf[1] = std::sin(aoa);
    pc.vinf[2] = 0.0;
    if (dim == 2) return prec ? run<double, 2>(n, pc) : run<float, 2>(n, pc);
    return prec ? run<double, 3>(n, pc) : run<float, 3>(n, pc);
}
