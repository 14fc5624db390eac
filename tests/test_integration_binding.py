"""The C++ binding INTEGRATION.md shows a maintainer adding on the reference side compiles
against include/philip_dg.h and links against the in-tree library (no GPU: nothing runs).
deal.II and PHiLiP's parameter classes are absent here, so minimal stand-ins with the
reference's names (all_parameters.h:104-306, parameters_euler.h, parameters_navier_stokes.h,
deal.II's distributed vector begin()) stand in for them."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STANDINS = r'''
#include <cstdint>
#include <vector>
namespace dealii { namespace LinearAlgebra { namespace distributed {
template <typename T> struct Vector {
    std::vector<T> v;
    T *begin() { return v.data(); }
    const T *begin() const { return v.data(); }
};
}}}
namespace PHiLiP { namespace Parameters {
struct EulerParam { double mach_inf = 0.1, gamma_gas = 1.4, angle_of_attack = 0.0; };
struct NavierStokesParam {
    double prandtl_number = 0.72, reynolds_number_inf = 1600.0, temperature_inf = 273.15;
    double nondimensionalized_isothermal_wall_temperature = 1.0;
    enum ThermalBoundaryCondition { adiabatic, isothermal };
    ThermalBoundaryCondition thermal_boundary_condition_type = adiabatic;
    bool use_constant_viscosity = false;
    double nondimensionalized_constant_viscosity = 1.0;
};
struct AllParameters {
    enum FluxNodes { GL, GLL };
    enum TwoPointNumericalFlux { KG, IR, CH, Ra };
    enum PartialDifferentialEquation { advection, diffusion, convection_diffusion, advection_vector,
        burgers_inviscid, burgers_viscous, burgers_rewienski, euler, mhd, navier_stokes };
    enum ConvectiveNumericalFlux { lax_friedrichs, roe, l2roe, central_flux, two_point_flux,
        two_point_flux_with_lax_friedrichs_dissipation, two_point_flux_with_roe_dissipation,
        two_point_flux_with_l2roe_dissipation };
    enum DissipativeNumericalFlux { symm_internal_penalty, bassi_rebay_2, central_visc_flux };
    enum NonPhysicalBehaviorEnum { return_big_number, abort_run, print_warning };
    FluxNodes flux_nodes_type = GLL;
    bool use_split_form = true;
    TwoPointNumericalFlux two_point_num_flux_type = CH;
    PartialDifferentialEquation pde_type = euler;
    ConvectiveNumericalFlux conv_num_flux_type = two_point_flux_with_roe_dissipation;
    DissipativeNumericalFlux diss_num_flux_type = symm_internal_penalty;
    NonPhysicalBehaviorEnum non_physical_behavior_type = return_big_number;
    double sipg_penalty_factor = 1.0;
    EulerParam euler_param;
    NavierStokesParam navier_stokes_param;
};
}}
'''

MAIN = r'''
struct MeshType {};
int main(int argc, char **) {
    if (argc < 100) return 0;  // linked, never run here
    PHiLiP::Parameters::AllParameters prm;
    PHiLiP::DGDevice<3, 1, double, MeshType> dg(prm, 3, 4, 0.0, 6.283185307179586);
    const int channel[6] = {1003, 1002, 1006, 1006, 0, 0};  // inflow, outflow, slip walls, periodic z
    PHiLiP::DGDevice<3, 1, double, MeshType> dg_channel(prm, 3, 4, 0.0, 1.0, channel);
    dealii::LinearAlgebra::distributed::Vector<double> u, r;
    dg.assemble_residual(u, r);
    dg.apply_inverse_global_mass_matrix(r);
    dg.allocate_ode_system(DGX_RK3_2_5F_3SSTARPLUS);
    dg.step_in_time(1e-3);
    dg.get_solution(u);
    dgx_request req{};
    return batch_and_serve({}, {}, {}, {}, {}, req, u.begin()) + (int)dg.current_iteration();
}
'''


def test_integration_binding_compiles_and_links(tmp_path):
    md = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", md, flags=re.S)
    assert len(blocks) >= 2
    src = tmp_path / "binding.cpp"
    src.write_text(STANDINS + "\n".join(blocks) + MAIN)
    lib = os.path.join(ROOT, "philip_amd", "libphilip_dg.so")
    if not os.path.exists(lib):
        pytest.fail("libphilip_dg.so not built (run __graft_entry__.build())")
    cmd = ["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
           str(tmp_path / "binding"), lib, "-Wl,--unresolved-symbols=ignore-in-shared-libs"]
    out = subprocess.run(cmd, capture_output=True, text=True)
    assert out.returncode == 0, out.stderr[-4000:]
