"""Pins the CPU oracle against the reference's own golden values.

Reference: src/testing/taylor_green_vortex_energy_check.cpp:37,41 compares the
kinetic energy and the vorticity-based dissipation rate at the final time with
the values in the .prm files at 1e-10 relative:
  tests/integration_tests_control_files/taylor_green_vortex_integration/
    viscous_taylor_green_vortex_energy_check_strong_quick.prm:66-67
    viscous_taylor_green_vortex_energy_check_nsfr_gll_flux_nodes_quick.prm:59-60
    viscous_taylor_green_vortex_energy_check_nsfr_gl_flux_nodes_quick.prm:59-60
All: 3D Navier-Stokes TGV, p=2, 4^3 cells on [0, 2pi]^3, M=0.1, Re=1600,
Pr=0.71, Sutherland viscosity, SIPG, SSPRK3, CFL 0.003, t_f=0.0125663706144.
"""
import pytest

import oracle as O

CASES = [
    ("strong_gl_roe_sipg", dict(flux_nodes=O.GL, conv_flux=O.ROE),
     1.2073987154899971e-01, 4.5422272551211095e-04),
    ("nsfr_ir_gll", dict(flux_nodes=O.GLL, use_split_form=1, two_point_flux=O.IR, conv_flux=O.TWO_POINT),
     1.2071670877794241e-01, 4.5392245903930087e-04),
    ("nsfr_ir_gl", dict(flux_nodes=O.GL, use_split_form=1, two_point_flux=O.IR, conv_flux=O.TWO_POINT),
     1.2073959489246434e-01, 4.5399594835606567e-04),
]


@pytest.mark.parametrize("name,kw,ke_ref,eps_ref", CASES, ids=[c[0] for c in CASES])
def test_viscous_tgv_golden(name, kw, ke_ref, eps_ref):
    o = O.Oracle(dim=3, poly_degree=2, pde=O.NAVIER_STOKES, diss_flux=O.SIPG, mach_inf=0.1,
                 reynolds_inf=1600.0, prandtl=0.71, n_cells_1d=4, **kw)
    u = o.ic_tgv()
    dt = o.constant_time_step(0.003)
    nsteps, _ = o.run("ssprk3", dt, 1.2566370614400000e-02, u)
    # 8 full steps plus the clipped last step (end_exactly_at_final_time)
    assert nsteps == 9
    q = o.integrated(u)
    assert abs(q["kinetic_energy"] - ke_ref) / ke_ref < 1e-10
    assert abs(q["vorticity_dissipation"] - eps_ref) / eps_ref < 1e-10
