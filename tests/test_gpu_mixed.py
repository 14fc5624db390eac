"""Mixed-precision device path (FP32 fluxes/contractions, FP64 state and RK
registers) against the FP64 oracle.

Bars: the kinetic-energy history of a stepped TGV stays within 1e-6 relative
of the FP64 oracle's at every step; the mixed residual's relative error is
reported and bounded by 5e-4 (FP32 rounding amplified by the same
conditioning that puts the FP64 floor near 1e-12; measured values are printed).
"""
import sys

import numpy as np
import pytest

import oracle as O
import philip_amd as P

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return max(np.abs(a - b).max() / np.abs(b).max(), np.linalg.norm(a - b) / np.linalg.norm(b))


CFG2 = dict(dim=3, poly_degree=3, use_split_form=1, two_point_flux=O.CH, conv_flux=O.TWO_POINT_ROE, mach_inf=0.1)


@pytest.mark.parametrize("scheme", [
    dict(conv_flux=O.LXF),
    dict(use_split_form=1, two_point_flux=O.CH, conv_flux=O.TWO_POINT_ROE),
    dict(use_split_form=1, two_point_flux=O.RA, conv_flux=O.TWO_POINT_LXF),
])
def test_mixed_residual_error(scheme):
    kw = dict(dim=3, poly_degree=3, n_cells_1d=4, mach_inf=0.1, **scheme)
    o = O.Oracle(**kw)
    d = P.DG(precision=P.MIXED, **kw)
    u = o.ic_tgv()
    d.set_solution(u)
    e = rel(d.assemble_residual(), o.residual(u))
    log(f"mixed residual rel err {e:.3e} for {scheme}")
    assert e <= 5e-4


@pytest.mark.parametrize("diss", [O.SIPG, O.BR2])
@pytest.mark.parametrize("p", [2, 3, 4, 5, 6])
def test_mixed_navier_stokes_residual_error(diss, p):
    """Split-form Navier-Stokes residual of the mixed path (line kernel with the viscous terms,
    float lifted gradient) against the FP64 oracle on a perturbed TGV state, 4^3 cells (the -x
    and +x neighbours of a cell differ): p=3 has elements sharing faces inside a block, p=4..6
    split each line over two or three items."""
    kw = dict(dim=3, poly_degree=p, n_cells_1d=4, pde=O.NAVIER_STOKES, diss_flux=diss,
              reynolds_inf=1600.0, prandtl=0.71, mach_inf=0.1, use_split_form=1, two_point_flux=O.RA,
              conv_flux=O.TWO_POINT_ROE)
    o = O.Oracle(**kw)
    d = P.DG(precision=P.MIXED, **kw)
    rng = np.random.default_rng(300 + p)
    v = o.ic_tgv().reshape(o.ncells, o.ns, o.Np)
    v = (v + 1e-3 * np.abs(v).max(axis=(0, 2), keepdims=True) * rng.standard_normal(v.shape)).ravel()
    d.set_solution(v)
    e = rel(d.assemble_residual(), o.residual(v))
    eq = max(rel(d.auxiliary_solution(k), o.auxiliary(k)) for k in range(3))
    log(f"mixed NS residual rel err {e:.3e}, lifted gradient {eq:.3e} (diss={diss}, p={p})")
    assert e <= 5e-4 and eq <= 1e-6


def test_mixed_ke_history_within_1e6():
    kw = dict(n_cells_1d=8, **CFG2)
    o = O.Oracle(**kw)
    d = P.DG(precision=P.MIXED, **kw)
    u = o.ic_tgv()
    d.set_initial_condition_tgv()
    assert rel(d.solution(), u) < 1e-14  # same initial data on both sides
    dt = o.constant_time_step(0.01)  # acoustic CFL ~0.14 at M = 0.1
    d.allocate_ode_system("rk3_2_5f_3sstarplus")
    worst = 0.0
    for step in range(20):
        o.step("rk3_2_5f_3sstarplus", dt, u)
        d.step_in_time(dt, 1)
        ke_o = o.integrated(u)["kinetic_energy"]
        ke_d = d.integrated()["kinetic_energy"]
        worst = max(worst, abs(ke_d - ke_o) / ke_o)
    log(f"mixed KE history worst rel diff over 20 LSRK3 steps: {worst:.3e}")
    assert worst <= 1e-6


def test_fp64_ke_history_config2_scheme():
    kw = dict(n_cells_1d=8, **CFG2)
    o = O.Oracle(**kw)
    d = P.DG(precision=P.FP64, **kw)
    u = o.ic_tgv()
    d.set_solution(u)
    dt = o.constant_time_step(0.01)
    d.allocate_ode_system("rk3_2_5f_3sstarplus")
    worst = 0.0
    for step in range(10):
        o.step("rk3_2_5f_3sstarplus", dt, u)
        d.step_in_time(dt, 1)
        ke_o = o.integrated(u)["kinetic_energy"]
        worst = max(worst, abs(d.integrated()["kinetic_energy"] - ke_o) / ke_o)
    log(f"FP64 KE history worst rel diff over 10 LSRK3 steps: {worst:.3e}")
    assert worst <= 1e-12


def test_mixed_entropy_drift_config2_scheme():
    """Entropy-conserving split form: the numerical-entropy change over 10 LSRK3 steps in the mixed
    path against the FP64 oracle's (the drift relative to |S_0|; bar 1e-6 of |S_0|)."""
    kw = dict(n_cells_1d=8, **CFG2)
    o = O.Oracle(**kw)
    d = P.DG(precision=P.MIXED, **kw)
    u = o.ic_tgv()
    d.set_solution(u)
    s0 = o.numerical_entropy(u)
    assert abs(d.numerical_entropy() - s0) <= 1e-13 * abs(s0)
    dt = o.constant_time_step(0.01)
    d.allocate_ode_system("rk3_2_5f_3sstarplus")
    worst = 0.0
    for _ in range(10):
        o.step("rk3_2_5f_3sstarplus", dt, u)
        d.step_in_time(dt, 1)
        worst = max(worst, abs(d.numerical_entropy() - o.numerical_entropy(u)) / abs(s0))
    log(f"mixed numerical-entropy drift vs FP64 oracle over 10 LSRK3 steps: {worst:.3e} of |S0|")
    assert worst <= 1e-6


def test_mixed_navier_stokes_enstrophy_and_dissipation():
    """Navier-Stokes TGV (Re 1600, BR2, Ranocha split form + EC/Roe, RK4): enstrophy and the
    vorticity-based dissipation rate of the mixed path against the FP64 oracle over 5 steps
    (bar 1e-5 relative; measured values printed)."""
    kw = dict(dim=3, poly_degree=3, n_cells_1d=4, pde=O.NAVIER_STOKES, diss_flux=O.BR2, use_constant_viscosity=1,
              constant_viscosity=1.0, reynolds_inf=1600.0, prandtl=0.71, mach_inf=0.1, use_split_form=1,
              two_point_flux=O.RA, conv_flux=O.TWO_POINT_ROE)
    o = O.Oracle(**kw)
    d = P.DG(precision=P.MIXED, **kw)
    u = o.ic_tgv()
    d.set_solution(u)
    dt = o.constant_time_step(0.01)
    d.allocate_ode_system("rk4")
    worst_e, worst_eps = 0.0, 0.0
    for _ in range(5):
        o.step("rk4", dt, u)
        d.step_in_time(dt, 1)
        qo, qd = o.integrated(u), d.integrated()
        worst_e = max(worst_e, abs(qd["enstrophy"] - qo["enstrophy"]) / qo["enstrophy"])
        worst_eps = max(worst_eps, abs(qd["vorticity_dissipation"] - qo["vorticity_dissipation"])
                        / qo["vorticity_dissipation"])
    log(f"mixed NS enstrophy rel err {worst_e:.3e}, dissipation rate rel err {worst_eps:.3e} over 5 RK4 steps")
    assert worst_e <= 1e-5 and worst_eps <= 1e-5
