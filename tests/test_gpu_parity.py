"""Device path vs the CPU FP64 oracle on identical seeded inputs, through the C ABI.

Bars (fixed, stated per check):
* integer connectivity bit-exact;
* FP64 residual: per DoF, |R_dev - R_oracle| <= 1e-12 x the sum of the absolute values of the
  terms summed into that DoF (two-point or flux-divergence volume terms, numerical and interior
  surface fluxes, viscous terms; ``Oracle.residual_magnitude``), the FP64 rounding scale of a
  sum evaluated in another order.  The conventional max-norm relative error
  ||R_dev - R_oracle||_inf / ||R_oracle||_inf is printed beside it: at M = 0.1 the fluxes are
  ~70x the residual (the pressure p_inf = 1/(gamma M^2) cancels), so that figure reaches a few
  1e-12 for split forms without any term being wrong;
* lifted right-hand side (M^-1 R): the same per-DoF bar with the magnitude lifted alike;
* time-stepped solutions: ||u_dev - u_oracle||_inf / ||u_oracle||_inf <= 1e-12;
* the reference's golden viscous-TGV kinetic energy reproduced at its own 1e-10 tolerance.
3D cases of every degree run on 4^3 (or 3^3) meshes, whose -x and +x neighbours of a cell are
different cells.
"""
import os
import sys

import numpy as np
import pytest

import oracle as O
import philip_amd as P

pytestmark = pytest.mark.gpu

TOL_RESIDUAL = 1e-12   # per DoF, relative to the DoF's term magnitude
TOL_STATE = 1e-12      # stepped solutions, max-norm relative


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def perturbed(u, o, seed, amp=1e-3):
    rng = np.random.default_rng(seed)
    v = u.reshape(o.ncells, o.ns, o.Np).copy()
    scale = np.abs(v).max(axis=(0, 2), keepdims=True)
    v += amp * scale * rng.standard_normal(v.shape)
    return v.ravel()


def rel(a, b):
    return max(np.abs(a - b).max() / np.abs(b).max(), np.linalg.norm(a - b) / np.linalg.norm(b))


def dof_rel_err(o, a, b, mag):
    """max over DoFs of |a - b| / magnitude; a DoF whose terms are all below 1e-8 of its state's
    largest magnitude is measured against that floor."""
    m = mag.reshape(o.ncells, o.ns, o.Np)
    den = np.maximum(m, 1e-8 * m.max(axis=(0, 2), keepdims=True))
    return float((np.abs(a - b).reshape(m.shape) / den).max())


def check_residual(o, d, u, label, lifted=True):
    """Device residual (and M^-1 R) against the oracle at the fixed per-DoF bar; returns the
    per-DoF error and the max-norm relative error."""
    r_ref = o.residual(u)
    mag = o.residual_magnitude(u)
    d.set_solution(u)
    r = d.assemble_residual()
    e = dof_rel_err(o, r, r_ref, mag)
    e_max = np.abs(r - r_ref).max() / np.abs(r_ref).max()
    msg = f"  {label}: per-DoF err {e:.2e} (bar {TOL_RESIDUAL:.0e}), max-norm rel err {e_max:.2e}"
    if lifted:
        el = dof_rel_err(o, d.apply_inverse_mass_to_rhs(), o.inverse_mass(r_ref), o.inverse_mass(mag))
        msg += f", lifted per-DoF err {el:.2e}"
        assert el <= TOL_RESIDUAL, label
    log(msg)
    assert e <= TOL_RESIDUAL, label
    return r, e, e_max


SCHEMES = {
    "lxf": dict(conv_flux=O.LXF),
    "roe": dict(conv_flux=O.ROE),
    "central": dict(conv_flux=O.CENTRAL),
    "ch_roe": dict(use_split_form=1, two_point_flux=O.CH, conv_flux=O.TWO_POINT_ROE),
    "ch_lxf": dict(use_split_form=1, two_point_flux=O.CH, conv_flux=O.TWO_POINT_LXF),
    "ra_ec": dict(use_split_form=1, two_point_flux=O.RA, conv_flux=O.TWO_POINT),
    "ir_ec": dict(use_split_form=1, two_point_flux=O.IR, conv_flux=O.TWO_POINT),
    "kg_roe": dict(use_split_form=1, two_point_flux=O.KG, conv_flux=O.TWO_POINT_ROE),
}


def make_pair(**kw):
    o = O.Oracle(**kw)
    d = P.DG(**kw)
    return o, d


def test_connectivity_bit_exact():
    for dim, N in [(3, 4), (3, 8), (3, 3), (2, 8), (2, 6)]:
        log(f"connectivity dim={dim} N={N}")
        o, d = make_pair(dim=dim, poly_degree=1, n_cells_1d=N)
        assert np.array_equal(d.neighbors(), o.neighbors())


@pytest.mark.parametrize("scheme", list(SCHEMES))
@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 6])
def test_residual_3d(scheme, p):
    kw = dict(dim=3, poly_degree=p, n_cells_1d=4, **SCHEMES[scheme])
    log(f"residual 3d scheme={scheme} p={p} 4^3")
    o, d = make_pair(**kw)
    u = perturbed(o.ic_tgv(), o, seed=10 + p)
    r, _, _ = check_residual(o, d, u, f"{scheme} p={p}")
    # run-to-run determinism: no reductions or races, so two launches agree bitwise
    assert np.array_equal(r, d.assemble_residual())


@pytest.mark.parametrize("scheme", ["lxf", "roe", "ch_roe", "ir_ec", "ra_ec"])
@pytest.mark.parametrize("p", [2, 3, 4, 5, 6])
def test_residual_2d_vortex(scheme, p):
    kw = dict(dim=2, poly_degree=p, n_cells_1d=4, domain_left=-5.0, domain_right=5.0, mach_inf=1.195228609334,
              **SCHEMES[scheme])
    log(f"residual 2d scheme={scheme} p={p}")
    o, d = make_pair(**kw)
    u = perturbed(o.ic_vortex(strength=3.0, u_inf=0.7, v_inf=0.9), o, seed=p)
    check_residual(o, d, u, f"2d {scheme} p={p}", lifted=False)


def ns_kw(p, diss, split, n_cells=4, **extra):
    return dict(dim=3, poly_degree=p, n_cells_1d=n_cells, pde=O.NAVIER_STOKES, diss_flux=diss,
                reynolds_inf=1600.0, prandtl=0.71, mach_inf=0.1, **extra,
                **(dict(use_split_form=1, two_point_flux=O.RA, conv_flux=O.TWO_POINT_ROE) if split else
                   dict(conv_flux=O.ROE)))


@pytest.mark.parametrize("diss", [O.SIPG, O.BR2])
@pytest.mark.parametrize("split", [0, 1])
@pytest.mark.parametrize("p", [2, 3, 4, 5, 6])
def test_residual_navier_stokes(diss, split, p):
    kw = ns_kw(p, diss, split)
    log(f"residual NS diss={diss} split={split} p={p} 4^3")
    o, d = make_pair(**kw)
    u = perturbed(o.ic_tgv(), o, seed=100 + p)
    check_residual(o, d, u, f"NS diss={diss} split={split} p={p}", lifted=False)
    for k in range(3):
        q, q_ref = d.auxiliary_solution(k), o.auxiliary(k)
        assert np.abs(q - q_ref).max() <= TOL_RESIDUAL * np.abs(q_ref).max()


@pytest.mark.parametrize("p", [3, 5])
def test_residual_constant_viscosity_and_lexicographic_mesh(p):
    kw = ns_kw(p, O.BR2, 1, n_cells=3, use_constant_viscosity=1, constant_viscosity=1.0)
    o, d = make_pair(**kw)
    u = perturbed(o.ic_tgv(), o, seed=7)
    check_residual(o, d, u, f"NS constant viscosity 3^3 p={p}", lifted=False)


@pytest.mark.parametrize("method", ["euler", "heun", "ssprk3", "rk4", "rk4_3_5_3sstar", "rk3_2_5f_3sstarplus",
                                    "rk4_3_9f_3sstarplus", "rk5_4_10f_3sstarplus"])
def test_time_stepping(method):
    kw = dict(dim=3, poly_degree=3, n_cells_1d=4, use_split_form=1, two_point_flux=O.CH,
              conv_flux=O.TWO_POINT_ROE)
    log(f"stepping {method}")
    o, d = make_pair(**kw)
    u = o.ic_tgv()
    d.set_solution(u)
    # the reference's TGV rule dt = CFL*dx ignores the sound speed (1/M = 10): CFL 0.01 is an
    # acoustic CFL of about 0.14 (CFL 0.1 is unstable and amplifies rounding ~20x per step)
    dt = o.constant_time_step(0.01)
    d.allocate_ode_system(method)
    d.step_in_time(dt, 3)
    for _ in range(3):
        o.step(method, dt, u)
    e = np.abs(d.solution() - u).max() / np.abs(u).max()
    log(f"  rel err {e:.3e} (bar {TOL_STATE:.0e})")
    assert e <= TOL_STATE
    assert d.current_iteration == 3 and abs(d.current_time - 3 * dt) < 1e-15


def test_lsrk_error_register_does_not_change_solution():
    kw = dict(dim=3, poly_degree=2, n_cells_1d=4, use_split_form=1, two_point_flux=O.CH,
              conv_flux=O.TWO_POINT_ROE)
    d1, d2 = P.DG(**kw), P.DG(**kw)
    o = O.Oracle(**kw)
    u = o.ic_tgv()
    for d, track in [(d1, False), (d2, True)]:
        d.set_solution(u)
        d.allocate_ode_system("rk3_2_5f_3sstarplus")
        d.set_track_error_estimate(track)
        d.step_in_time(0.01, 2)
    assert np.array_equal(d1.solution(), d2.solution())


def test_integrated_quantities_match_oracle():
    o, d = make_pair(dim=3, poly_degree=3, n_cells_1d=4)
    u = perturbed(o.ic_tgv(), o, seed=3)
    d.set_solution(u)
    qo, qd = o.integrated(u), d.integrated()
    for k in qo:
        assert abs(qd[k] - qo[k]) <= 1e-13 * abs(qo[k]), k


@pytest.mark.parametrize("dim", [2, 3])
def test_numerical_entropy_matches_oracle(dim):
    kw = dict(dim=dim, poly_degree=3, n_cells_1d=4) if dim == 3 else dict(
        dim=2, poly_degree=4, n_cells_1d=5, domain_left=-5.0, domain_right=5.0, mach_inf=1.195228609334)
    o, d = make_pair(**kw)
    u = perturbed(o.ic_tgv() if dim == 3 else o.ic_vortex(strength=3.0, u_inf=0.5, v_inf=0.4), o, seed=dim)
    d.set_solution(u)
    ref = o.numerical_entropy(u)
    assert abs(d.numerical_entropy() - ref) <= 1e-13 * abs(ref)


@pytest.mark.parametrize("dim", [2, 3])
def test_max_wave_speed_and_adaptive_step_match_oracle(dim):
    kw = dict(dim=3, poly_degree=3, n_cells_1d=4, mach_inf=0.2) if dim == 3 else dict(
        dim=2, poly_degree=5, n_cells_1d=4, domain_left=-5.0, domain_right=5.0, mach_inf=1.195228609334)
    o, d = make_pair(**kw)
    u = perturbed(o.ic_tgv() if dim == 3 else o.ic_vortex(strength=3.0, u_inf=0.5, v_inf=0.4), o, seed=5)
    d.set_solution(u)
    for overint in (0, 10):
        ref = o.max_wave_speed(u, overint)
        assert abs(d.max_wave_speed(overint) - ref) <= 1e-14 * ref
    ref_dt = o.adaptive_time_step(u, 0.2)
    assert abs(d.adaptive_time_step(0.2) - ref_dt) <= 1e-14 * ref_dt


def test_adaptive_run_matches_oracle_steps():
    kw = dict(dim=3, poly_degree=2, n_cells_1d=4, use_split_form=1, two_point_flux=O.CH, conv_flux=O.TWO_POINT_ROE)
    o, d = make_pair(**kw)
    u = o.ic_tgv()
    d.set_solution(u)
    steps = d.run("ssprk3", None, 0.02, adaptive_cfl=0.1)
    t, n = 0.0, 0
    while t < 0.02:
        dt = o.adaptive_time_step(u, 0.1)
        if t + dt > 0.02:
            dt = 0.02 - t
        o.step("ssprk3", dt, u)
        t += dt
        n += 1
    assert steps == n
    assert np.abs(d.solution() - u).max() / np.abs(u).max() <= TOL_STATE


def test_restart_continues_bitwise(tmp_path):
    """Restart equivalence (the reference's taylor_green_vortex_restart_check compares at 1e-10):
    N steps straight equal N/2 steps, save, a fresh solver restored from the file, N/2 steps."""
    kw = dict(dim=3, poly_degree=3, n_cells_1d=4, use_split_form=1, two_point_flux=O.CH, conv_flux=O.TWO_POINT_ROE,
              precision=P.MIXED)
    d1 = P.DG(**kw)
    d1.set_initial_condition_tgv()
    dt = d1.constant_time_step(0.01)
    table1 = []
    d1.run("rk3_2_5f_3sstarplus", dt, 6 * dt, table=table1)
    d2 = P.DG(**kw)
    d2.set_initial_condition_tgv()
    table2 = []
    d2.run("rk3_2_5f_3sstarplus", dt, 3 * dt, table=table2)
    d2.save_restart(tmp_path / "restart.npz", table2)
    d3 = P.DG(**kw)
    table3 = d3.load_restart(tmp_path / "restart.npz")
    d3.allocate_ode_system("rk3_2_5f_3sstarplus")
    d3.run("rk3_2_5f_3sstarplus", dt, 6 * dt, table=table3, allocate=False)
    assert np.array_equal(d1.solution(), d3.solution())
    # (the accumulated time may fall an ulp short of 6 dt: a clipped extra step is taken in both runs)
    assert d3.current_iteration == d1.current_iteration and len(table3) == len(table1) == d1.current_iteration
    for a, b in zip(table1, table3):
        assert abs(a["kinetic_energy"] - b["kinetic_energy"]) <= 1e-10 * a["kinetic_energy"]
    P.write_unsteady_table(tmp_path / "table.txt", table3)
    assert P.read_unsteady_table(tmp_path / "table.txt") == table3
    with pytest.raises(P.DeviceError):
        P.DG(**dict(kw, n_cells_1d=8)).load_restart(tmp_path / "restart.npz")


def test_reference_golden_nsfr_gll_on_device():
    """The reference's own golden value, reproduced by the device path end to end
    (viscous_taylor_green_vortex_energy_check_nsfr_gll_flux_nodes_quick.prm:59-60,
    taylor_green_vortex_energy_check.cpp:37-41, 1e-10 relative)."""
    d = P.DG(dim=3, poly_degree=2, n_cells_1d=4, pde=P.NAVIER_STOKES, diss_flux=P.SIPG, mach_inf=0.1,
             reynolds_inf=1600.0, prandtl=0.71, use_split_form=1, two_point_flux=P.IR, conv_flux=P.TWO_POINT)
    o = O.Oracle(dim=3, poly_degree=2, n_cells_1d=4)
    # initial condition from the oracle (test input); the device integrates and reduces
    o2 = O.Oracle(dim=3, poly_degree=2, n_cells_1d=4, pde=O.NAVIER_STOKES, mach_inf=0.1)
    d.set_solution(o2.ic_tgv())
    dt = d.constant_time_step(0.003)
    nsteps = d.run("ssprk3", dt, 1.2566370614400000e-02)
    assert nsteps == 9
    q = d.integrated()
    log(f"  KE {q['kinetic_energy']!r} eps {q['vorticity_dissipation']!r}")
    assert abs(q["kinetic_energy"] - 1.2071670877794241e-01) / 1.2071670877794241e-01 < 1e-10
    assert abs(q["vorticity_dissipation"] - 4.5392245903930087e-04) / 4.5392245903930087e-04 < 1e-10
    del o


def test_invalid_configurations_raise():
    with pytest.raises(P.DeviceError):
        P.DG(dim=3, poly_degree=2, flux_nodes=P.GL)
    with pytest.raises(P.DeviceError):
        P.DG(dim=3, poly_degree=2, conv_flux=P.TWO_POINT, use_split_form=0)
    with pytest.raises(P.DeviceError):
        P.DG(dim=3, poly_degree=9)
