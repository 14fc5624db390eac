"""Relaxation Runge-Kutta on the device against the oracle's restatement of the reference's
root-finding RRK (rrk_ode_solver_base.cpp:16-35, root_finding_rrk_ode_solver.cpp:16-390,
runge_kutta_ode_solver.cpp:101-109).  With an entropy-conservative split form and surface flux
the semi-discrete entropy is conserved, and RRK conserves the integrated numerical entropy of
the fully discrete solution up to the root tolerance (plain RK4 drifts): a property the oracle's
restatement satisfies (tests/test_oracle_rrk.py) and the device must too (1e-12 relative drift
over six steps).

The relaxation parameter is the root of a difference of integrated entropies that nearly
cancel, found to a 5e-10 residual tolerance: in the oracle itself a 1e-15 relative perturbation
of the initial state moves gamma by 2e-8 .. 2.5e-7 over six steps, and the state by up to 1e-9.
Bars: gamma, state and time within five times the oracle's own largest spread over six such
perturbations (measured in the test, on a single-threaded and so run-to-run deterministic
oracle); a wrong entropy, direction or root moves gamma by O(dt^p) ~ 1e-3."""
import os
import sys

import numpy as np
import pytest

import oracle as O
import philip_amd as P

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, scope="module")
def deterministic_oracle():
    """One oracle thread: the OpenMP neighbour scatter adds face terms in a varying order, and these
    ill-conditioned comparisons would see that rounding as run-to-run noise in the reference."""
    O.set_threads(1)
    yield
    O.set_threads(min(16, os.cpu_count() or 1))


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("flux,method", [(O.TWO_POINT, "rk4"), (O.TWO_POINT_ROE, "rk4"), (O.TWO_POINT, "ssprk3")])
def test_rrk_matches_oracle(flux, method):
    kw = dict(dim=3, poly_degree=2, n_cells_1d=4, use_split_form=1, two_point_flux=O.CH, conv_flux=flux,
              mach_inf=0.1)
    o, d = O.Oracle(**kw), P.DG(**kw)
    u = o.ic_tgv(isothermal=True)
    d.set_solution(u)
    d.allocate_ode_system(method)
    d.set_relaxation_runge_kutta(True)
    s0 = o.numerical_entropy(u)
    dt = o.constant_time_step(0.05)

    def oracle_run(v):
        g, t, gs = 1.0, 0.0, []
        for _ in range(6):
            adv, g = o.step_rrk(method, dt, v, g)
            t += adv
            gs.append(g)
        return np.array(gs), v, t

    u0 = u.copy()
    gs_dev = []
    for _ in range(6):
        d.step_in_time(dt, 1)
        gs_dev.append(d.relaxation_parameter)
    gs_ref, u, t_ref = oracle_run(u)
    rng = np.random.default_rng(3)
    sg = su = st = 0.0
    for pert in (1e-15, 1e-15, 1e-15, 1e-14, 1e-14, 1e-14):
        gp, up, tp = oracle_run(u0 * (1.0 + pert * rng.choice([-1.0, 1.0], u0.shape)))
        sg = max(sg, np.abs(gp - gs_ref).max())
        su = max(su, np.abs(up - u).max() / np.abs(u).max())
        st = max(st, abs(tp - t_ref))
    worst_g = np.abs(np.array(gs_dev) - gs_ref).max()
    e_u = np.abs(d.solution() - u).max() / np.abs(u).max()
    e_t = abs(d.current_time - t_ref)
    drift = abs(d.numerical_entropy() - s0) / abs(s0)
    log(f"RRK {method} flux={flux}: gamma {gs_ref[-1]:.12f}, |gamma_dev - gamma_ref| {worst_g:.2e} "
        f"(oracle spread {sg:.2e}), state {e_u:.2e} ({su:.2e}), time {e_t:.1e} ({st:.1e}), "
        f"entropy drift {drift:.2e}")
    assert worst_g <= 5.0 * sg and e_u <= max(5.0 * su, 1e-11) and e_t <= max(5.0 * st, 1e-11 * t_ref)
    if flux == O.TWO_POINT:
        assert drift <= 1e-12
