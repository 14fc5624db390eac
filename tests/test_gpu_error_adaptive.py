"""Error-adaptive low-storage RK on the device against the oracle's restatement of
FlowSolver::run with error_adaptive_time_step (flow_solver.cpp:505-507, 554-607;
low_storage_runge_kutta_ode_solver.cpp:89-178).

The PID step-size controller and the embedded-error norm are checked on steps whose error is
resolved (dt = 1e-2 .. 1.5e-2 on the 4^3 p=2 TGV, error norms 0.5 .. 125): the next step size
agrees with the oracle's to 1e-9 relative, where perturbing the oracle's own initial state by
1e-15 .. 1e-13 moves it by at most ~1e-10, and a wrong exponent, weight or history term moves it
by O(1e-2) or more.  All three PID terms are exercised (beta = 0.70, -0.40, 0.10).

The reference's full loop is then run from its automatic initial step.  That step is tiny
(~8e-6 here), so the first embedded-error estimate is at rounding level and every later step
size inherits that noise: in the oracle itself, a 1e-15 relative perturbation of the initial
state moves the step sizes by up to ~1e-2.  The loop is therefore checked by the step count,
the final state (1e-11 relative) and step sizes within five times the oracle's own spread under
such perturbations (measured in the test).  The initial step itself is a ratio of right-hand-side
max-norms and agrees to 1e-11."""
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

METHODS = ["rk3_2_5f_3sstarplus", "rk4_3_5_3sstar", "rk4_3_9f_3sstarplus"]
KW = dict(dim=3, poly_degree=2, n_cells_1d=4, use_split_form=1, two_point_flux=O.CH, conv_flux=O.TWO_POINT_ROE)


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def pid_next(eps, beta, k, dt):
    return pow(eps[0], beta[0] / k) * pow(eps[1], beta[1] / k) * pow(eps[2], beta[2] / k) * dt


@pytest.mark.parametrize("method", METHODS)
def test_error_adaptive_controller_matches_oracle(method):
    o, d = O.Oracle(**KW), P.DG(**KW)
    u = o.ic_tgv()
    d.set_solution(u)
    tol = dict(atol=1e-7, rtol=1e-7, beta=(0.70, -0.40, 0.10))
    d.allocate_ode_system(method)
    d.set_error_adaptive_parameters(**tol)
    k = O.RK_ORDER[method]
    eps = [1.0, 1.0, 1.0]
    worst = 0.0
    for dt in (1e-2, 1.2e-2, 1.5e-2):
        d.step_in_time(dt, 1)
        got = d.error_adaptive_step_size(dt)
        w = o.step_error_norm(method, dt, u, tol["atol"], tol["rtol"])
        eps = [1.0 / w, eps[0], eps[1]]
        want = pid_next(eps, tol["beta"], k, dt)
        worst = max(worst, abs(got - want) / want)
        log(f"{method}: dt {dt:.3e} error norm {w:.3e} next {want:.6e} device {got:.6e}")
    e_u = np.abs(d.solution() - u).max() / np.abs(u).max()
    log(f"  worst next-step rel diff {worst:.2e} (bar 1e-9), state {e_u:.2e} (bar 1e-11)")
    assert worst <= 1e-9 and e_u <= 1e-11


@pytest.mark.parametrize("method", METHODS)
def test_error_adaptive_run_matches_oracle(method):
    o, d = O.Oracle(**KW), P.DG(**KW)
    u = o.ic_tgv()
    d.set_solution(u)
    tol = dict(atol=1e-7, rtol=1e-7, beta=(0.70, -0.23, 0.0))
    d.allocate_ode_system(method)
    d.set_error_adaptive_parameters(**tol)
    h_dev, h_ref = d.automatic_initial_step_size(), o.initial_step_size(u)
    assert abs(h_dev - h_ref) <= 1e-11 * h_ref
    final = 0.02
    u0 = u.copy()
    steps = d.run(method, None, final, error_adaptive=tol, allocate=False)
    dts = o.run_error_adaptive(method, final, u, atol=tol["atol"], rtol=tol["rtol"], beta=tol["beta"])
    # the oracle's own step-size spread under rounding-level perturbations of the initial state
    rng = np.random.default_rng(5)
    spread = 0.0
    for pert in (1e-15, 1e-15, 1e-14, 1e-14, 1e-13, 1e-13):
        up = u0 * (1.0 + pert * rng.choice([-1.0, 1.0], u0.shape))
        dp = o.run_error_adaptive(method, final, up, atol=tol["atol"], rtol=tol["rtol"], beta=tol["beta"])
        assert len(dp) == len(dts)
        spread = max(spread, max(abs(a - b) / b for a, b in zip(dp, dts)))
    e_dt = max(abs(a - b) / b for a, b in zip(d.last_dts, dts))
    e_u = np.abs(d.solution() - u).max() / np.abs(u).max()
    log(f"{method}: {steps} steps (oracle {len(dts)}), dt {dts[0]:.3e} .. {max(dts):.3e}, "
        f"worst dt rel diff {e_dt:.2e} (oracle spread {spread:.2e}), final state {e_u:.2e}")
    assert steps == len(dts) > 5 and e_dt <= 5.0 * spread and e_u <= 1e-11
