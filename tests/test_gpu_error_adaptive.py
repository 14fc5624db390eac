"""Error-adaptive low-storage RK on the device against the oracle's restatement of
FlowSolver::run with error_adaptive_time_step (flow_solver.cpp:505-507, 554-607;
low_storage_runge_kutta_ode_solver.cpp:89-178): the initial step size, the sequence of step
sizes chosen by the embedded-error PID controller (3S*+ with the FSAL residual, and the 3S*
method whose embedded solution comes from sum_stages) and the final solution.  Bars: step sizes
1e-12 relative each (the error norms are FP64 sums in another order), final state 1e-11."""
import sys

import numpy as np
import pytest

import oracle as O
import philip_amd as P

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("method", ["rk3_2_5f_3sstarplus", "rk4_3_5_3sstar", "rk4_3_9f_3sstarplus"])
def test_error_adaptive_run_matches_oracle(method):
    kw = dict(dim=3, poly_degree=2, n_cells_1d=4, use_split_form=1, two_point_flux=O.CH, conv_flux=O.TWO_POINT_ROE)
    o, d = O.Oracle(**kw), P.DG(**kw)
    u = o.ic_tgv()
    d.set_solution(u)
    tol = dict(atol=1e-7, rtol=1e-7, beta=(0.70, -0.23, 0.0))
    d.allocate_ode_system(method)
    d.set_error_adaptive_parameters(**tol)
    h_dev, h_ref = d.automatic_initial_step_size(), o.initial_step_size(u)
    assert abs(h_dev - h_ref) <= 1e-13 * h_ref
    final = 0.02
    steps = d.run(method, None, final, error_adaptive=tol, allocate=False)
    dts = o.run_error_adaptive(method, final, u, atol=tol["atol"], rtol=tol["rtol"], beta=tol["beta"])
    e_dt = max(abs(a - b) / b for a, b in zip(d.last_dts, dts))
    e_u = np.abs(d.solution() - u).max() / np.abs(u).max()
    log(f"{method}: {steps} steps (oracle {len(dts)}), dt {dts[0]:.3e} .. {max(dts):.3e}, "
        f"worst dt rel diff {e_dt:.2e}, final state {e_u:.2e}")
    assert steps == len(dts) > 5 and e_dt <= 1e-12 and e_u <= 1e-11
