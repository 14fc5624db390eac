"""The oracle's relaxation RK (root_finding_rrk_ode_solver.cpp:16-390) pinned by the property it
exists for: with an entropy-conservative split form and surface flux (semi-discrete entropy
conservation, euler_entropy_conserving_split_forms_check.cpp) the fully discrete integrated
numerical entropy is conserved to round-off, while plain RK4 drifts; the relaxation parameter is
1 + O(dt^p) and the time advanced is gamma dt."""
import numpy as np

import oracle as O


def test_rrk_conserves_entropy_where_rk4_drifts():
    kw = dict(dim=3, poly_degree=2, n_cells_1d=3, use_split_form=1, two_point_flux=O.CH, conv_flux=O.TWO_POINT,
              mach_inf=0.1)
    o = O.Oracle(**kw)
    u = o.ic_tgv(isothermal=True)
    u2 = u.copy()
    s0 = o.numerical_entropy(u)
    dt = o.constant_time_step(0.05)
    g, worst = 1.0, 0.0
    for _ in range(4):
        adv, g = o.step_rrk("rk4", dt, u, g)
        assert abs(adv - g * dt) <= 1e-15 * dt and abs(g - 1.0) < 1e-2
        worst = max(worst, abs(o.numerical_entropy(u) - s0) / abs(s0))
        o.step("rk4", dt, u2)
    drift_rk4 = abs(o.numerical_entropy(u2) - s0) / abs(s0)
    assert worst < 1e-13 and drift_rk4 > 1e3 * worst


def test_rrk_gamma_tends_to_one_with_dt():
    # gamma - 1 = O(dt^p): halving dt shrinks |gamma - 1| (RK4, p = 4) by far more than 2
    kw = dict(dim=3, poly_degree=2, n_cells_1d=3, use_split_form=1, two_point_flux=O.CH, conv_flux=O.TWO_POINT,
              mach_inf=0.3)
    o = O.Oracle(**kw)
    dev = []
    for dt in (0.02, 0.01):
        u = o.ic_tgv(isothermal=True)
        _, g = o.step_rrk("rk4", dt, u, 1.0)
        dev.append(abs(g - 1.0))
    assert dev[1] < dev[0] / 4
