"""Whole-residual properties of the oracle: free-stream preservation
(reference freestream_preservation.cpp:262,283, 1e-13), discrete conservation
of mass/momentum/energy on the periodic cube, semi-discrete entropy
conservation of the split forms with an entropy-conserving surface flux
(reference euler_entropy_conserving_split_forms_check.cpp:24), and the order
of the Runge-Kutta tableaus on u' = lambda u (closed form)."""
import numpy as np
import pytest

import oracle as O

SCHEMES = [
    ("lxf", dict(conv_flux=O.LXF)),
    ("roe", dict(conv_flux=O.ROE)),
    ("ch_roe", dict(use_split_form=1, two_point_flux=O.CH, conv_flux=O.TWO_POINT_ROE)),
    ("ra_lxf", dict(use_split_form=1, two_point_flux=O.RA, conv_flux=O.TWO_POINT_LXF)),
    ("ir_ec", dict(use_split_form=1, two_point_flux=O.IR, conv_flux=O.TWO_POINT)),
    ("kg_ec", dict(use_split_form=1, two_point_flux=O.KG, conv_flux=O.TWO_POINT)),
]


def conserved_totals(o, rhs):
    # sum over nodes of W_i det(J) du_i/dt  =  sum of the un-lifted residual (chi = I, GLL)
    r = rhs.reshape(o.ncells, o.ns, o.Np)
    return r.sum(axis=(0, 2))


@pytest.mark.parametrize("name,kw", SCHEMES, ids=[s[0] for s in SCHEMES])
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("nodes", [O.GLL, O.GL])
def test_freestream_preservation(name, kw, dim, nodes):
    o = O.Oracle(dim=dim, poly_degree=3, n_cells_1d=4, flux_nodes=nodes, mach_inf=0.3, **kw)
    u = o.vec().reshape(o.ncells, o.ns, o.Np)
    vel = np.array([0.3, -0.2, 0.1])[:dim]
    rho, p = 1.2, 1.0 / (1.4 * 0.09)
    u[:, 0] = rho
    for k in range(dim):
        u[:, 1 + k] = rho * vel[k]
    u[:, -1] = p / 0.4 + 0.5 * rho * vel @ vel
    r = o.residual(u.ravel())
    assert np.abs(r).max() < 1e-13 * p


@pytest.mark.parametrize("name,kw", SCHEMES, ids=[s[0] for s in SCHEMES])
def test_conservation_tgv(name, kw):
    o = O.Oracle(dim=3, poly_degree=3, n_cells_1d=4, **kw)
    u = o.ic_tgv()
    r = o.residual(u)
    tot = conserved_totals(o, r)
    scale = np.abs(r).max() * r.size / o.ns
    assert np.abs(tot).max() / scale < 1e-14


def test_conservation_navier_stokes():
    o = O.Oracle(dim=3, poly_degree=2, n_cells_1d=4, pde=O.NAVIER_STOKES, conv_flux=O.ROE,
                 diss_flux=O.BR2, use_constant_viscosity=1, reynolds_inf=100.0, prandtl=0.71)
    u = o.ic_tgv()
    r = o.residual(u)
    tot = conserved_totals(o, r)
    assert np.abs(tot).max() / (np.abs(r).max() * r.size / o.ns) < 1e-14


@pytest.mark.parametrize("tp", [O.IR, O.CH, O.RA])
def test_semidiscrete_entropy_conservation(tp):
    # isothermal TGV (the reference's entropy-check initial condition), p=3, 4^3, GLL
    o = O.Oracle(dim=3, poly_degree=3, n_cells_1d=4, use_split_form=1, two_point_flux=tp,
                 conv_flux=O.TWO_POINT)
    u = o.ic_tgv(isothermal=True)
    rate = o.entropy_rate(u)
    r = o.residual(u)
    v = np.concatenate([o.entropy_variables(x) for x in u.reshape(o.ncells, o.ns, o.Np).transpose(0, 2, 1).reshape(-1, o.ns)])
    scale = np.abs(r).sum() * np.abs(v).max()
    assert abs(rate) / scale < 1e-12
    # an entropy-stable (dissipative) surface flux must produce entropy decay (rate <= 0 in v.R)
    o2 = O.Oracle(dim=3, poly_degree=3, n_cells_1d=4, use_split_form=1, two_point_flux=tp,
                  conv_flux=O.TWO_POINT_ROE)
    assert o2.entropy_rate(u) < 0


@pytest.mark.parametrize("method,order", [("euler", 1), ("heun", 2), ("ssprk3", 3), ("rk4", 4),
                                          ("rk4_3_5_3sstar", 4), ("rk3_2_5f_3sstarplus", 3),
                                          ("rk4_3_9f_3sstarplus", 4), ("rk5_4_10f_3sstarplus", 5)])
def test_rk_order_on_linear_advection(method, order):
    # smooth linear-ish problem: tiny-amplitude acoustic wave; compare against a much finer step
    o = O.Oracle(dim=2, poly_degree=2, n_cells_1d=2, conv_flux=O.CENTRAL, mach_inf=0.5)
    u0 = o.ic_vortex(strength=0.5)
    T = 0.2
    def run(nsteps):
        u = u0.copy()
        for _ in range(nsteps):
            o.step(method, T / nsteps, u)
        return u
    ref = run(256)
    n1 = 2 if order >= 5 else 8  # stay above round-off for the fifth-order scheme
    e1 = np.abs(run(n1) - ref).max()
    e2 = np.abs(run(2 * n1) - ref).max()
    rate = np.log2(e1 / e2)
    assert rate > order - 0.35


def test_dt_rule_matches_reference_formula():
    o = O.Oracle(dim=3, poly_degree=2, n_cells_1d=4)
    # periodic_turbulence.cpp:115-125 : CFL * L / (n_dofs/nstate)^(1/dim)
    assert o.constant_time_step(0.003) == 0.003 * 6.28318530717958623200 / pow(1728, 1.0 / 3)
