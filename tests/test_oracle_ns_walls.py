"""Navier-Stokes walls in the oracle (NavierStokes::boundary_face_values_viscous_flux,
navier_stokes.cpp:1354-1430, the adiabatic flux of navier_stokes.cpp:1266-1310, and the viscous
parts of assemble_boundary_term_strong / _auxiliary_equation, strong_dg.cpp:790-870, 2276-2300),
pinned by what walls do to the conserved totals of a closed box:

* a quiescent uniform gas in a closed no-slip box is at rest (residual zero);
* mass is conserved by every wall;
* a gas at rest with a temperature field keeps its energy behind adiabatic walls, gains energy
  from hotter isothermal walls and loses it to colder ones;
* a uniform stream along slip walls keeps its momentum (and is a steady state); along no-slip
  walls it is slowed down (the wall shear opposes the flow)."""
import numpy as np
import pytest

import oracle as O

GAM = 1.4
MACH = 0.3
PINF = 1.0 / (GAM * MACH * MACH)


def ns(bcs, thermal=0, tw=1.0, split=1, diss=O.SIPG, p=2, n=3):
    kw = dict(use_split_form=1, two_point_flux=O.RA, conv_flux=O.TWO_POINT_ROE) if split else dict(conv_flux=O.ROE)
    return O.Oracle(dim=3, poly_degree=p, n_cells_1d=n, pde=O.NAVIER_STOKES, diss_flux=diss, reynolds_inf=100.0,
                    prandtl=0.71, mach_inf=MACH, boundary_type=bcs, thermal_bc=thermal, wall_temperature=tw,
                    domain_left=0.0, domain_right=1.0, **kw)


def state(o, rho, vel, p):
    u = o.vec().reshape(o.ncells, o.ns, o.Np)
    rho = np.broadcast_to(rho, (o.ncells, o.Np))
    u[:, 0] = rho
    v2 = 0.0
    for k in range(3):
        u[:, 1 + k] = rho * vel[k]
        v2 = v2 + np.asarray(vel[k]) ** 2
    u[:, 4] = p / (GAM - 1) + 0.5 * rho * v2
    return u.ravel()


def totals(o, u):
    r = o.residual(u).reshape(o.ncells, o.ns, o.Np)
    return r.sum(axis=(0, 2)), np.abs(r).max() * r.size / o.ns


CASES = [(s, d) for s in (0, 1) for d in (O.SIPG, O.BR2)]


@pytest.mark.parametrize("split,diss", CASES)
def test_quiescent_gas_in_a_closed_box_is_at_rest(split, diss):
    for thermal in (0, 1):
        o = ns([1001] * 6, thermal=thermal, split=split, diss=diss)
        u = state(o, 1.0, (0.0, 0.0, 0.0), PINF)   # T = gamma M^2 p / rho = 1 = the wall temperature
        assert np.abs(o.residual(u)).max() < 1e-13 * PINF


@pytest.mark.parametrize("split,diss", CASES)
@pytest.mark.parametrize("bcs", [[1001] * 6, [1006] * 6, [1001, 1001, 1006, 1006, 0, 0]])
def test_walls_conserve_mass(split, diss, bcs):
    o = ns(bcs, split=split, diss=diss)
    rng = np.random.default_rng(1)
    shape = (o.ncells, o.Np)
    u = state(o, rng.uniform(0.8, 1.2, shape), [rng.uniform(-0.3, 0.3, shape) for _ in range(3)],
              PINF * rng.uniform(0.9, 1.1, shape))
    tot, scale = totals(o, u)
    assert abs(tot[0]) < 1e-14 * scale


@pytest.mark.parametrize("split,diss", CASES)
def test_heat_through_walls_at_rest(split, diss):
    o = ns([1001] * 6, split=split, diss=diss)
    x = o.node_coords()
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * x[..., 0]) * np.cos(np.pi * x[..., 1])   # T = 1/rho at p_inf
    u = state(o, rho, (0.0, 0.0, 0.0), PINF)
    tot, scale = totals(o, u)
    assert abs(tot[4]) < 1e-14 * scale           # adiabatic: no heat crosses the walls
    hot = totals(ns([1001] * 6, thermal=1, tw=1.3, split=split, diss=diss), u)[0][4]
    cold = totals(ns([1001] * 6, thermal=1, tw=0.7, split=split, diss=diss), u)[0][4]
    assert hot > 1e-3 * scale and cold < -1e-3 * scale


@pytest.mark.parametrize("split,diss", CASES)
def test_stream_along_walls(split, diss):
    # walls on the y faces, periodic in x and z, a uniform stream in x
    o_slip = ns([0, 0, 1006, 1006, 0, 0], split=split, diss=diss)
    u = state(o_slip, 1.0, (0.4, 0.0, 0.0), PINF)
    r = o_slip.residual(u)
    assert np.abs(r).max() < 1e-13 * PINF        # slip walls: a steady state
    o_ns = ns([0, 0, 1001, 1001, 0, 0], split=split, diss=diss)
    tot, scale = totals(o_ns, u)
    assert abs(tot[0]) < 1e-14 * scale
    assert tot[1] < -1e-3 * scale                # no-slip walls slow the stream down
    assert abs(tot[2]) < 1e-12 * scale           # symmetric walls: no net normal force
