"""Navier-Stokes walls on the device (no-slip 1001, adiabatic or isothermal, and slip 1006; the
lifted gradient's wall term and the viscous wall flux of stage2_kernel) against the FP64 oracle,
through the C ABI.

Bars: the lifted gradient and the FP64 residual per DoF within 1e-12 (the bar of the periodic
Navier-Stokes parity tests); stepped solutions within 1e-12 max-norm relative; the mixed path
within 5e-4 relative (its periodic bar)."""
import sys

import numpy as np
import pytest

import oracle as O
import philip_amd as P

pytestmark = pytest.mark.gpu

GAM = 1.4
MACH = 0.3
PINF = 1.0 / (GAM * MACH * MACH)
TOL = 1e-12


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def kw_for(bcs, thermal=0, tw=1.0, split=1, diss=O.SIPG, p=3, n=3, **extra):
    sch = dict(use_split_form=1, two_point_flux=O.RA, conv_flux=O.TWO_POINT_ROE) if split else dict(conv_flux=O.ROE)
    return dict(dim=3, poly_degree=p, n_cells_1d=n, pde=O.NAVIER_STOKES, diss_flux=diss, reynolds_inf=100.0,
                prandtl=0.71, mach_inf=MACH, boundary_type=bcs, thermal_bc=thermal, wall_temperature=tw,
                domain_left=0.0, domain_right=1.0, **sch, **extra)


def random_field(o, seed):
    rng = np.random.default_rng(seed)
    u = o.vec().reshape(o.ncells, o.ns, o.Np)
    rho = rng.uniform(0.8, 1.2, (o.ncells, o.Np))
    v = rng.uniform(-0.3, 0.3, (o.ncells, 3, o.Np))
    p = PINF * rng.uniform(0.9, 1.1, (o.ncells, o.Np))
    u[:, 0] = rho
    u[:, 1:4] = rho[:, None] * v
    u[:, 4] = p / (GAM - 1) + 0.5 * rho * (v * v).sum(axis=1)
    return u.ravel()


def dof_rel_err(o, a, b, mag):
    m = mag.reshape(o.ncells, o.ns, o.Np)
    den = np.maximum(m, 1e-8 * m.max(axis=(0, 2), keepdims=True))
    return float((np.abs(a - b).reshape(m.shape) / den).max())


WALLS = [("noslip_adiabatic", [1001] * 6, 0, 1.0), ("noslip_isothermal", [1001] * 6, 1, 1.2),
         ("slip", [1006] * 6, 0, 1.0), ("channel", [0, 0, 1001, 1001, 1006, 1006], 1, 0.9)]


@pytest.mark.parametrize("name,bcs,thermal,tw", WALLS, ids=[w[0] for w in WALLS])
@pytest.mark.parametrize("split", [0, 1])
@pytest.mark.parametrize("diss", [O.SIPG, O.BR2])
@pytest.mark.parametrize("p", [2, 3, 5])
def test_residual_with_walls(name, bcs, thermal, tw, split, diss, p):
    kw = kw_for(bcs, thermal, tw, split, diss, p=p, n=3 if p < 5 else 2)
    log(f"NS walls {name} split={split} diss={diss} p={p}")
    o, d = O.Oracle(**kw), P.DG(**kw)
    u = random_field(o, seed=7 + p)
    r_ref = o.residual(u)
    d.set_solution(u)
    r = d.assemble_residual()
    for k in range(3):
        q, q_ref = d.auxiliary_solution(k), o.auxiliary(k)
        assert np.abs(q - q_ref).max() <= TOL * np.abs(q_ref).max()
    e = dof_rel_err(o, r, r_ref, o.residual_magnitude(u))
    log(f"  per-DoF err {e:.2e} (bar {TOL:.0e})")
    assert e <= TOL


@pytest.mark.parametrize("split", [0, 1])
def test_time_stepping_with_walls(split):
    kw = kw_for([1001, 1001, 1006, 1006, 1001, 1001], thermal=1, tw=1.1, split=split)
    o, d = O.Oracle(**kw), P.DG(**kw)
    u = random_field(o, seed=3)
    d.set_solution(u)
    dt = 2e-4
    d.allocate_ode_system("rk4")
    d.step_in_time(dt, 3)
    for _ in range(3):
        o.step("rk4", dt, u)
    e = np.abs(d.solution() - u).max() / np.abs(u).max()
    log(f"  stepping split={split}: rel err {e:.2e}")
    assert e <= TOL


def test_mixed_precision_with_walls():
    for split in (0, 1):
        kw = kw_for([1001] * 6, split=split, p=3)
        o = O.Oracle(**kw)
        d = P.DG(precision=P.MIXED, **kw)
        u = random_field(o, seed=11)
        d.set_solution(u)
        r, r_ref = d.assemble_residual(), o.residual(u)
        e = max(np.abs(r - r_ref).max() / np.abs(r_ref).max(), np.linalg.norm(r - r_ref) / np.linalg.norm(r_ref))
        log(f"  mixed split={split}: rel err {e:.2e} (bar 5e-4)")
        assert e <= 5e-4
