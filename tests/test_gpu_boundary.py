"""Boundary faces on the device (Euler::boundary_face_values, euler.cpp:1511-1567, in
assemble_boundary_term_strong, strong_dg.cpp:1583-2364) against the FP64 oracle, through the
C ABI, on random states inside the physical range.

Bars: connectivity bit-exact; FP64 residual per DoF within 1e-12 of the DoF's term magnitude
(the bar of the periodic parity tests); stepped solutions within 1e-12 max-norm relative;
mixed precision (FP32 fluxes, boundary states in FP32) within 5e-4 relative, the bar of the
periodic mixed residual; a closed slip-wall box conserves mass and energy to rounding on the
device as the oracle's does."""
import sys

import numpy as np
import pytest

import oracle as O
import philip_amd as P

pytestmark = pytest.mark.gpu

TOL_RESIDUAL = 1e-12
TOL_STATE = 1e-12
GAM = 1.4
BCS = [1001, 1002, 1003, 1004, 1005, 1006, 1007]


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def dof_rel_err(o, a, b, mag):
    m = mag.reshape(o.ncells, o.ns, o.Np)
    den = np.maximum(m, 1e-8 * m.max(axis=(0, 2), keepdims=True))
    return float((np.abs(a - b).reshape(m.shape) / den).max())


def random_field(o, seed, mach, vel=0.5):
    rng = np.random.default_rng(seed)
    pinf = 1.0 / (GAM * mach * mach)
    u = o.vec().reshape(o.ncells, o.ns, o.Np)
    rho = rng.uniform(0.8, 1.2, (o.ncells, o.Np))
    v = rng.uniform(-vel, vel, (o.ncells, o.dim, o.Np))
    p = pinf * rng.uniform(0.8, 1.2, (o.ncells, o.Np))
    u[:, 0] = rho
    u[:, 1:1 + o.dim] = rho[:, None] * v
    u[:, -1] = p / (GAM - 1) + 0.5 * rho * (v * v).sum(axis=1)
    return u.ravel()


SPLIT = dict(use_split_form=1, two_point_flux=O.CH, conv_flux=O.TWO_POINT_ROE)
SCHEMES = {
    "lxf": dict(conv_flux=O.LXF),
    "roe": dict(conv_flux=O.ROE),
    "ch_roe": SPLIT,
    "ra_lxf": dict(use_split_form=1, two_point_flux=O.RA, conv_flux=O.TWO_POINT_LXF),
    "ir_ec": dict(use_split_form=1, two_point_flux=O.IR, conv_flux=O.TWO_POINT),
}


def kw_for(dim, p, bcs, scheme, n_cells=4, mach=0.5, aoa=0.3, **extra):
    return dict(dim=dim, poly_degree=p, n_cells_1d=n_cells, mach_inf=mach, angle_of_attack=aoa,
                domain_left=0.0, domain_right=1.0, boundary_type=bcs, **SCHEMES[scheme], **extra)


def test_connectivity_bit_exact():
    for dim, N, bcs, order in [(3, 4, [1006, 1006, 0, 0, 1004, 1004], O.ORDER_DEALII),
                               (3, 3, [1003, 1002, 1006, 1006, 0, 0], O.ORDER_DEALII),
                               (2, 8, [1005, 1005, 1007, 1007], O.ORDER_DEALII),
                               (2, 6, [0, 0, 1001, 1001], O.ORDER_LEXICOGRAPHIC)]:
        log(f"connectivity dim={dim} N={N} bcs={bcs}")
        kw = dict(dim=dim, poly_degree=1, n_cells_1d=N, boundary_type=bcs, cell_order=order)
        o, d = O.Oracle(**kw), P.DG(**kw)
        assert np.array_equal(d.neighbors(), o.neighbors())
        assert (d.neighbors() < 0).any()


def check(o, d, u, label):
    r_ref = o.residual(u)
    mag = o.residual_magnitude(u)
    d.set_solution(u)
    r = d.assemble_residual()
    e = dof_rel_err(o, r, r_ref, mag)
    el = dof_rel_err(o, d.apply_inverse_mass_to_rhs(), o.inverse_mass(r_ref), o.inverse_mass(mag))
    log(f"  {label}: per-DoF err {e:.2e}, lifted {el:.2e} (bar {TOL_RESIDUAL:.0e})")
    assert e <= TOL_RESIDUAL and el <= TOL_RESIDUAL, label
    return r


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("scheme", ["lxf", "ch_roe"])
@pytest.mark.parametrize("dim,p", [(2, 3), (3, 2), (3, 3)])
def test_residual_each_boundary_type(bc, scheme, dim, p):
    """Every type on every face of the box (subsonic free stream; the random state makes both
    subsonic and supersonic faces for the outflow and farfield branches)."""
    kw = kw_for(dim, p, [bc] * (2 * dim), scheme)
    log(f"boundary {bc} {scheme} dim={dim} p={p}")
    o, d = O.Oracle(**kw), P.DG(**kw)
    u = random_field(o, seed=bc + 10 * p + dim, mach=0.5, vel=2.0)
    r = check(o, d, u, f"{bc} {scheme} {dim}d p={p}")
    assert np.array_equal(r, d.assemble_residual())


@pytest.mark.parametrize("scheme", list(SCHEMES))
@pytest.mark.parametrize("dim,p", [(2, 2), (2, 5), (3, 1), (3, 4), (3, 5), (3, 6)])
def test_residual_mixed_boundaries(scheme, dim, p):
    """Different types per direction, one direction periodic in 3D; p=3..5 in 3D run the
    occupancy-bound FP64 line kernel, p=6 splits each line over two items."""
    bcs = [1003, 1002, 1006, 1006] if dim == 2 else [1004, 1007, 0, 0, 1001, 1005]
    kw = kw_for(dim, p, bcs, scheme)
    log(f"mixed boundaries {scheme} dim={dim} p={p} {bcs}")
    o, d = O.Oracle(**kw), P.DG(**kw)
    u = random_field(o, seed=50 + p + dim, mach=0.5)
    check(o, d, u, f"{scheme} {dim}d p={p} {bcs}")


def test_supersonic_free_stream_boundaries():
    for bcs in ([1003, 1002, 1004, 1004], [1004, 1004, 1005, 1005]):
        kw = kw_for(2, 3, bcs, "ch_roe", mach=1.6, aoa=0.2)
        log(f"supersonic {bcs}")
        o, d = O.Oracle(**kw), P.DG(**kw)
        u = random_field(o, seed=77, mach=1.6, vel=1.5)
        check(o, d, u, f"M=1.6 {bcs}")


@pytest.mark.parametrize("method", ["rk4", "rk4_3_5_3sstar", "rk3_2_5f_3sstarplus"])
@pytest.mark.parametrize("scheme", ["roe", "ch_roe"])
def test_time_stepping_with_boundaries(method, scheme):
    kw = kw_for(3, 3, [1004, 1004, 1006, 1006, 0, 0], scheme, mach=0.5)
    log(f"stepping {method} {scheme} with boundaries")
    o, d = O.Oracle(**kw), P.DG(**kw)
    u = random_field(o, seed=9, mach=0.5, vel=0.3)
    d.set_solution(u)
    dt = 0.002
    d.allocate_ode_system(method)
    d.step_in_time(dt, 3)
    for _ in range(3):
        o.step(method, dt, u)
    e = np.abs(d.solution() - u).max() / np.abs(u).max()
    log(f"  rel err {e:.3e} (bar {TOL_STATE:.0e})")
    assert e <= TOL_STATE


@pytest.mark.parametrize("scheme", ["lxf", "ch_roe", "ir_ec"])
def test_closed_slip_wall_box_conserves_mass_and_energy(scheme):
    kw = kw_for(3, 3, [1006] * 6, scheme, mach=0.4)
    o, d = O.Oracle(**kw), P.DG(**kw)
    u = random_field(o, seed=21, mach=0.4)
    d.set_solution(u)
    r = d.assemble_residual().reshape(o.ncells, o.ns, o.Np)
    tot = r.sum(axis=(0, 2))
    scale = np.abs(r).max() * r.size / o.ns
    log(f"  {scheme}: mass {tot[0] / scale:.1e} energy {tot[-1] / scale:.1e}")
    assert abs(tot[0]) / scale < 1e-14 and abs(tot[-1]) / scale < 1e-14


@pytest.mark.parametrize("dim,p", [(2, 3), (3, 3), (3, 4)])
def test_mixed_precision_with_boundaries(dim, p):
    bcs = [1003, 1002, 1006, 1006] if dim == 2 else [1004, 1004, 1006, 1006, 1005, 1005]
    for scheme in ("lxf", "ch_roe"):
        kw = kw_for(dim, p, bcs, scheme, mach=0.5)
        o = O.Oracle(**kw)
        d = P.DG(precision=P.MIXED, **kw)
        u = random_field(o, seed=31 + p, mach=0.5)
        d.set_solution(u)
        r, r_ref = d.assemble_residual(), o.residual(u)
        e = max(np.abs(r - r_ref).max() / np.abs(r_ref).max(),
                np.linalg.norm(r - r_ref) / np.linalg.norm(r_ref))
        log(f"  mixed {scheme} {dim}d p={p}: rel err {e:.2e} (bar 5e-4)")
        assert e <= 5e-4


def test_invalid_boundaries_raise():
    with pytest.raises(P.DeviceError):
        P.DG(dim=2, poly_degree=2, boundary_type=[1006, 0, 0, 0])
    with pytest.raises(P.DeviceError):
        P.DG(dim=2, poly_degree=2, boundary_type=[1000, 1000, 0, 0])
    with pytest.raises(P.DeviceError):
        P.DG(dim=3, poly_degree=2, pde=P.NAVIER_STOKES, boundary_type=[1004, 1004, 0, 0, 0, 0])
    with pytest.raises(P.DeviceError):
        P.DG(dim=3, poly_degree=2, pde=P.NAVIER_STOKES, boundary_type=[1001, 1001, 0, 0, 0, 0], thermal_bc=2)
