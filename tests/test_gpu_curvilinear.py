"""Curvilinear grids on the device against the FP64 oracle, through the C ABI: the reference's
curved periodic grid (NonsymmetricCurvedGridManifold, beta = 0.1) built on the device, caller-
supplied node coordinates, metric terms (the 3D conservative curl form), and the curvilinear split
and conservative forms with their face terms (stage2_kernel with the metric, strong_dg.cpp:1267-1380,
3162-3350), and Navier-Stokes on the curved grid (the lifted gradient through the metric,
strong_dg.cpp:700-870; viscous fluxes contracted with the cofactor and the face fluxes along the
physical normal with each cell's penalty from its extent, dg_base.cpp:397-413).

Bars: node coordinates within 1e-14 and det J within 1e-12 relative, the cofactor within 1e-11
(the curl form's cancellation at p = 6); FP64 residual per DoF within 1e-12
of the DoF's term magnitude (the bar of the Cartesian parity tests); stepped solutions within
1e-12 max-norm relative; free-stream preservation at the reference's 1e-13 (L2 of the residual);
semi-discrete entropy conservation at the oracle's 1e-12; the mixed path at its 5e-4 bar."""
import sys

import numpy as np
import pytest

import oracle as O
import philip_amd as P

pytestmark = pytest.mark.gpu

GAM = 1.4
TOL = 1e-12
BOX = {2: (-1.0, 1.0), 3: (0.0, 2 * np.pi)}
SCHEMES = {
    "lxf": dict(conv_flux=O.LXF),
    "roe": dict(conv_flux=O.ROE),
    "central": dict(conv_flux=O.CENTRAL),
    "ch_roe": dict(use_split_form=1, two_point_flux=O.CH, conv_flux=O.TWO_POINT_ROE),
    "ra_lxf": dict(use_split_form=1, two_point_flux=O.RA, conv_flux=O.TWO_POINT_LXF),
    "ir_ec": dict(use_split_form=1, two_point_flux=O.IR, conv_flux=O.TWO_POINT),
    "kg_roe": dict(use_split_form=1, two_point_flux=O.KG, conv_flux=O.TWO_POINT_ROE),
}


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def kw_for(dim, p, scheme, n_cells=None, warp=1, **extra):
    lo, hi = BOX[dim]
    return dict(dim=dim, poly_degree=p, n_cells_1d=n_cells or (4 if dim == 2 else 3), domain_left=lo,
                domain_right=hi, grid_warp=warp, mach_inf=0.3, **SCHEMES[scheme], **extra)


def random_field(o, seed, mach=0.3, vel=0.4):
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


def dof_rel_err(o, a, b, mag):
    m = mag.reshape(o.ncells, o.ns, o.Np)
    den = np.maximum(m, 1e-8 * m.max(axis=(0, 2), keepdims=True))
    return float((np.abs(a - b).reshape(m.shape) / den).max())


def check(o, d, u, label):
    r_ref = o.residual(u)
    mag = o.residual_magnitude(u)
    d.set_solution(u)
    r = d.assemble_residual()
    e = dof_rel_err(o, r, r_ref, mag)
    el = dof_rel_err(o, d.apply_inverse_mass_to_rhs(), o.inverse_mass(r_ref), o.inverse_mass(mag))
    log(f"  {label}: per-DoF err {e:.2e}, lifted {el:.2e} (bar {TOL:.0e})")
    assert e <= TOL and el <= TOL, label
    return r


@pytest.mark.parametrize("dim,p", [(2, 2), (2, 4), (3, 1), (3, 3), (3, 6)])
def test_grid_and_metric_terms_match_oracle(dim, p):
    kw = kw_for(dim, p, "lxf")
    o, d = O.Oracle(**kw), P.DG(**kw)
    x, xo = d.grid_nodes(), o.node_coords()
    assert np.abs(x - xo).max() <= 1e-14 * np.abs(xo).max()
    C, J = d.metric_terms()
    Co, Jo = o.metric_terms()
    ec = np.abs(C - Co).max() / np.abs(Co).max()
    ej = np.abs(J - Jo).max() / np.abs(Jo).max()
    log(f"  metric {dim}d p={p}: cofactor {ec:.1e}, det J {ej:.1e}")
    # the 3D curl form differentiates products X_l dX_m/dxi of size |X| |J| (|X| up to 2 pi) with
    # derivative rows whose absolute sums reach ~50 at p = 6: its rounding, relative to max |C|,
    # is bounded by 1e-11 there (the oracle's own affine-grid check sees 4e-13 at p = 3)
    assert ec <= 1e-11 and ej <= TOL


@pytest.mark.parametrize("scheme", list(SCHEMES))
@pytest.mark.parametrize("dim,p", [(2, 2), (2, 3), (2, 5), (3, 1), (3, 2), (3, 3), (3, 4), (3, 5), (3, 6)])
def test_residual_on_the_curved_grid(scheme, dim, p):
    kw = kw_for(dim, p, scheme)
    log(f"curved grid {scheme} {dim}d p={p}")
    o, d = O.Oracle(**kw), P.DG(**kw)
    u = random_field(o, seed=10 * p + dim)
    r = check(o, d, u, f"{scheme} {dim}d p={p}")
    assert np.array_equal(r, d.assemble_residual())


def quadratic_map(dim, x):
    y = x.copy()
    if dim == 2:
        y[..., 0] = x[..., 0] + 0.05 * x[..., 0] * x[..., 1]
        y[..., 1] = x[..., 1] + 0.04 * x[..., 0] ** 2
    else:
        y[..., 0] = x[..., 0] + 0.03 * x[..., 1] * x[..., 2]
        y[..., 1] = x[..., 1] + 0.02 * x[..., 0] * x[..., 2]
        y[..., 2] = x[..., 2] + 0.025 * x[..., 0] * x[..., 1]
    return y


@pytest.mark.parametrize("scheme", ["roe", "ch_roe"])
@pytest.mark.parametrize("dim", [2, 3])
def test_caller_node_coordinates(scheme, dim):
    kw = kw_for(dim, 3, scheme, warp=0)
    kw.update(domain_left=0.0, domain_right=1.0)
    o, d = O.Oracle(**kw), P.DG(**kw)
    x = quadratic_map(dim, o.node_coords())
    o.set_grid_nodes(x)
    d.set_grid_nodes(x)
    assert np.array_equal(d.grid_nodes(), x)
    check(o, d, random_field(o, seed=3), f"caller grid {scheme} {dim}d")


@pytest.mark.parametrize("scheme", ["lxf", "ch_roe"])
def test_boundaries_on_the_curved_grid(scheme):
    # the 2D curved grid keeps the box sides straight (its nodes slide along them): slip walls on
    # y, a characteristic farfield on x
    kw = kw_for(2, 3, scheme, boundary_type=[1004, 1004, 1006, 1006])
    o, d = O.Oracle(**kw), P.DG(**kw)
    check(o, d, random_field(o, seed=5), f"curved grid with boundaries {scheme}")


@pytest.mark.parametrize("btype", [1001, 1002, 1003, 1004, 1005, 1006, 1007])
@pytest.mark.parametrize("scheme", ["roe", "ra_lxf"])
@pytest.mark.parametrize("dim", [2, 3])
def test_every_boundary_type_on_curved_caller_nodes(btype, scheme, dim):
    # every face of the unit box bent by the quadratic map, all of one type (Euler)
    kw = kw_for(dim, 3, scheme, warp=0, boundary_type=[btype] * (2 * dim))
    kw.update(domain_left=0.0, domain_right=1.0)
    o, d = O.Oracle(**kw), P.DG(**kw)
    x = quadratic_map(dim, o.node_coords())
    o.set_grid_nodes(x)
    d.set_grid_nodes(x)
    check(o, d, random_field(o, seed=btype, vel=0.3), f"curved caller nodes {dim}d, all faces {btype}, {scheme}")


@pytest.mark.parametrize("method", ["rk4", "rk4_3_5_3sstar", "rk3_2_5f_3sstarplus"])
@pytest.mark.parametrize("scheme", ["roe", "ch_roe"])
def test_time_stepping_on_the_curved_grid(method, scheme):
    kw = kw_for(3, 3, scheme)
    o, d = O.Oracle(**kw), P.DG(**kw)
    u = random_field(o, seed=9, vel=0.2)
    d.set_solution(u)
    dt = 0.002
    d.allocate_ode_system(method)
    d.step_in_time(dt, 3)
    for _ in range(3):
        o.step(method, dt, u)
    e = np.abs(d.solution() - u).max() / np.abs(u).max()
    log(f"  stepping {method} {scheme}: rel err {e:.2e} (bar {TOL:.0e})")
    assert e <= TOL


@pytest.mark.parametrize("scheme", ["lxf", "roe", "ch_roe", "ir_ec", "kg_roe"])
@pytest.mark.parametrize("dim", [2, 3])
def test_free_stream_preserved_on_device(scheme, dim):
    kw = kw_for(dim, 4, scheme)
    d = P.DG(**kw)
    uinf = np.r_[1.2, 1.2 * np.array([0.3, -0.2, 0.1][:dim]), 0.0]
    uinf[-1] = 1.0 / (GAM * 0.09) / (GAM - 1) + 0.5 * 1.2 * np.sum(np.array([0.3, -0.2, 0.1][:dim]) ** 2)
    ncells, npn = d.ncells, d.ndofs // (d.ncells * (dim + 2))
    u = np.tile(uinf[:, None], (ncells, 1, npn)).ravel()
    d.set_solution(u)
    r = d.assemble_residual()
    l2 = np.linalg.norm(r) / np.sqrt(r.size)
    log(f"  free stream {scheme} {dim}d: residual rms {l2:.1e}")
    assert l2 < 1e-13 * uinf[-1]


@pytest.mark.parametrize("tp", [O.IR, O.CH, O.RA])
def test_entropy_conservation_on_device(tp):
    kw = dict(kw_for(3, 3, "ir_ec"), two_point_flux=tp)
    o, d = O.Oracle(**kw), P.DG(**kw)
    u = random_field(o, seed=4)
    d.set_solution(u)
    r = d.assemble_residual()
    v = np.concatenate([o.entropy_variables(x) for x in u.reshape(o.ncells, o.ns, o.Np).transpose(0, 2, 1).reshape(-1, o.ns)])
    v = v.reshape(o.ncells, o.Np, o.ns).transpose(0, 2, 1).ravel()
    rate = abs(v @ r) / (np.abs(r).sum() * np.abs(v).max())
    log(f"  entropy rate {tp}: {rate:.1e}")
    assert rate < 1e-12


def test_numerical_entropy_and_relaxation_on_the_curved_grid():
    kw = kw_for(3, 3, "ir_ec")
    o, d = O.Oracle(**kw), P.DG(**kw)
    u = random_field(o, seed=6, vel=0.2)
    d.set_solution(u)
    s_dev, s_ref = d.numerical_entropy(), o.numerical_entropy(u)
    assert abs(s_dev - s_ref) <= 1e-13 * abs(s_ref)
    # relaxation RK keeps the entropy of the EC scheme to the root tolerance over a few steps
    d.allocate_ode_system("rk4")
    d.set_relaxation_runge_kutta(True)
    d.step_in_time(0.002, 3)
    s1 = d.numerical_entropy()
    log(f"  RRK entropy change {abs(s1 - s_dev) / abs(s_dev):.1e}")
    assert abs(s1 - s_dev) <= 1e-12 * abs(s_dev)


@pytest.mark.parametrize("dim,p", [(2, 3), (3, 3), (3, 5)])
def test_mixed_precision_on_the_curved_grid(dim, p):
    for scheme in ("lxf", "ch_roe"):
        kw = kw_for(dim, p, scheme)
        o = O.Oracle(**kw)
        d = P.DG(precision=P.MIXED, **kw)
        u = random_field(o, seed=31 + p)
        d.set_solution(u)
        r, r_ref = d.assemble_residual(), o.residual(u)
        e = max(np.abs(r - r_ref).max() / np.abs(r_ref).max(), np.linalg.norm(r - r_ref) / np.linalg.norm(r_ref))
        log(f"  mixed {scheme} {dim}d p={p}: rel err {e:.2e} (bar 5e-4)")
        assert e <= 5e-4


def test_unsupported_combinations_raise():
    with pytest.raises(P.DeviceError):
        P.DG(dim=3, poly_degree=2, grid_warp=2)
    plain = P.DG(dim=3, poly_degree=2)
    with pytest.raises(P.DeviceError):
        plain.metric_terms()


def test_initial_condition_at_the_curved_nodes():
    kw = kw_for(3, 3, "ch_roe", n_cells=4)
    o, d = O.Oracle(**kw), P.DG(**kw)
    d.set_initial_condition_tgv()
    u, uo = d.solution(), o.ic_tgv()
    assert np.abs(u - uo).max() <= 1e-14 * np.abs(uo).max()
    # the curved nodes, not the box's: the state differs from the affine grid's
    assert np.abs(uo - O.Oracle(**dict(kw, grid_warp=0)).ic_tgv()).max() > 1e-3


# ---------------------------------------------------------------------------------------------
# Navier-Stokes on curvilinear grids
# ---------------------------------------------------------------------------------------------

def ns_kw(split, diss, p, n=3, **extra):
    sch = dict(use_split_form=1, two_point_flux=O.RA, conv_flux=O.TWO_POINT_ROE) if split else dict(conv_flux=O.ROE)
    kw = dict(dim=3, poly_degree=p, n_cells_1d=n, pde=O.NAVIER_STOKES, diss_flux=diss, reynolds_inf=100.0,
              prandtl=0.71, mach_inf=0.3, domain_left=0.0, domain_right=2 * np.pi, grid_warp=1, **sch)
    kw.update(extra)
    return kw


def check_ns(o, d, u, label):
    r = check(o, d, u, label)
    for k in range(3):
        q, q_ref = d.auxiliary_solution(k), o.auxiliary(k)
        eq = np.abs(q - q_ref).max() / np.abs(q_ref).max()
        assert eq <= TOL, (label, k, eq)
    return r


@pytest.mark.parametrize("split", [0, 1])
@pytest.mark.parametrize("diss", [O.SIPG, O.BR2])
@pytest.mark.parametrize("p", [2, 3, 5])
def test_navier_stokes_on_the_curved_grid(split, diss, p):
    kw = ns_kw(split, diss, p, n=3 if p < 5 else 2)
    log(f"curved grid NS split={split} diss={diss} p={p}")
    o, d = O.Oracle(**kw), P.DG(**kw)
    r = check_ns(o, d, random_field(o, seed=40 + p, vel=0.3), f"NS split={split} diss={diss} p={p}")
    assert np.array_equal(r, d.assemble_residual())


WALL_SETS = [("noslip_adiabatic", [1001] * 6, 0, 1.0), ("noslip_isothermal", [1001] * 6, 1, 1.2),
             ("channel", [0, 0, 1001, 1001, 1006, 1006], 1, 0.9)]


@pytest.mark.parametrize("name,bcs,thermal,tw", WALL_SETS, ids=[w[0] for w in WALL_SETS])
@pytest.mark.parametrize("split", [0, 1])
@pytest.mark.parametrize("diss", [O.SIPG, O.BR2])
def test_navier_stokes_walls_on_caller_nodes(name, bcs, thermal, tw, split, diss):
    # curved walls: the quadratic map bends every face of the unit box
    kw = ns_kw(split, diss, 3, grid_warp=0, domain_right=1.0, boundary_type=bcs, thermal_bc=thermal,
               wall_temperature=tw)
    log(f"curved walls {name} split={split} diss={diss}")
    o, d = O.Oracle(**kw), P.DG(**kw)
    x = quadratic_map(3, o.node_coords())
    o.set_grid_nodes(x)
    d.set_grid_nodes(x)
    check_ns(o, d, random_field(o, seed=6, vel=0.3), f"curved walls {name} split={split}")


@pytest.mark.parametrize("split", [0, 1])
def test_navier_stokes_stepping_on_the_curved_grid(split):
    kw = ns_kw(split, O.SIPG, 3)
    o, d = O.Oracle(**kw), P.DG(**kw)
    u = random_field(o, seed=12, vel=0.2)
    d.set_solution(u)
    dt = 1e-3
    d.allocate_ode_system("rk4")
    d.step_in_time(dt, 3)
    for _ in range(3):
        o.step("rk4", dt, u)
    e = np.abs(d.solution() - u).max() / np.abs(u).max()
    log(f"  NS stepping split={split}: rel err {e:.2e} (bar {TOL:.0e})")
    assert e <= TOL


def test_navier_stokes_mixed_precision_on_the_curved_grid():
    for split in (0, 1):
        kw = ns_kw(split, O.SIPG, 3)
        o = O.Oracle(**kw)
        d = P.DG(precision=P.MIXED, **kw)
        u = random_field(o, seed=13, vel=0.3)
        d.set_solution(u)
        r, r_ref = d.assemble_residual(), o.residual(u)
        e = max(np.abs(r - r_ref).max() / np.abs(r_ref).max(), np.linalg.norm(r - r_ref) / np.linalg.norm(r_ref))
        log(f"  NS mixed split={split}: rel err {e:.2e} (bar 5e-4)")
        assert e <= 5e-4


# ---------------------------------------------------------------------------------------------
# Integrated quantities on curved cells
# ---------------------------------------------------------------------------------------------

def iq_vec(q):
    return np.array([q["kinetic_energy"], q["enstrophy"], q["incompressible_kinetic_energy"],
                     q["vorticity_dissipation"]])


@pytest.mark.parametrize("dim,p", [(2, 2), (2, 5), (3, 2), (3, 3), (3, 5), (3, 6)])
def test_integrated_quantities_on_the_curved_grid(dim, p):
    kw = kw_for(dim, p, "ch_roe", pde=O.NAVIER_STOKES if dim == 3 else O.EULER, reynolds_inf=1600.0)
    o, d = O.Oracle(**kw), P.DG(**kw)
    u = random_field(o, seed=50 + p, vel=0.3)
    d.set_solution(u)
    for overint in (0, 10):
        got, ref = iq_vec(d.integrated(overint)), iq_vec(o.integrated(u, overint))
        e = np.abs(got - ref) / np.abs(ref)
        log(f"  integrated {dim}d p={p} overint={overint}: rel err {e.max():.2e} (bar 1e-12)")
        # the 3D cofactor at the points comes from the curl form's cancellation, as at the nodes
        assert e.max() <= 1e-12


def test_integrated_quantities_on_caller_nodes_and_allocation_free():
    from hipmem import DeviceBuffer
    kw = kw_for(3, 3, "ch_roe", warp=0, pde=O.NAVIER_STOKES, reynolds_inf=1600.0)
    kw.update(domain_left=0.0, domain_right=1.0)
    o, d = O.Oracle(**kw), P.DG(**kw)
    x = quadratic_map(3, o.node_coords())
    o.set_grid_nodes(x)
    d.set_grid_nodes(x)
    u = random_field(o, seed=8, vel=0.3)
    d.set_solution(u)
    got, ref = iq_vec(d.integrated(10)), iq_vec(o.integrated(u, 10))
    assert (np.abs(got - ref) / np.abs(ref)).max() <= 1e-12
    # the allocation-free entry point runs the same kernels on curved cells
    lib = P.load()
    out = DeviceBuffer.from_array(np.full(4, -1.0))
    assert lib.dgx_integrated_quantities_static(d.h, None, 10, out.value) == 0, lib.dgx_last_error(d.h)
    d.synchronize()
    assert np.array_equal(out.download(), got)
