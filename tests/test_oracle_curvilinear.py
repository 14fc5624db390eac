"""Curvilinear grids in the oracle (metric_operators, operators.cpp:2520-2884; the curvilinear
split form of strong_dg.cpp:1267-1380 and the facet terms :3162-3350; the collocated inverse
mass with det J), pinned by what the mathematics fixes:

* on the affine box given as node coordinates, the curvilinear path reproduces the Cartesian
  residual (which the reference's golden vectors pin);
* for a polynomial mapping whose metric products the nodes resolve (degree 2 at p = 3), the
  conservative curl form returns the exact cofactor matrix det(J) J^-T and det J;
* the discrete metric identities sum_n D_n C[m][n] = 0 hold on the reference's curved periodic
  grid, and with them free-stream preservation, discrete conservation and, for the entropy
  conserving split forms, semi-discrete entropy conservation;
* rotating the grid and the momentum rotates the residual (a transposed cofactor or normal
  fails this);
* the lifted residual of a smooth state converges to -div F(u) at the nodes under refinement."""
import numpy as np
import pytest

import oracle as O

GAM = 1.4
SCHEMES = [
    ("lxf", dict(conv_flux=O.LXF)),
    ("roe", dict(conv_flux=O.ROE)),
    ("ch_roe", dict(use_split_form=1, two_point_flux=O.CH, conv_flux=O.TWO_POINT_ROE)),
    ("ra_lxf", dict(use_split_form=1, two_point_flux=O.RA, conv_flux=O.TWO_POINT_LXF)),
    ("ir_ec", dict(use_split_form=1, two_point_flux=O.IR, conv_flux=O.TWO_POINT)),
    ("kg_ec", dict(use_split_form=1, two_point_flux=O.KG, conv_flux=O.TWO_POINT)),
]
BOX = {2: (-1.0, 1.0), 3: (0.0, 2 * np.pi)}


def cons(rho, v, p):
    v = np.asarray(v, dtype=float)
    return np.concatenate([[rho], rho * v, [p / (GAM - 1) + 0.5 * rho * v @ v]])


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


@pytest.mark.parametrize("name,kw", SCHEMES, ids=[s[0] for s in SCHEMES])
@pytest.mark.parametrize("dim", [2, 3])
def test_affine_box_as_node_coordinates_is_the_cartesian_residual(name, kw, dim):
    o = O.Oracle(dim=dim, poly_degree=3, n_cells_1d=3, **kw)
    oc = O.Oracle(dim=dim, poly_degree=3, n_cells_1d=3, **kw)
    oc.set_grid_nodes(o.node_coords())
    C, J = oc.metric_terms()
    h = o.p.domain_right / 3
    # the curl form differentiates the products X_l dX_m/dxi, whose rounding grows with |X| (up to
    # 2 pi here): 1e-12 relative
    assert np.abs(J - h ** dim).max() < 1e-13 * h ** dim
    assert np.abs(C[..., :dim, :dim] - h ** (dim - 1) * np.eye(dim)).max() < 1e-12 * h ** (dim - 1)
    u = random_field(o, 1)
    r, rc = o.residual(u), oc.residual(u)
    assert np.abs(rc - r).max() < 1e-12 * np.abs(r).max()
    assert np.abs(oc.inverse_mass(rc) - o.inverse_mass(r)).max() < 1e-12 * np.abs(o.inverse_mass(r)).max()


def quadratic_map(dim, x):
    """A degree-2 polynomial mapping of the box and its Jacobian d(x')/dx."""
    y = x.copy()
    Jx = np.tile(np.eye(3), x.shape[:-1] + (1, 1))
    if dim == 2:
        a, b = 0.05, 0.04
        y[..., 0] = x[..., 0] + a * x[..., 0] * x[..., 1]
        y[..., 1] = x[..., 1] + b * x[..., 0] ** 2
        Jx[..., 0, 0] += a * x[..., 1]
        Jx[..., 0, 1] += a * x[..., 0]
        Jx[..., 1, 0] += 2 * b * x[..., 0]
    else:
        a, b, c = 0.03, 0.02, 0.025
        y[..., 0] = x[..., 0] + a * x[..., 1] * x[..., 2]
        y[..., 1] = x[..., 1] + b * x[..., 0] * x[..., 2]
        y[..., 2] = x[..., 2] + c * x[..., 0] * x[..., 1]
        Jx[..., 0, 1] += a * x[..., 2]
        Jx[..., 0, 2] += a * x[..., 1]
        Jx[..., 1, 0] += b * x[..., 2]
        Jx[..., 1, 2] += b * x[..., 0]
        Jx[..., 2, 0] += c * x[..., 1]
        Jx[..., 2, 1] += c * x[..., 0]
    return y, Jx


@pytest.mark.parametrize("dim", [2, 3])
def test_curl_form_cofactor_is_exact_for_a_quadratic_map(dim):
    o = O.Oracle(dim=dim, poly_degree=3, n_cells_1d=2, domain_left=0.0, domain_right=1.0)
    x = o.node_coords()
    y, Jx = quadratic_map(dim, x)
    oc = O.Oracle(dim=dim, poly_degree=3, n_cells_1d=2, domain_left=0.0, domain_right=1.0)
    oc.set_grid_nodes(y)
    C, J = oc.metric_terms()
    h = 0.5
    d = dim
    Jref = Jx[..., :d, :d] * h                       # dx'_i / dxi_j on the unit reference cell
    det = np.linalg.det(Jref)
    Cexp = det[..., None, None] * np.swapaxes(np.linalg.inv(Jref), -1, -2)   # C[m][n] = det (J^-1)[n][m]
    assert np.abs(J - det).max() < 1e-13
    assert np.abs(C[..., :d, :d] - Cexp).max() < 1e-13


def d_ref(o, f):
    """Reference-direction derivatives of nodal values f [cell][node] with the GLL derivative
    matrix on the unit cell."""
    ops = o.operators()
    D = ops["chi_grad"]
    n, dim = o.n, o.dim
    g = f.reshape((o.ncells,) + (n,) * dim)   # axes (cell, [z,] y, x)
    out = []
    for k in range(dim):
        ax = dim - k
        out.append(np.moveaxis(np.tensordot(D, np.moveaxis(g, ax, 0), axes=(1, 0)), 0, ax).reshape(o.ncells, -1))
    return out


@pytest.mark.parametrize("dim", [2, 3])
def test_discrete_metric_identities_on_the_curved_grid(dim):
    lo, hi = BOX[dim]
    o = O.Oracle(dim=dim, poly_degree=3, n_cells_1d=4, domain_left=lo, domain_right=hi, grid_warp=1)
    C, J = o.metric_terms()
    assert J.min() > 0
    scale = np.abs(C).max()
    for m in range(dim):
        div = sum(d_ref(o, C[:, :, m, n])[n] for n in range(dim))
        assert np.abs(div).max() < 1e-13 * scale * o.n
    if dim == 3:
        O.set_mutation(O.MUT_CROSS_PRODUCT_METRIC)
        try:
            om = O.Oracle(dim=dim, poly_degree=3, n_cells_1d=4, domain_left=lo, domain_right=hi, grid_warp=1)
        finally:
            O.set_mutation(O.MUT_NONE)
        Cm = om.metric_terms()[0]
        div = sum(d_ref(om, Cm[:, :, 0, n])[n] for n in range(dim))
        assert np.abs(div).max() > 1e-6 * scale  # the pointwise cofactor does not satisfy them
    # the grid is curved: the cofactor varies by O(beta) from node to node
    assert np.abs(C[..., 0, 0] - C[..., 0, 0].mean()).max() > 1e-2 * scale


@pytest.mark.parametrize("name,kw", SCHEMES, ids=[s[0] for s in SCHEMES])
@pytest.mark.parametrize("dim", [2, 3])
def test_free_stream_preserved_on_the_curved_grid(name, kw, dim):
    lo, hi = BOX[dim]
    o = O.Oracle(dim=dim, poly_degree=3, n_cells_1d=3, domain_left=lo, domain_right=hi, grid_warp=1,
                 mach_inf=0.3, **kw)
    uinf = cons(1.2, [0.3, -0.2, 0.1][:dim], 1.0 / (GAM * 0.09))
    u = np.tile(uinf[:, None], (o.ncells, 1, o.Np)).ravel()
    r = o.residual(u)
    # the reference's bar (freestream_preservation.cpp:262): 1e-13 on the residual's L2 norm
    assert np.linalg.norm(r) < 1e-13 * uinf[-1] * np.sqrt(r.size)
    assert np.abs(o.metric_terms()[0][..., 0, 1]).max() > 1e-3  # the grid is curved
    if dim == 3:
        # the planted mistake of the pointwise cofactor det(J) J^-T (no curl form) breaks the metric
        # identities and with them the free stream
        O.set_mutation(O.MUT_CROSS_PRODUCT_METRIC)
        try:
            om = O.Oracle(dim=dim, poly_degree=3, n_cells_1d=3, domain_left=lo, domain_right=hi, grid_warp=1,
                          mach_inf=0.3, **kw)
            assert np.linalg.norm(om.residual(u)) > 1e-6 * uinf[-1] * np.sqrt(r.size)
        finally:
            O.set_mutation(O.MUT_NONE)


@pytest.mark.parametrize("name,kw", SCHEMES, ids=[s[0] for s in SCHEMES])
def test_conservation_on_the_curved_periodic_grid(name, kw):
    o = O.Oracle(dim=3, poly_degree=2, n_cells_1d=3, domain_left=0.0, domain_right=2 * np.pi, grid_warp=1, **kw)
    u = random_field(o, 3)
    r = o.residual(u).reshape(o.ncells, o.ns, o.Np)
    tot = r.sum(axis=(0, 2))
    scale = np.abs(r).max() * r.size / o.ns
    assert np.abs(tot).max() / scale < 1e-14


@pytest.mark.parametrize("tp", [O.IR, O.CH, O.RA])
@pytest.mark.parametrize("dim", [2, 3])
def test_entropy_conservation_on_the_curved_grid(tp, dim):
    lo, hi = BOX[dim]
    kw = dict(dim=dim, poly_degree=3, n_cells_1d=3, domain_left=lo, domain_right=hi, grid_warp=1,
              use_split_form=1, two_point_flux=tp)
    o = O.Oracle(conv_flux=O.TWO_POINT, **kw)
    u = random_field(o, 4)
    v = np.concatenate([o.entropy_variables(x) for x in u.reshape(o.ncells, o.ns, o.Np).transpose(0, 2, 1).reshape(-1, o.ns)])
    v = v.reshape(o.ncells, o.Np, o.ns).transpose(0, 2, 1).ravel()

    def rate():
        r = o.residual(u)
        return abs(v @ r) / (np.abs(r).sum() * np.abs(v).max())

    # the bar of the Cartesian check (rounding: ~1e-13 there and here)
    assert rate() < 1e-12
    # the planted mistake of a two-point flux without the metric average breaks it
    O.set_mutation(O.MUT_NO_METRIC_AVERAGE)
    try:
        assert rate() > 1e-6
    finally:
        O.set_mutation(O.MUT_NONE)
    # with Roe dissipation the entropy decays
    o2 = O.Oracle(conv_flux=O.TWO_POINT_ROE, **kw)
    assert o2.entropy_rate(u) < 0


def rotation(dim):
    if dim == 2:
        t = 0.7
        return np.array([[np.cos(t), -np.sin(t)], [np.sin(t), np.cos(t)]])
    a, b = 0.6, -0.4
    Rz = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    Rx = np.array([[1, 0, 0], [0, np.cos(b), -np.sin(b)], [0, np.sin(b), np.cos(b)]])
    return Rz @ Rx


@pytest.mark.parametrize("name,kw", SCHEMES, ids=[s[0] for s in SCHEMES])
@pytest.mark.parametrize("dim", [2, 3])
def test_rotated_grid_rotates_the_residual(name, kw, dim):
    """In 2D the cofactor is linear in the Jacobian, so any grid rotates exactly (the curved grid
    here); the 3D curl form is exact -- hence rotates exactly -- where the nodes resolve the
    products X_l dX_m, so the 3D grid is the multilinear map of the cofactor test."""
    lo, hi = BOX[dim] if dim == 2 else (0.0, 1.0)
    base = dict(dim=dim, poly_degree=3, n_cells_1d=3, domain_left=lo, domain_right=hi, **kw)
    o = O.Oracle(**base, grid_warp=1) if dim == 2 else O.Oracle(**base)
    if dim == 3:
        o.set_grid_nodes(quadratic_map(dim, o.node_coords())[0])
    orot = O.Oracle(**base)
    R = rotation(dim)
    x = o.node_coords()
    xr = x.copy()
    xr[..., :dim] = x[..., :dim] @ R.T
    orot.set_grid_nodes(xr)
    u = random_field(o, 5).reshape(o.ncells, o.ns, o.Np)
    ur = u.copy()
    ur[:, 1:1 + dim] = np.einsum("ij,cjn->cin", R, u[:, 1:1 + dim])
    r = o.residual(u.ravel()).reshape(u.shape)
    rr = orot.residual(ur.ravel()).reshape(u.shape)
    expect = r.copy()
    expect[:, 1:1 + dim] = np.einsum("ij,cjn->cin", R, r[:, 1:1 + dim])
    assert np.abs(rr - expect).max() < 1e-12 * np.abs(r).max()


def smooth_state_2d(x, y):
    rho = 1.0 + 0.2 * np.sin(np.pi * x) * np.cos(np.pi * y)
    vx = 0.3 + 0.1 * np.cos(np.pi * y)
    vy = -0.2 + 0.1 * np.sin(np.pi * x)
    p = 5.0 + 0.5 * np.cos(np.pi * (x + y))
    return rho, vx, vy, p


def flux_2d(x, y):
    rho, vx, vy, p = smooth_state_2d(x, y)
    E = p / (GAM - 1) + 0.5 * rho * (vx * vx + vy * vy)
    Fx = np.stack([rho * vx, rho * vx * vx + p, rho * vx * vy, (E + p) * vx])
    Fy = np.stack([rho * vy, rho * vx * vy, rho * vy * vy + p, (E + p) * vy])
    return Fx, Fy


@pytest.mark.parametrize("name,kw", [SCHEMES[0], SCHEMES[2]], ids=["lxf", "ch_roe"])
def test_smooth_residual_converges_to_the_flux_divergence(name, kw):
    # p = 4 on 8^2 and 16^2 cells of the curved grid (whose sin(2 pi x) shear needs a few cells per
    # wavelength before the rate is asymptotic): the nodal error of du/dt falls as h^p
    errs = []
    for N in (8, 16):
        o = O.Oracle(dim=2, poly_degree=4, n_cells_1d=N, domain_left=-1.0, domain_right=1.0, grid_warp=1, **kw)
        x = o.node_coords()
        X, Y = x[..., 0], x[..., 1]
        rho, vx, vy, p = smooth_state_2d(X, Y)
        u = np.stack([rho, rho * vx, rho * vy, p / (GAM - 1) + 0.5 * rho * (vx * vx + vy * vy)], axis=1)
        dudt = o.inverse_mass(o.residual(u.ravel())).reshape(u.shape)
        eps = 1e-6
        divF = (flux_2d(X + eps, Y)[0] - flux_2d(X - eps, Y)[0]) / (2 * eps) \
            + (flux_2d(X, Y + eps)[1] - flux_2d(X, Y - eps)[1]) / (2 * eps)
        err = np.abs(dudt + np.moveaxis(divF, 0, 1)).max() / np.abs(divF).max()
        errs.append(err)
    rate = np.log2(errs[0] / errs[1])
    assert errs[1] < 1e-2 and rate > 3.5, (errs, rate)


# --- Navier-Stokes on curvilinear grids: the lifted gradient C grad_ref u / J and the viscous
# fluxes contracted with the cofactor (strong_dg.cpp:690-870 with the metric terms) ---

def ns_kw(split=1, diss=O.SIPG, **extra):
    sch = dict(use_split_form=1, two_point_flux=O.RA, conv_flux=O.TWO_POINT_ROE) if split else dict(conv_flux=O.ROE)
    kw = dict(dim=3, poly_degree=3, n_cells_1d=3, pde=O.NAVIER_STOKES, diss_flux=diss, reynolds_inf=100.0,
              prandtl=0.71, mach_inf=0.3, **sch)
    kw.update(extra)
    return kw


def test_lifted_gradient_of_a_linear_field_is_exact_on_a_resolved_map():
    # multilinear map: the curl form is the exact cofactor, the node values of a linear field are its
    # interpolant, so C grad_ref u / J is the field's gradient at every node of the cells away from
    # the periodic wrap (where the linear field jumps)
    o = O.Oracle(**ns_kw(domain_left=0.0, domain_right=1.0, n_cells_1d=4))
    x = quadratic_map(3, o.node_coords())[0]
    o.set_grid_nodes(x)
    a = np.array([[0.3, -0.2, 0.1], [0.05, 0.2, -0.1], [-0.1, 0.04, 0.3], [0.2, 0.1, -0.05], [1.0, -0.5, 0.25]])
    base = np.array([1.0, 0.1, -0.1, 0.05, 9.0])
    u = (base[None, :, None] + np.einsum("sk,cnk->csn", a, x)).ravel()
    o.residual(u)
    inner = np.all((o.cell_ijk() >= 1) & (o.cell_ijk() <= 2), axis=1)
    assert inner.sum() == 8
    for k in range(3):
        q = o.auxiliary(k).reshape(o.ncells, o.ns, o.Np)[inner]
        assert np.abs(q - a[None, :, k, None]).max() < 1e-12


@pytest.mark.parametrize("split,diss", [(0, O.SIPG), (1, O.SIPG), (1, O.BR2)])
def test_navier_stokes_free_stream_and_conservation_on_the_curved_grid(split, diss):
    o = O.Oracle(**ns_kw(split, diss, domain_left=0.0, domain_right=2 * np.pi, grid_warp=1))
    uinf = cons(1.2, [0.3, -0.2, 0.1], 1.0 / (GAM * 0.09))
    u = np.tile(uinf[:, None], (o.ncells, 1, o.Np)).ravel()
    r = o.residual(u)
    assert np.linalg.norm(r) < 1e-13 * uinf[-1] * np.sqrt(r.size)
    assert max(np.abs(o.auxiliary(k)).max() for k in range(3)) < 1e-12
    v = random_field(o, 6)
    r = o.residual(v).reshape(o.ncells, o.ns, o.Np)
    assert np.abs(r.sum(axis=(0, 2))).max() / (np.abs(r).max() * r.size / o.ns) < 1e-14


@pytest.mark.parametrize("split,diss", [(0, O.SIPG), (1, O.BR2)])
def test_navier_stokes_rotated_grid_rotates_the_residual(split, diss):
    base = ns_kw(split, diss, domain_left=0.0, domain_right=1.0)
    o, orot = O.Oracle(**base), O.Oracle(**base)
    x = quadratic_map(3, o.node_coords())[0]
    o.set_grid_nodes(x)
    R = rotation(3)
    orot.set_grid_nodes(x @ R.T)
    u = random_field(o, 7).reshape(o.ncells, o.ns, o.Np)
    ur = u.copy()
    ur[:, 1:4] = np.einsum("ij,cjn->cin", R, u[:, 1:4])
    r = o.residual(u.ravel()).reshape(u.shape)
    rr = orot.residual(ur.ravel()).reshape(u.shape)
    expect = r.copy()
    expect[:, 1:4] = np.einsum("ij,cjn->cin", R, r[:, 1:4])
    assert np.abs(rr - expect).max() < 1e-12 * np.abs(r).max()


# ---------------------------------------------------------------------------------------------
# Integrated quantities on curved cells (periodic_turbulence.cpp:439-671 with the metric at the
# QGauss points)
# ---------------------------------------------------------------------------------------------

def linear_state(o, y, a=0.3, b=-0.2, c=0.25):
    """rho = 1 + 0.2 y0, velocity linear in the physical coordinates: (a y1, c y0) in 2D,
    (a y1, b y2, c y0) in 3D, uniform pressure; sampled at the nodes y [cell][node][3]."""
    dim = o.dim
    rho = 1.0 + 0.2 * y[..., 0]
    vel = [a * y[..., 1], c * y[..., 0]] if dim == 2 else [a * y[..., 1], b * y[..., 2], c * y[..., 0]]
    pinf = 1.0 / (1.4 * 0.3 ** 2)
    u = o.vec().reshape(o.ncells, o.ns, o.Np)
    u[:, 0] = rho
    v2 = 0.0
    for k, v in enumerate(vel):
        u[:, 1 + k] = rho * v
        v2 = v2 + v * v
    u[:, -1] = pinf / 0.4 + 0.5 * rho * v2
    return u.ravel()


def analytic_integrals(dim, n_cells, a=0.3, b=-0.2, c=0.25, npts=14):
    """The same integrals of linear_state over the mapped unit box, by Gauss-Legendre quadrature of
    the analytic map and its analytic Jacobian (quadratic_map): KE, enstrophy, incompressible KE
    per unit box volume."""
    xg, wg = np.polynomial.legendre.leggauss(npts)
    xg, wg = 0.5 * (xg + 1.0), 0.5 * wg
    h = 1.0 / n_cells
    grids = np.meshgrid(*([xg] * dim), indexing="ij")
    wts = np.prod(np.meshgrid(*([wg] * dim), indexing="ij"), axis=0).ravel()
    om2 = (c - a) ** 2 if dim == 2 else a * a + b * b + c * c
    ke = ens = ike = 0.0
    for cell in np.ndindex(*([n_cells] * dim)):
        x = np.zeros((wts.size, 3))
        for k in range(dim):
            x[:, k] = h * (cell[k] + grids[k].ravel())
        y, Jx = quadratic_map(dim, x)
        det = np.linalg.det(Jx[:, :dim, :dim] * h)
        rho = 1.0 + 0.2 * y[:, 0]
        vel = [a * y[:, 1], c * y[:, 0]] if dim == 2 else [a * y[:, 1], b * y[:, 2], c * y[:, 0]]
        v2 = sum(v * v for v in vel)
        ke += np.sum(wts * det * 0.5 * rho * v2)
        ens += np.sum(wts * det * 0.5 * rho * om2)
        ike += np.sum(wts * det * 0.5 * v2)
    return np.array([ke, ens, ike])


def integrated_vec(q):
    return np.array([q["kinetic_energy"], q["enstrophy"], q["incompressible_kinetic_energy"]])


@pytest.mark.parametrize("dim", [2, 3])
def test_integrated_quantities_on_a_resolved_map_match_the_analytic_integrals(dim):
    # the map and the state are resolved at p = 3 (2D: the map is quadratic, the momentum cubic in
    # x; 3D: the map is multilinear, the curl form's products quadratic), so the metric at the
    # QGauss points, the physical velocity gradient and the integrands are exact polynomials
    o = O.Oracle(dim=dim, poly_degree=3, n_cells_1d=2, domain_left=0.0, domain_right=1.0)
    y = quadratic_map(dim, o.node_coords())[0]
    o.set_grid_nodes(y)
    u = linear_state(o, y)
    ref = analytic_integrals(dim, 2)
    got = integrated_vec(o.integrated(u))
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max(), (got, ref)
    # planted mistakes: the cofactor transposed in the gradient transform; the box's det J
    for mut, k in ((O.MUT_TRANSPOSED_METRIC, 1), (O.MUT_CELL_CONSTANT_DETJ, 0)):
        O.set_mutation(mut)
        try:
            bad = integrated_vec(o.integrated(u))
        finally:
            O.set_mutation(O.MUT_NONE)
        assert abs(bad[k] - ref[k]) > 1e-6 * abs(ref[k]), (mut, bad, ref)


@pytest.mark.parametrize("warp", [0, 1])
@pytest.mark.parametrize("dim,p,n", [(2, 3, 3), (3, 3, 3), (3, 6, 2)])
def test_integrated_quantities_of_a_uniform_state_on_the_curved_periodic_grid(dim, p, n, warp):
    # the periodic warp is a periodic displacement of the box: its det J integrates to the box
    # volume, so a uniform state gives 1/2 rho |v|^2 and no enstrophy (p = 6: QGauss(17), more
    # points than the basis holds, which the Gauss rule's scratch once overflowed)
    lo, hi = BOX[dim]
    o = O.Oracle(dim=dim, poly_degree=p, n_cells_1d=n, domain_left=lo, domain_right=hi, grid_warp=warp)
    rho, vel = 1.1, np.array([0.3, -0.2, 0.1])[:dim]
    u = o.vec().reshape(o.ncells, o.ns, o.Np)
    u[:, 0] = rho
    u[:, 1:1 + dim] = rho * vel[None, :, None]
    u[:, -1] = 10.0 + 0.5 * rho * vel @ vel
    q = integrated_vec(o.integrated(u.ravel()))
    assert abs(q[0] - 0.5 * rho * vel @ vel) <= 1e-13
    assert abs(q[1]) <= 1e-20
    assert abs(q[2] - 0.5 * vel @ vel) <= 1e-13


@pytest.mark.parametrize("dim", [2, 3])
def test_integrated_quantities_of_the_affine_box_given_as_nodes(dim):
    lo, hi = BOX[dim]
    kw = dict(dim=dim, poly_degree=3, n_cells_1d=3, domain_left=lo, domain_right=hi)
    o, oc = O.Oracle(**kw), O.Oracle(**kw)
    oc.set_grid_nodes(o.node_coords())
    rng = np.random.default_rng(5)
    u = o.ic_tgv().reshape(o.ncells, o.ns, o.Np) if dim == 3 else None
    if u is None:
        u = o.vec().reshape(o.ncells, o.ns, o.Np)
        u[:, 0] = 1.0
        u[:, 1:3] = 0.1 * np.sin(o.node_coords()[..., :2]).transpose(0, 2, 1)
        u[:, 3] = 20.0
    u = (u * (1.0 + 1e-2 * rng.standard_normal(u.shape))).ravel()
    a, b = integrated_vec(o.integrated(u)), integrated_vec(oc.integrated(u))
    assert np.abs(a - b).max() <= 1e-13 * np.abs(a).max()


@pytest.mark.parametrize("diss", [O.SIPG, O.BR2])
@pytest.mark.parametrize("split", [0, 1])
@pytest.mark.parametrize("walls", [0, 1])
def test_navier_stokes_affine_box_as_node_coordinates_is_the_cartesian_residual(diss, split, walls):
    # the curvilinear Navier-Stokes path (lifted gradient through the metric, contracted viscous
    # fluxes, per-cell penalty from the cell extents) reproduces the Cartesian one on the box:
    # every extent is h there, so the SIPG penalty is the box's deg (deg + 1) / h
    extra = dict(boundary_type=[1001, 1001, 0, 0, 1006, 1006], thermal_bc=1, wall_temperature=1.1) if walls else {}
    kw = ns_kw(split, diss, domain_left=0.0, domain_right=1.0, **extra)
    o, oc = O.Oracle(**kw), O.Oracle(**kw)
    oc.set_grid_nodes(o.node_coords())
    u = random_field(o, 4, vel=0.3)
    r, rc = o.residual(u), oc.residual(u)
    assert np.abs(rc - r).max() < 1e-12 * np.abs(r).max()
    for k in range(3):
        assert np.abs(oc.auxiliary(k) - o.auxiliary(k)).max() < 1e-12 * np.abs(o.auxiliary(k)).max()
    if diss == O.SIPG:
        # the penalty matters: a different penalty factor changes the residual
        rp = O.Oracle(**dict(kw, sipg_penalty_factor=2.0))
        rp.set_grid_nodes(o.node_coords())
        assert np.abs(rp.residual(u) - rc).max() > 1e-6 * np.abs(r).max()
