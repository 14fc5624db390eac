"""Boundary faces of the oracle (Euler::boundary_face_values, euler.cpp:1511-1567, and
assemble_boundary_term_strong, strong_dg.cpp:1583-2364), pinned by what each condition is
defined to do rather than by its formula:

* the exterior states: a free-stream or extrapolated state, the mirror state of a slip wall,
  the Riemann invariants and entropy of the characteristic farfield, the total pressure,
  total temperature and invariants of the subsonic inflow, the back pressure of the outflow;
* the assembled boundary term: a box with slip walls on every face has the residual of the
  doubled periodic box holding the mirror image of the state (the periodic residual is pinned
  by the reference's golden vectors), mass and energy are conserved in a closed slip-wall box,
  and a uniform free stream is preserved by the farfield, characteristic and extrapolation
  boundaries."""
import numpy as np
import pytest

import oracle as O

GAM = 1.4


def prim(u, dim):
    rho = u[0]
    v = u[1:1 + dim] / rho
    p = (GAM - 1) * (u[dim + 1] - 0.5 * rho * v @ v)
    return rho, v, p


def cons(rho, v, p):
    v = np.asarray(v, dtype=float)
    return np.concatenate([[rho], rho * v, [p / (GAM - 1) + 0.5 * rho * v @ v]])


def oracle(dim, mach=0.5, aoa=0.0, bc=1005, **kw):
    return O.Oracle(dim=dim, poly_degree=2, n_cells_1d=2, mach_inf=mach, angle_of_attack=aoa,
                    boundary_type=[bc] * (2 * dim), **kw)


def random_states(dim, n, rng, mach):
    pinf = 1.0 / (GAM * mach * mach)
    for _ in range(n):
        rho = rng.uniform(0.7, 1.3)
        v = rng.uniform(-0.8, 0.8, dim)
        p = pinf * rng.uniform(0.7, 1.3)
        yield cons(rho, v, p)


def normals(dim):
    for k in range(dim):
        for s in (-1.0, 1.0):
            n = np.zeros(dim)
            n[k] = s
            yield n
    n = np.ones(dim) / np.sqrt(dim)
    yield n


@pytest.mark.parametrize("dim", [2, 3])
def test_farfield_is_the_free_stream(dim):
    mach, aoa = 0.4, 0.3
    o = oracle(dim, mach, aoa)
    rng = np.random.default_rng(1)
    vinf = np.zeros(dim)
    vinf[0], vinf[1] = np.cos(aoa), np.sin(aoa)
    expect = cons(1.0, vinf, 1.0 / (GAM * mach * mach))
    for u in random_states(dim, 5, rng, mach):
        for n in normals(dim):
            np.testing.assert_allclose(o.boundary_state(1005, u, n), expect, rtol=1e-15, atol=1e-15)


@pytest.mark.parametrize("dim", [2, 3])
def test_extrapolation_is_the_interior_state(dim):
    o = oracle(dim, bc=1007)
    rng = np.random.default_rng(2)
    for u in random_states(dim, 5, rng, 0.5):
        for n in normals(dim):
            assert np.array_equal(o.boundary_state(1007, u, n), u)


@pytest.mark.parametrize("bc", [1001, 1006])
@pytest.mark.parametrize("dim", [2, 3])
def test_slip_wall_is_the_mirror_state(bc, dim):
    o = oracle(dim, bc=bc)
    rng = np.random.default_rng(3)
    for u in random_states(dim, 5, rng, 0.5):
        rho, v, p = prim(u, dim)
        for n in normals(dim):
            rb, vb, pb = prim(o.boundary_state(bc, u, n), dim)
            assert abs(rb - rho) < 1e-15 and abs(pb - p) < 1e-12 * p
            # normal component reversed, tangential kept: the reflection I - 2 n n^T
            np.testing.assert_allclose(vb, v - 2 * (v @ n) * n, atol=1e-14)
            np.testing.assert_allclose(vb @ n, -(v @ n), atol=1e-14)


def entropy(rho, p):
    return p / rho ** GAM


def snd(rho, p):
    return np.sqrt(GAM * p / rho)


@pytest.mark.parametrize("dim", [2, 3])
def test_riemann_farfield(dim):
    mach, aoa = 0.5, 0.2
    o = oracle(dim, mach, aoa, bc=1004)
    pinf = 1.0 / (GAM * mach * mach)
    vinf = np.zeros(dim)
    vinf[0], vinf[1] = np.cos(aoa), np.sin(aoa)
    uinf = cons(1.0, vinf, pinf)
    # the free stream is a fixed point for any normal
    for n in normals(dim):
        np.testing.assert_allclose(o.boundary_state(1004, uinf, n), uinf, rtol=1e-14, atol=1e-14)
    rng = np.random.default_rng(4)
    cinf = snd(1.0, pinf)
    for u in random_states(dim, 6, rng, mach):
        rho, v, p = prim(u, dim)
        for n in normals(dim):
            ub = o.boundary_state(1004, u, n)
            rb, vb, pb = prim(ub, dim)
            vn_i, vn_e, ci = v @ n, vinf @ n, snd(rho, p)
            r_out, r_in = vn_i + 2 * ci / (GAM - 1), vn_e - 2 * cinf / (GAM - 1)
            vn_b, c_b = 0.5 * (r_out + r_in), 0.25 * (GAM - 1) * (r_out - r_in)
            if abs(vn_b) >= abs(c_b):
                # supersonic through the face: the upwind state whole
                np.testing.assert_allclose(ub, u if vn_b >= 0 else uinf, rtol=1e-14, atol=1e-14)
                continue
            # subsonic: both invariants carried in, entropy and tangential velocity upwind
            assert abs(vb @ n + 2 * snd(rb, pb) / (GAM - 1) - r_out) < 1e-12 * abs(r_out)
            assert abs(vb @ n - 2 * snd(rb, pb) / (GAM - 1) - r_in) < 1e-12 * abs(r_in)
            src_rho, src_v, src_p = (1.0, vinf, pinf) if vn_b < 0 else (rho, v, p)
            assert abs(entropy(rb, pb) / entropy(src_rho, src_p) - 1) < 1e-12
            np.testing.assert_allclose(vb - (vb @ n) * n, src_v - (src_v @ n) * n, atol=1e-13)
    # supersonic outflow and inflow states
    n = np.zeros(dim)
    n[0] = 1.0
    fast = cons(1.0, 3.0 * cinf * n, pinf)
    np.testing.assert_allclose(o.boundary_state(1004, fast, n), fast, rtol=1e-15)
    np.testing.assert_allclose(o.boundary_state(1004, cons(1.0, -3.0 * cinf * n, pinf), n),
                               o.boundary_state(1004, cons(1.0, -3.0 * cinf * n, pinf), n))


@pytest.mark.parametrize("dim", [2, 3])
def test_subsonic_inflow_totals_and_invariant(dim):
    mach = 0.5
    o = oracle(dim, mach, bc=1003)
    pinf = 1.0 / (GAM * mach * mach)
    p0 = pinf * (1 + 0.5 * (GAM - 1) * mach ** 2) ** (GAM / (GAM - 1))
    T0 = (p0 / pinf) ** ((GAM - 1) / GAM)  # free-stream temperature 1
    # the free stream entering through the face is a fixed point
    n = np.zeros(dim)
    n[0] = -1.0
    vinf = np.zeros(dim)
    vinf[0] = 1.0
    uinf = cons(1.0, vinf, pinf)
    np.testing.assert_allclose(o.boundary_state(1003, uinf, n), uinf, rtol=1e-13, atol=1e-13)
    rng = np.random.default_rng(5)
    for u in random_states(dim, 6, rng, mach):
        rho, v, p = prim(u, dim)
        for n in normals(dim):
            rb, vb, pb = prim(o.boundary_state(1003, u, n), dim)
            # temperature T = gamma M^2 p / rho; velocity along the inward normal
            Tb = GAM * mach ** 2 * pb / rb
            speed = vb @ (-n)
            np.testing.assert_allclose(vb, speed * (-n), atol=1e-13)
            # total pressure and temperature of the reservoir: p0 = p (T0/T)^(gamma/(gamma-1))
            assert abs(pb * (T0 / Tb) ** (GAM / (GAM - 1)) / p0 - 1) < 1e-12
            # the boundary Mach number in the sense of the invariant's sound speed c_b:
            # T0/T = 1 + (gamma-1)/2 M_b^2, c_b = |v_b| / M_b
            mb = np.sqrt(2 / (GAM - 1) * (T0 / Tb - 1))
            cb = abs(speed) / mb  # c_b > 0; v_b may point outward (then M_b < 0)
            # the outgoing invariant and the total enthalpy of the interior are kept
            ci = snd(rho, p)
            r_i = v @ (-n) + 2 * ci / (GAM - 1)
            H_i = (u[-1] + p) / rho
            assert abs(speed + 2 * cb / (GAM - 1) - r_i) < 1e-11 * abs(r_i)
            assert abs(cb * cb / (GAM - 1) + 0.5 * speed ** 2 - H_i) < 1e-11 * H_i


def test_supersonic_inflow_is_the_free_stream():
    mach = 1.6
    o = oracle(2, mach, bc=1003)
    pinf = 1.0 / (GAM * mach * mach)
    rng = np.random.default_rng(6)
    for u in random_states(2, 3, rng, mach):
        for n in ([-1.0, 0.0], [0.0, 1.0]):
            n = np.array(n)
            np.testing.assert_allclose(o.boundary_state(1003, u, n), cons(1.0, -n, pinf), rtol=1e-13)


@pytest.mark.parametrize("dim", [2, 3])
def test_pressure_outflow(dim):
    mach = 0.5
    o = oracle(dim, mach, bc=1002)
    pinf = 1.0 / (GAM * mach * mach)
    rng = np.random.default_rng(7)
    for u in random_states(dim, 6, rng, mach):
        rho, v, p = prim(u, dim)
        for n in normals(dim):
            ub = o.boundary_state(1002, u, n)
            rb, vb, pb = prim(ub, dim)
            if np.sqrt(v @ v) > snd(rho, p):
                assert np.array_equal(ub, u)
                continue
            # back pressure 0.99 of the free-stream static pressure, interior velocity and
            # temperature
            assert abs(pb / (0.99 * pinf) - 1) < 1e-13
            np.testing.assert_allclose(vb, v, rtol=1e-14, atol=1e-15)
            assert abs((pb / rb) / (p / rho) - 1) < 1e-13
    fast = cons(1.0, np.r_[3.0 * snd(1.0, pinf), np.zeros(dim - 1)], pinf)
    assert np.array_equal(o.boundary_state(1002, fast, np.r_[1.0, np.zeros(dim - 1)]), fast)


def test_connectivity_marks_box_faces():
    o = O.Oracle(dim=3, poly_degree=1, n_cells_1d=4, boundary_type=[1006, 1006, 0, 0, 1004, 1004])
    ijk, nbr = o.cell_ijk(), o.neighbors()
    for e in range(o.ncells):
        for f in range(6):
            k, hi = f // 2, f % 2
            on_box = ijk[e, k] == (3 if hi else 0)
            expect_bc = [1006, 1006, 0, 0, 1004, 1004][f]
            if on_box and expect_bc:
                assert nbr[e, f] == -expect_bc
            else:
                assert nbr[e, f] >= 0
                d = (ijk[nbr[e, f]] - ijk[e]) % 4
                step = np.zeros(3, dtype=int)
                step[k] = 1 if hi else 3
                assert np.array_equal(d, step)


@pytest.mark.parametrize("bad", [[1006, 0, 0, 0], [999, 999, 0, 0], [1008, 1008, 0, 0]])
def test_invalid_boundary_types_rejected(bad):
    with pytest.raises(ValueError):
        O.Oracle(dim=2, poly_degree=1, n_cells_1d=2, boundary_type=bad)
    with pytest.raises(ValueError):
        # Navier-Stokes takes the walls only (the reference sets no viscous boundary state for the others)
        O.Oracle(dim=2, poly_degree=1, n_cells_1d=2, pde=O.NAVIER_STOKES, boundary_type=[1004, 1004, 0, 0])


SCHEMES = [
    ("lxf", dict(conv_flux=O.LXF)),
    ("roe", dict(conv_flux=O.ROE)),
    ("ch_roe", dict(use_split_form=1, two_point_flux=O.CH, conv_flux=O.TWO_POINT_ROE)),
    ("ra_lxf", dict(use_split_form=1, two_point_flux=O.RA, conv_flux=O.TWO_POINT_LXF)),
    ("ir_ec", dict(use_split_form=1, two_point_flux=O.IR, conv_flux=O.TWO_POINT)),
]


def random_field(o, rng, mach):
    pinf = 1.0 / (GAM * mach * mach)
    u = o.vec().reshape(o.ncells, o.ns, o.Np)
    rho = rng.uniform(0.8, 1.2, (o.ncells, o.Np))
    v = rng.uniform(-0.5, 0.5, (o.ncells, o.dim, o.Np))
    p = pinf * rng.uniform(0.8, 1.2, (o.ncells, o.Np))
    u[:, 0] = rho
    u[:, 1:1 + o.dim] = rho[:, None] * v
    u[:, -1] = p / (GAM - 1) + 0.5 * rho * (v * v).sum(axis=1)
    return u.ravel()


def mirror_into_doubled(o, big, u):
    """State of the doubled periodic box [-L, L]^d: the cell at ijk of the box [0, L]^d is
    placed at N + ijk, and its mirror images across every subset of the coordinate planes at
    N - 1 - ijk in the mirrored directions (node order reversed there, momentum negated)."""
    N, n, d = o.p.n_cells_1d, o.n, o.dim
    uc = u.reshape(o.ncells, o.ns, *([n] * d)[::-1])  # node axes (z,) y, x
    lookup = {tuple(c[:d]): i for i, c in enumerate(big.cell_ijk())}
    ub = big.vec().reshape(big.ncells, big.ns, *([n] * d)[::-1])
    ijk = o.cell_ijk()
    for e in range(o.ncells):
        for mask in range(1 << d):
            s = uc[e].copy()
            pos = []
            for k in range(d):
                if mask >> k & 1:
                    s = np.flip(s, axis=d - k)  # node axis of direction k (axes: state, (z,) y, x)
                    s[1 + k] = -s[1 + k]
                    pos.append(N - 1 - ijk[e, k])
                else:
                    pos.append(N + ijk[e, k])
            ub[lookup[tuple(pos)]] = s
    return ub.ravel(), lookup


@pytest.mark.parametrize("name,kw", SCHEMES, ids=[s[0] for s in SCHEMES])
@pytest.mark.parametrize("dim", [2, 3])
def test_slip_wall_box_equals_mirrored_periodic_box(name, kw, dim):
    mach = 0.4
    N = 2 if dim == 3 else 4
    L = 1.0
    o = O.Oracle(dim=dim, poly_degree=2, n_cells_1d=N, mach_inf=mach, domain_left=0.0, domain_right=L,
                 boundary_type=[1006] * (2 * dim), cell_order=O.ORDER_LEXICOGRAPHIC, **kw)
    big = O.Oracle(dim=dim, poly_degree=2, n_cells_1d=2 * N, mach_inf=mach, domain_left=-L, domain_right=L,
                   cell_order=O.ORDER_LEXICOGRAPHIC, **kw)
    rng = np.random.default_rng(11)
    u = random_field(o, rng, mach)
    ub, lookup = mirror_into_doubled(o, big, u)
    r = o.residual(u).reshape(o.ncells, o.ns, o.Np)
    rb = big.residual(ub).reshape(big.ncells, big.ns, big.Np)
    ijk = o.cell_ijk()
    scale = np.abs(r).max()
    for e in range(o.ncells):
        be = lookup[tuple(N + ijk[e, :dim])]
        np.testing.assert_allclose(r[e], rb[be], rtol=0, atol=1e-12 * scale)


@pytest.mark.parametrize("name,kw", SCHEMES, ids=[s[0] for s in SCHEMES])
def test_closed_slip_wall_box_conserves_mass_and_energy(name, kw):
    o = O.Oracle(dim=3, poly_degree=3, n_cells_1d=2, mach_inf=0.4, boundary_type=[1006] * 6, **kw)
    u = random_field(o, np.random.default_rng(12), 0.4)
    r = o.residual(u).reshape(o.ncells, o.ns, o.Np)
    tot = r.sum(axis=(0, 2))
    scale = np.abs(r).max() * r.size / o.ns
    assert abs(tot[0]) / scale < 1e-14 and abs(tot[-1]) / scale < 1e-14
    # the walls do exert a pressure force: momentum is not conserved
    assert np.abs(tot[1:4]).max() / scale > 1e-6


@pytest.mark.parametrize("bc", [1004, 1005, 1007])
@pytest.mark.parametrize("name,kw", SCHEMES, ids=[s[0] for s in SCHEMES])
@pytest.mark.parametrize("dim", [2, 3])
def test_free_stream_preserved_by_open_boundaries(bc, name, kw, dim):
    mach, aoa = 0.4, 0.35
    o = O.Oracle(dim=dim, poly_degree=2, n_cells_1d=2, mach_inf=mach, angle_of_attack=aoa,
                 boundary_type=[bc] * (2 * dim), **kw)
    vinf = np.zeros(dim)
    vinf[0], vinf[1] = np.cos(aoa), np.sin(aoa)
    uinf = cons(1.0, vinf, 1.0 / (GAM * mach * mach))
    u = np.tile(uinf[:, None], (o.ncells, 1, o.Np)).ravel()
    r = o.residual(u)
    assert np.abs(r).max() < 1e-12 * uinf[-1]
