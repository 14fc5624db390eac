"""Oracle integrated numerical entropy (periodic_turbulence.cpp:765-868):
sum over cells and volume quadrature points of w_q det J (-rho ln(p rho^-gamma)).
Pinned by the closed form for a uniform state and by an independent numpy
construction of the quadrature (numpy's Gauss-Lobatto/Gauss-Legendre rules and
Lagrange interpolation from the GLL solution points) on a non-uniform state."""
import numpy as np
import numpy.polynomial.legendre as L
import pytest

import oracle as O


def gll(n):
    """Gauss-Lobatto nodes/weights on [0,1] from numpy (roots of P'_{n-1} and the end points)."""
    c = np.zeros(n)
    c[-1] = 1
    x = np.concatenate(([-1.0], np.sort(L.legroots(L.legder(c))), [1.0]))
    w = 2.0 / (n * (n - 1) * L.legval(x, c) ** 2)
    return 0.5 * (x + 1), 0.5 * w


def lagrange_matrix(xs, xq):
    A = np.ones((len(xq), len(xs)))
    for j, xj in enumerate(xs):
        for m, xm in enumerate(xs):
            if m != j:
                A[:, j] *= (xq - xm) / (xj - xm)
    return A


def entropy_numpy(o, u, gam=1.4, gl=False):
    n, dim = o.n, o.dim
    xs, ws = gll(n)
    if gl:
        xq, wq = L.leggauss(n)
        xq, wq = 0.5 * (xq + 1), 0.5 * wq
    else:
        xq, wq = xs, ws
    A = lagrange_matrix(xs, xq)
    h = (o.p.domain_right - o.p.domain_left) / o.p.n_cells_1d
    U = u.reshape(o.ncells, o.ns, *([n] * dim))  # node index x fastest -> last axis
    if dim == 3:
        Uq = np.einsum("ia,jb,kc,esabc->esijk", A, A, A, U[..., :, :, :].transpose(0, 1, 2, 3, 4))
        W = np.einsum("i,j,k->ijk", wq, wq, wq)
    else:
        Uq = np.einsum("ia,jb,esab->esij", A, A, U)
        W = np.einsum("i,j->ij", wq, wq)
    rho = Uq[:, 0]
    ke = 0.5 * (Uq[:, 1:1 + dim] ** 2).sum(1) / rho
    p = (gam - 1) * (Uq[:, dim + 1] - ke)
    s = np.log(p * rho ** (-gam))
    return float((-rho * s * W[None] * h ** dim).sum())


def test_uniform_state_closed_form():
    o = O.Oracle(dim=3, poly_degree=3, n_cells_1d=3)
    rho, vel, p, gam = 1.3, np.array([0.2, -0.1, 0.05]), 7.0, 1.4
    E = p / (gam - 1) + 0.5 * rho * vel @ vel
    u = np.zeros((o.ncells, o.ns, o.Np))
    u[:, 0], u[:, 1:4], u[:, 4] = rho, (rho * vel)[None, :, None], E
    exact = -rho * np.log(p * rho ** -gam) * (2 * np.pi) ** 3
    assert abs(o.numerical_entropy(u.ravel()) - exact) <= 1e-13 * abs(exact)


@pytest.mark.parametrize("dim,p", [(2, 4), (3, 2), (3, 3)])
def test_gll_matches_independent_quadrature(dim, p):
    o = O.Oracle(dim=dim, poly_degree=p, n_cells_1d=3)
    rng = np.random.default_rng(p)
    u = o.ic_tgv() if dim == 3 else o.ic_vortex(strength=3.0, u_inf=0.5, v_inf=0.4)
    u = u * (1 + 1e-2 * rng.standard_normal(u.size))
    ref = entropy_numpy(o, u)
    assert abs(o.numerical_entropy(u) - ref) <= 1e-12 * abs(ref)


def test_gl_flux_nodes_interpolate_to_gauss_points():
    o = O.Oracle(dim=3, poly_degree=3, n_cells_1d=2, flux_nodes=O.GL)
    u = o.ic_tgv()
    ref = entropy_numpy(o, u, gl=True)
    assert abs(o.numerical_entropy(u) - ref) <= 1e-12 * abs(ref)
    # and it differs from the nodal (GLL) rule, so the interpolation is exercised
    assert abs(entropy_numpy(o, u, gl=False) - ref) > 1e-9 * abs(ref)


def test_max_wave_speed_uniform_state_and_independent_interpolation():
    o = O.Oracle(dim=3, poly_degree=2, n_cells_1d=2, mach_inf=0.2)
    gam = 1.4
    rho, vel, p = 1.1, np.array([0.3, -0.4, 0.1]), 9.0
    u = np.zeros((o.ncells, o.ns, o.Np))
    u[:, 0], u[:, 1:4], u[:, 4] = rho, (rho * vel)[None, :, None], p / (gam - 1) + 0.5 * rho * vel @ vel
    exact = np.linalg.norm(vel) + np.sqrt(gam * p / rho)
    assert abs(o.max_wave_speed(u.ravel()) - exact) <= 1e-14 * exact
    # non-uniform: numpy interpolation of the GLL nodal values to QGauss(p+11) points
    w = o.ic_tgv()
    n, nx = o.n, o.n + 10
    xs, _ = gll(n)
    xq, _ = L.leggauss(nx)
    A = lagrange_matrix(xs, 0.5 * (xq + 1))
    U = w.reshape(o.ncells, o.ns, n, n, n)
    Uq = np.einsum("ia,jb,kc,esabc->esijk", A, A, A, U)
    v = Uq[:, 1:4] / Uq[:, 0:1]
    pq = (gam - 1) * (Uq[:, 4] - 0.5 * Uq[:, 0] * (v ** 2).sum(1))
    ref = float((np.sqrt((v ** 2).sum(1)) + np.sqrt(gam * pq / Uq[:, 0])).max())
    assert abs(o.max_wave_speed(w) - ref) <= 1e-13 * ref
    h = 2 * np.pi / (o.ndofs / o.ns) ** (1 / 3)
    assert abs(o.adaptive_time_step(w, 0.3) - 0.3 * h / ref) <= 1e-13 * 0.3 * h / ref
