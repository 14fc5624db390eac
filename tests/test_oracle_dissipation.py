"""The oracle's surface dissipation pinned without retyping its formulas.

* Lax-Friedrichs (convective_numerical_flux.cpp:188-213, euler.cpp:1049-1062): the dissipation
  -1/2 lambda (u_R - u_L) with lambda the larger spectral radius of the two states' normal flux
  Jacobians, each Jacobian built by complex-step differentiation of a numpy Euler flux and its
  spectral radius taken by numpy's eigenvalue solver.
* Roe-Pike with Harten's fix (convective_numerical_flux.cpp:216-232, 332-480): the Roe matrix is
  the normal flux Jacobian at the Roe-averaged state; the test first checks the Roe property
  f(u_R).n - f(u_L).n = A (u_R - u_L), then builds |A| as a matrix function of A (Sylvester's
  formula over its three distinct eigenvalues, Harten's fix applied to their absolute values;
  the oracle uses characteristic wave strengths instead) and compares -1/2 |A| (u_R - u_L).
* The whole residual: with an entropy-conservative split volume term and surface flux
  F* = F_ec.n + D, the semi-discrete entropy production sum_i v_i . R_i of the oracle equals
  sum over faces and face nodes of W_f h^(d-1) (v_R - v_L) . D, D computed as above in numpy.
* Each planted mistake (lambda = min instead of max, a dropped 1/2, no Harten fix, no shear
  wave) makes these checks fail.
* Incompressible kinetic energy (periodic_turbulence.cpp:439-671): closed form on a uniform
  state and an independent numpy quadrature on a non-uniform one.
"""
import contextlib

import numpy as np
import numpy.polynomial.legendre as L
import pytest

import oracle as O

GAM = 1.4


def euler_flux_n(u, n, gam=GAM):
    """f(u).n of the Euler equations (works for complex u: complex-step derivatives)."""
    rho, m, E = u[0], u[1:-1], u[-1]
    v = m / rho
    p = (gam - 1) * (E - 0.5 * rho * (v @ v))
    vn = v @ n
    return np.concatenate([[rho * vn], m * vn + p * n, [(E + p) * vn]])


def jacobian(u, n, h=1e-30):
    u = np.asarray(u, dtype=np.complex128)
    J = np.zeros((u.size, u.size))
    for j in range(u.size):
        du = np.zeros(u.size, dtype=np.complex128)
        du[j] = 1j * h
        J[:, j] = euler_flux_n(u + du, n).imag / h
    return J


def spectral_radius(u, n):
    return float(np.abs(np.linalg.eigvals(jacobian(u, n))).max())


def lxf_dissipation(a, b, n):
    lam = max(spectral_radius(a, n), spectral_radius(b, n))
    return -0.5 * lam * (b - a)


def roe_state(a, b, gam=GAM):
    """A state with the Roe averages of velocity and total enthalpy (the Jacobian of the Euler
    flux depends on the state only through them)."""
    d = a.size - 2
    def vh(u):
        rho, v = u[0], u[1:-1] / u[0]
        p = (gam - 1) * (u[-1] - 0.5 * rho * v @ v)
        return np.sqrt(rho), v, (u[-1] + p) / rho
    sa, va, ha = vh(a)
    sb, vb, hb = vh(b)
    v = (sa * va + sb * vb) / (sa + sb)
    H = (sa * ha + sb * hb) / (sa + sb)
    p = (gam - 1) / gam * (H - 0.5 * v @ v)  # for rho = 1
    return np.concatenate([[1.0], v, [p / (gam - 1) + 0.5 * v @ v]]), d


def family_eigs(u, n, gam=GAM):
    """|vn - c|, |vn|, |vn + c| of a state."""
    v = u[1:-1] / u[0]
    p = (gam - 1) * (u[-1] - 0.5 * u[0] * v @ v)
    c = np.sqrt(gam * p / u[0])
    vn = v @ n
    return np.abs([vn - c, vn, vn + c])


def roe_dissipation(a, b, n, harten=True, gam=GAM):
    """-1/2 |A| (u_R - u_L) with A the Roe matrix and |A| = p(A), p the Lagrange polynomial
    through the three distinct eigenvalues vn - c, vn, vn + c of A and their (Harten-fixed)
    absolute values (Sylvester's formula: A is diagonalisable)."""
    ur, d = roe_state(a, b)
    A = jacobian(ur, n)
    v = ur[1:-1]
    H = ur[-1] + (gam - 1) * (ur[-1] - 0.5 * v @ v)  # rho = 1: H = E + p
    c = np.sqrt((gam - 1) * (H - 0.5 * v @ v))
    vn = v @ n
    lam = np.array([vn - c, vn, vn + c])
    # the eigenvalues of the complex-step Jacobian are these (closed form), d-fold in the middle
    ev = np.sort(np.linalg.eigvals(A).real)
    assert np.abs(ev - np.sort(np.concatenate([[lam[0], lam[2]], [lam[1]] * d]))).max() <= 1e-10 * c
    ea, eb, er = family_eigs(a, n), family_eigs(b, n), np.abs(lam)
    absl = np.abs(lam)
    if harten:
        for f in range(3):
            eps = max(abs(er[f] - ea[f]), abs(eb[f] - er[f]))
            if absl[f] < eps:
                absl[f] = 0.5 * (absl[f] ** 2 / eps + eps)
    I = np.eye(A.shape[0])
    absA = np.zeros_like(A)
    for f in range(3):
        P = np.eye(A.shape[0])
        for g in range(3):
            if g != f:
                P = P @ (A - lam[g] * I) / (lam[f] - lam[g])
        absA += absl[f] * P
    return -0.5 * absA @ (b - a), A


def states(dim, rng, m, mach=0.3):
    pinf = 1.0 / (GAM * mach ** 2)
    out = []
    for _ in range(m):
        rho = rng.uniform(0.5, 2.0)
        vel = rng.uniform(-1, 1, dim)
        p = rng.uniform(0.5, 2.0) * pinf
        out.append(np.concatenate([[rho], rho * vel, [p / (GAM - 1) + 0.5 * rho * vel @ vel]]))
    return out


def near_pairs(dim, rng, m, mach=0.3, amp=0.05):
    """Pairs of close states (face traces of a smooth field): the Harten fix is active for the
    middle wave where |vn| is small compared with its jump."""
    out = []
    for a in states(dim, rng, m, mach):
        b = a * (1.0 + amp * rng.uniform(-1, 1, a.size))
        b[1] = a[1] * rng.uniform(-1, 1)  # normal velocity of either sign, possibly small
        out.append((a, b))
    return out


@contextlib.contextmanager
def mutation(m):
    O.set_mutation(m)
    try:
        yield
    finally:
        O.set_mutation(O.MUT_NONE)


def central(a, b, n):
    return 0.5 * (euler_flux_n(a, n) + euler_flux_n(b, n))


def lxf_point_error(o, pairs, which=O.LXF):
    err = 0.0
    for a, b, n in pairs:
        f = o.numerical_flux(which, a, b, n)
        base = central(a, b, n) if which == O.LXF else o.numerical_flux(O.TWO_POINT, a, b, n)
        D = lxf_dissipation(a, b, n)
        err = max(err, np.abs((f - base) - D).max() / np.abs(D).max())
    return err


def roe_point_error(o, pairs, which=O.ROE):
    err = 0.0
    for a, b, n in pairs:
        f = o.numerical_flux(which, a, b, n)
        base = central(a, b, n) if which == O.ROE else o.numerical_flux(O.TWO_POINT, a, b, n)
        D, _ = roe_dissipation(a, b, n)
        err = max(err, np.abs((f - base) - D).max() / np.abs(D).max())
    return err


def point_pairs(dim, seed, m=30):
    rng = np.random.default_rng(seed)
    pairs = []
    for a, b in near_pairs(dim, rng, m) + list(zip(*[iter(states(dim, rng, 2 * m))] * 2)):
        n = rng.standard_normal(dim)
        n /= np.linalg.norm(n)
        pairs.append((a, b, n))
    return pairs


def test_complex_step_jacobian_spectral_radius_is_vn_plus_c():
    # the numpy construction itself: spectral radius |v.n| + c (closed form)
    rng = np.random.default_rng(1)
    for u in states(3, rng, 10):
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)
        v = u[1:4] / u[0]
        p = 0.4 * (u[4] - 0.5 * u[0] * v @ v)
        assert abs(spectral_radius(u, n) - (abs(v @ n) + np.sqrt(1.4 * p / u[0]))) < 1e-12 * spectral_radius(u, n)


@pytest.mark.parametrize("dim", [2, 3])
def test_roe_matrix_property(dim):
    # f(u_R).n - f(u_L).n = A_Roe (u_R - u_L): pins the numpy Roe average used below
    rng = np.random.default_rng(2)
    for a, b, n in point_pairs(dim, 3):
        _, A = roe_dissipation(a, b, n)
        df = euler_flux_n(b, n) - euler_flux_n(a, n)
        assert np.abs(A @ (b - a) - df).max() <= 1e-11 * np.abs(df).max()


@pytest.mark.parametrize("dim", [2, 3])
def test_lxf_dissipation_pointwise(dim):
    o = O.Oracle(dim=dim, poly_degree=1, n_cells_1d=2, mach_inf=0.3, two_point_flux=O.CH)
    pairs = point_pairs(dim, 4)
    assert lxf_point_error(o, pairs, O.LXF) < 1e-12
    assert lxf_point_error(o, pairs, O.TWO_POINT_LXF) < 1e-12
    for m in (O.MUT_LXF_MIN, O.MUT_LXF_NO_HALF):
        with mutation(m):
            assert lxf_point_error(o, pairs, O.LXF) > 1e-6


@pytest.mark.parametrize("dim", [2, 3])
def test_roe_dissipation_pointwise(dim):
    o = O.Oracle(dim=dim, poly_degree=1, n_cells_1d=2, mach_inf=0.3, two_point_flux=O.CH)
    pairs = point_pairs(dim, 5)
    assert roe_point_error(o, pairs, O.ROE) < 1e-10
    assert roe_point_error(o, pairs, O.TWO_POINT_ROE) < 1e-10
    for m in (O.MUT_ROE_NO_HALF, O.MUT_ROE_NO_HARTEN, O.MUT_ROE_NO_SHEAR):
        with mutation(m):
            assert roe_point_error(o, pairs, O.ROE) > 1e-6


def entropy_vars(u, gam=GAM):
    """(gamma - 1) x the standard entropy variables of the entropy -rho s / (gamma - 1)."""
    rho, m, E = u[0], u[1:-1], u[-1]
    v = m / rho
    p = (gam - 1) * (E - 0.5 * rho * v @ v)
    s = np.log(p * rho ** -gam)
    return (gam - 1) * np.concatenate([[(gam - s) / (gam - 1) - 0.5 * rho * v @ v / p], m / p, [-rho / p]])


def gll01(n):
    c = np.zeros(n)
    c[-1] = 1
    x = np.concatenate(([-1.0], np.sort(L.legroots(L.legder(c))), [1.0]))
    w = 2.0 / (n * (n - 1) * L.legval(x, c) ** 2)
    return 0.5 * (x + 1), 0.5 * w


def dissipation_entropy_rate(o, u, diss):
    """sum over faces and face nodes of W_f h^(d-1) (v_R - v_L) . D(u_L, u_R, n), the owner
    (lower cell index) being L and n its outward normal."""
    d, n = o.dim, o.n
    _, w = gll01(n)
    h = (o.p.domain_right - o.p.domain_left) / o.p.n_cells_1d
    U = u.reshape(o.ncells, o.ns, o.Np)
    total = 0.0
    for cell, face, nb, _ in o.face_list():
        k, side = face // 2, face % 2
        nrm = np.zeros(d)
        nrm[k] = 1.0 if side else -1.0
        for fn in range(n ** (d - 1)):
            rest = [fn % n, fn // n]
            ci, cj = [], []
            r = 0
            for kk in range(d):
                if kk == k:
                    ci.append(n - 1 if side else 0)
                    cj.append(0 if side else n - 1)
                else:
                    ci.append(rest[r])
                    cj.append(rest[r])
                    r += 1
            i = sum(ci[kk] * n ** kk for kk in range(d))
            j = sum(cj[kk] * n ** kk for kk in range(d))
            wf = np.prod([w[ci[kk]] for kk in range(d) if kk != k]) * h ** (d - 1)
            uL, uR = U[cell, :, i], U[nb, :, j]
            total += wf * (entropy_vars(uR) - entropy_vars(uL)) @ diss(uL, uR, nrm)
    return total


def perturbed_tgv(o, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    u = o.ic_tgv().reshape(o.ncells, o.ns, o.Np)
    u = u * (1.0 + amp * rng.uniform(-1, 1, u.shape))
    return u.ravel()


@pytest.mark.parametrize("tp", [O.CH, O.RA, O.IR])
def test_entropy_production_lxf_and_roe(tp):
    # 3^3 cells (lexicographic order): the -x and +x neighbours of a cell differ
    kw = dict(dim=3, poly_degree=2, n_cells_1d=3, use_split_form=1, two_point_flux=tp, mach_inf=0.3)
    u = perturbed_tgv(O.Oracle(**kw), seed=tp)
    for which, diss, muts in [(O.TWO_POINT_LXF, lxf_dissipation, (O.MUT_LXF_MIN, O.MUT_LXF_NO_HALF)),
                              (O.TWO_POINT_ROE, lambda a, b, n: roe_dissipation(a, b, n)[0],
                               (O.MUT_ROE_NO_HALF, O.MUT_ROE_NO_SHEAR))]:
        o = O.Oracle(conv_flux=which, **kw)
        expect = dissipation_entropy_rate(o, u, diss)
        assert expect < 0
        got = o.entropy_rate(u)
        assert abs(got - expect) <= 1e-9 * abs(expect), (which, got, expect)
        for m in muts:
            with mutation(m):
                assert abs(o.entropy_rate(u) - expect) > 1e-4 * abs(expect), (which, m)


def test_incompressible_kinetic_energy_closed_form_and_quadrature():
    o = O.Oracle(dim=3, poly_degree=3, n_cells_1d=3)
    rho, vel, p = 1.3, np.array([0.2, -0.1, 0.05]), 7.0
    u = np.zeros((o.ncells, o.ns, o.Np))
    u[:, 0], u[:, 1:4], u[:, 4] = rho, (rho * vel)[None, :, None], p / 0.4 + 0.5 * rho * vel @ vel
    q = o.integrated(u.ravel())
    assert abs(q["incompressible_kinetic_energy"] - 0.5 * vel @ vel) <= 1e-12 * 0.5 * vel @ vel
    assert abs(q["kinetic_energy"] - 0.5 * rho * vel @ vel) <= 1e-12 * 0.5 * rho * vel @ vel
    # non-uniform state: numpy interpolation of the GLL nodal values to QGauss(p+11)
    w = perturbed_tgv(o, seed=9)
    n, nx = o.n, o.n + 10
    xs, _ = gll01(n)
    xq, wq = L.leggauss(nx)
    xq, wq = 0.5 * (xq + 1), 0.5 * wq
    A = np.ones((nx, n))
    for j in range(n):
        for m_ in range(n):
            if m_ != j:
                A[:, j] *= (xq - xs[m_]) / (xs[j] - xs[m_])
    U = w.reshape(o.ncells, o.ns, n, n, n)
    Uq = np.einsum("ia,jb,kc,esabc->esijk", A, A, A, U)
    v2 = ((Uq[:, 1:4] / Uq[:, 0:1]) ** 2).sum(1)
    W = np.einsum("i,j,k->ijk", wq, wq, wq)
    ref = float((0.5 * v2 * W[None]).sum()) * (2 * np.pi / 3) ** 3 / (2 * np.pi) ** 3
    got = o.integrated(w)["incompressible_kinetic_energy"]
    assert abs(got - ref) <= 1e-13 * ref
