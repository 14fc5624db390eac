"""Point physics of the oracle pinned by identities the mathematics fixes:
entropy-variable round trip (reference euler_convert_entropy_var_conservative.cpp:54,
1e-12), two-point flux consistency and symmetry, Tadmor's entropy-conservation
condition for IR/CH/Ra, numerical-flux conservation F(uL,uR,n) = -F(uR,uL,-n)
(numerical_flux_conservation.cpp:21,37, 1e-12) and consistency F(u,u,n) = f(u).n,
log-mean series branch against the direct formula."""
import numpy as np
import pytest

import oracle as O


def random_states(o, rng, m, near=False):
    ns, d = o.ns, o.dim
    out = []
    pinf = 1.0 / (1.4 * o.p.mach_inf ** 2)
    for _ in range(m):
        rho = rng.uniform(0.5, 2.0)
        vel = rng.uniform(-1, 1, d)
        p = rng.uniform(0.5, 2.0) * pinf
        out.append(np.concatenate([[rho], rho * vel, [p / 0.4 + 0.5 * rho * vel @ vel]]))
    return out


@pytest.fixture(params=[2, 3])
def orc(request):
    return O.Oracle(dim=request.param, poly_degree=1, n_cells_1d=2, mach_inf=0.3)


def test_entropy_variable_round_trip(orc):
    rng = np.random.default_rng(1)
    for u in random_states(orc, rng, 200):
        v = orc.entropy_variables(u)
        assert np.abs(orc.conservative_from_entropy(v) - u).max() / np.abs(u).max() < 1e-12


def test_entropy_variables_closed_form(orc):
    rng = np.random.default_rng(2)
    g = 1.4
    for u in random_states(orc, rng, 20):
        rho, m, E = u[0], u[1:-1], u[-1]
        p = 0.4 * (E - 0.5 * m @ m / rho)
        s = np.log(p * rho ** -g)
        vstd = np.concatenate([[(g - s) / (g - 1) - 0.5 * rho * (m / rho) @ (m / rho) / p], m / p, [-rho / p]])
        assert np.allclose(orc.entropy_variables(u), (g - 1) * vstd, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("which", [O.KG, O.IR, O.CH, O.RA])
def test_two_point_consistency_and_symmetry(orc, which):
    rng = np.random.default_rng(3)
    us = random_states(orc, rng, 50)
    for a, b in zip(us[::2], us[1::2]):
        f = orc.convective_flux(a)
        assert np.abs(orc.two_point_flux(which, a, a) - f).max() / np.abs(f).max() < 1e-13
        fab, fba = orc.two_point_flux(which, a, b), orc.two_point_flux(which, b, a)
        assert np.abs(fab - fba).max() / np.abs(fab).max() < 1e-13


@pytest.mark.parametrize("which", [O.IR, O.CH, O.RA])
def test_tadmor_entropy_conservation(orc, which):
    # (v_L - v_R) . F#(u_L,u_R) = psi_L - psi_R ; entropy potential psi = (gamma-1) rho u.
    # Exact up to round-off when every log mean takes the log branch; the series branch of the
    # reference's log mean (euler.cpp:629-647) truncates at u^3/7, relative error < u^4/9 < 1.2e-9.
    rng = np.random.default_rng(4)
    far = near = 0
    for _ in range(400):
        a, b = random_states(orc, rng, 2)
        pa, pb = orc.pressure(a), orc.pressure(b)
        if which == O.IR:
            ratios = [np.sqrt(a[0] / pa) / np.sqrt(b[0] / pb), np.sqrt(a[0] * pa) / np.sqrt(b[0] * pb)]
        else:
            ratios = [a[0] / b[0], (a[0] / pa) / (b[0] / pb)]
        is_far = all(abs((r - 1) / (r + 1)) ** 2 >= 1e-2 for r in ratios)
        F = orc.two_point_flux(which, a, b)
        dv = orc.entropy_variables(a) - orc.entropy_variables(b)
        err = np.abs(dv @ F - 0.4 * (a[1:-1] - b[1:-1])).max() / (np.abs(dv) @ np.abs(F)).max()
        if is_far:
            far += 1
            assert err < 1e-13
        else:
            near += 1
            assert err < 2e-9
    assert far > 20 and near > 20


def test_kennedy_gruber_is_not_entropy_conservative(orc):
    rng = np.random.default_rng(5)
    a, b = random_states(orc, rng, 2)
    F = orc.two_point_flux(O.KG, a, b)
    dv = orc.entropy_variables(a) - orc.entropy_variables(b)
    assert np.abs(dv @ F - 0.4 * (a[1:-1] - b[1:-1])).max() / (np.abs(dv).max() * np.abs(F).max()) > 1e-8


def test_log_mean_branches():
    assert O.log_mean(1.0, 1.0) == 1.0
    for a, b in [(1.0, 1.05), (2.0, 1.9), (1.0, 3.0), (0.2, 5.0), (1.0, 1.2), (7.0, 6.999)]:
        direct = (a - b) / np.log1p((a - b) / b)
        z = a / b
        uu = ((z - 1) / (z + 1)) ** 2
        bound = (uu ** 4 / 9 * 1.1 if uu < 1e-2 else 0.0) + 1e-14
        assert abs(O.log_mean(a, b) - direct) / direct < bound


@pytest.mark.parametrize("which", [O.LXF, O.ROE, O.CENTRAL, O.TWO_POINT, O.TWO_POINT_LXF, O.TWO_POINT_ROE])
@pytest.mark.parametrize("tp", [O.IR, O.CH, O.RA])
def test_numerical_flux_conservation_consistency(which, tp):
    for dim in (2, 3):
        o = O.Oracle(dim=dim, poly_degree=1, n_cells_1d=2, mach_inf=0.3, two_point_flux=tp)
        rng = np.random.default_rng(6)
        us = random_states(o, rng, 40)
        for a, b in zip(us[::2], us[1::2]):
            n = rng.standard_normal(dim)
            n /= np.linalg.norm(n)
            f1 = o.numerical_flux(which, a, b, n)
            f2 = o.numerical_flux(which, b, a, -n)
            assert np.abs(f1 + f2).max() / np.abs(f1).max() < 1e-12
            fa = o.numerical_flux(which, a, a, n)
            assert np.abs(fa - o.convective_flux(a) @ n).max() / np.abs(fa).max() < 1e-12


def test_roe_upwinds_supersonic():
    # for supersonic flow along n every eigenvalue is positive: Roe = upwind flux f(u_L).n
    o = O.Oracle(dim=3, poly_degree=1, n_cells_1d=2, mach_inf=0.3)
    rho, p = 1.0, 1.0
    a = np.array([rho, 3.0, 0.1, 0.0, p / 0.4 + 0.5 * (9 + 0.01)])
    b = np.array([1.02, 3.05, 0.1, 0.0, 1.01 / 0.4 + 0.5 * (3.05 ** 2 / 1.02 + 0.01 / 1.02)])
    n = np.array([1.0, 0.0, 0.0])
    f = o.numerical_flux(O.ROE, a, b, n)
    assert np.abs(f - o.convective_flux(a) @ n).max() / np.abs(f).max() < 2e-3  # Harten fix inactive far from sonic


def test_viscous_flux_closed_form():
    # uniform translation: velocity gradient zero => momentum viscous flux zero; linear shear u = y
    o = O.Oracle(dim=3, poly_degree=1, n_cells_1d=2, pde=O.NAVIER_STOKES, mach_inf=0.1,
                 reynolds_inf=100.0, prandtl=0.71, use_constant_viscosity=1, constant_viscosity=1.0)
    rho, u0 = 1.0, 0.5
    pinf = 1 / (1.4 * 0.01)
    u = np.array([rho, rho * u0, 0, 0, pinf / 0.4 + 0.5 * rho * u0 ** 2])
    g = np.zeros((5, 3))
    g[1, 1] = rho * 1.0           # d(rho u)/dy = 1
    g[4, 1] = rho * u0 * 1.0      # d(E)/dy = rho u du/dy  (pressure uniform)
    F = o.dissipative_flux(u, g)
    mu = 1.0 / 100.0
    assert abs(F[1, 1] - (-mu)) < 1e-14 and abs(F[2, 0] - (-mu)) < 1e-14
    assert abs(F[4, 1] - (-u0 * mu)) < 1e-13
    assert np.abs(F[0]).max() == 0.0
