"""Two properties the reference checks on its own solver, checked on the device path through
the C ABI (FP64 path):

* Free-stream preservation (tests/unit_tests/euler_unit_test/freestream_preservation.cpp:
  187-265): a uniform state at M = 0.3 and 36 degrees angle of attack has a residual whose L2
  norm, sqrt(sum over GLL nodes of R^2 w det J) / volume as get_residual_l2norm computes it for
  collocated nodes (dg_base.cpp:2965-3019), is at most 1e-13; every surface flux and volume form,
  p = 2..6, 2D and 3D.  The reference runs it on curved grids, which this path does not have.
* Semi-discrete entropy conservation of the split forms (euler_entropy_conserving_split_forms_
  check.cpp:20-66): with an entropy-conserving volume flux (Ismail-Roe, Chandrashekar, Ranocha)
  and its own flux at the faces, the entropy production sum_i v(u_i) . R_i vanishes to rounding:
  |sum| <= 1e-13 x sum_i |v(u_i) . R_i| on a TGV (3D) or uniform (2D) state perturbed by 1e-3
  relative noise (jumps small enough that the log mean's truncated series, euler.cpp:629-647,
  is exact to rounding; the oracle gives <= 4e-16 there); with Lax-Friedrichs or Roe dissipation
  added at the faces it is negative (<= -1e-6 x the same scale; the oracle gives -1e-4 .. -0.2).
"""
import sys

import numpy as np
import pytest

import oracle as O
import philip_amd as P

pytestmark = pytest.mark.gpu

GAMMA = 1.4


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def uniform_state(o, dim, mach=0.3, aoa_deg=36.0):
    a = np.deg2rad(aoa_deg)
    vel = np.array([np.cos(a), np.sin(a), 0.0])[:dim]
    p = 1.0 / (GAMMA * mach * mach)
    u = np.zeros((o.ncells, dim + 2, o.Np))
    u[:, 0] = 1.0
    u[:, 1:1 + dim] = vel[None, :, None]
    u[:, dim + 1] = p / (GAMMA - 1.0) + 0.5 * (vel ** 2).sum()
    return u.ravel()


def residual_l2norm(kw, o, r):
    """get_residual_l2norm for collocated GLL nodes: sqrt(sum R^2 w det J) / volume."""
    dim, p = kw["dim"], kw["poly_degree"]
    _, w = O.gauss_lobatto(p + 1)          # on [0, 1]
    wq = w
    for _ in range(dim - 1):
        wq = np.outer(wq, w).ravel()       # node weights, x fastest
    h = (kw.get("domain_right", 2 * np.pi) - kw.get("domain_left", 0.0)) / kw["n_cells_1d"]
    rr = r.reshape(o.ncells, dim + 2, o.Np)
    jxw = wq * h ** dim
    return np.sqrt((rr ** 2 * jxw[None, None, :]).sum()) / (o.ncells * h ** dim)


SCHEMES = {
    "lxf": dict(conv_flux=O.LXF),
    "roe": dict(conv_flux=O.ROE),
    "central": dict(conv_flux=O.CENTRAL),
    "ch_ec": dict(use_split_form=1, two_point_flux=O.CH, conv_flux=O.TWO_POINT),
    "ch_roe": dict(use_split_form=1, two_point_flux=O.CH, conv_flux=O.TWO_POINT_ROE),
    "ra_lxf": dict(use_split_form=1, two_point_flux=O.RA, conv_flux=O.TWO_POINT_LXF),
    "ir_ec": dict(use_split_form=1, two_point_flux=O.IR, conv_flux=O.TWO_POINT),
    "kg_roe": dict(use_split_form=1, two_point_flux=O.KG, conv_flux=O.TWO_POINT_ROE),
}


@pytest.mark.parametrize("dim,p", [(3, 2), (3, 3), (3, 4), (3, 5), (3, 6), (2, 3), (2, 6)])
def test_freestream_preservation(dim, p):
    worst = 0.0
    for name, sch in SCHEMES.items():
        kw = dict(dim=dim, poly_degree=p, n_cells_1d=4 if dim == 3 else 6, mach_inf=0.3, **sch)
        o, d = O.Oracle(**kw), P.DG(**kw)
        u = uniform_state(o, dim)
        d.set_solution(u)
        r = d.assemble_residual()
        nrm = residual_l2norm(kw, o, r)
        worst = max(worst, nrm)
        assert nrm <= 1e-13, (name, nrm)
        d.close()
    log(f"free stream {dim}D p={p}: worst residual L2 norm {worst:.2e} over {len(SCHEMES)} schemes (bar 1e-13)")


def entropy_variables(u, dim):
    rho, m, E = u[:, 0], u[:, 1:1 + dim], u[:, dim + 1]
    vel = m / rho[:, None]
    v2 = (vel ** 2).sum(1)
    pr = (GAMMA - 1.0) * (E - 0.5 * rho * v2)
    s = np.log(pr * rho ** -GAMMA)
    v = np.empty_like(u)
    v[:, 0] = (GAMMA - s) / (GAMMA - 1.0) - 0.5 * rho * v2 / pr
    v[:, 1:1 + dim] = rho[:, None] * vel / pr[:, None]
    v[:, dim + 1] = -rho / pr
    return v


@pytest.mark.parametrize("tp", [O.IR, O.CH, O.RA])
@pytest.mark.parametrize("dim,p", [(3, 3), (3, 5), (2, 4)])
def test_semi_discrete_entropy_conservation(tp, dim, p):
    rng = np.random.default_rng(11 + tp + p)
    out = []
    for cf in (O.TWO_POINT, O.TWO_POINT_LXF, O.TWO_POINT_ROE):
        kw = dict(dim=dim, poly_degree=p, n_cells_1d=4 if dim == 3 else 6, use_split_form=1, two_point_flux=tp,
                  conv_flux=cf, mach_inf=0.3)
        o, d = O.Oracle(**kw), P.DG(**kw)
        u = (o.ic_tgv() if dim == 3 else uniform_state(o, dim)).reshape(o.ncells, o.ns, o.Np)
        u = u * (1.0 + 1e-3 * rng.standard_normal(u.shape))
        d.set_solution(u.ravel())
        r = d.assemble_residual().reshape(u.shape)
        uq = u.transpose(0, 2, 1).reshape(-1, dim + 2)
        rq = r.transpose(0, 2, 1).reshape(-1, dim + 2)
        prod = (entropy_variables(uq, dim) * rq).sum(1)
        e, scale = prod.sum(), np.abs(prod).sum()
        out.append(e / scale)
        d.close()
    log(f"entropy production {dim}D p={p} two-point flux {tp}: EC {out[0]:.2e}, +LxF {out[1]:.2e}, "
        f"+Roe {out[2]:.2e} (relative to sum |v.R|)")
    assert abs(out[0]) <= 1e-13
    assert out[1] <= -1e-6 and out[2] <= -1e-6
