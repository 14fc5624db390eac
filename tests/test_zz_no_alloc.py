"""Allocation-free integrated quantities (dgx_integrated_quantities_static): same results as
dgx_integrated_quantities, bit for bit, on two calls in a row on one GPU (the second on a smaller
mesh of another degree, with another over-integration), through the C ABI; sizes beyond the
static scratch are refused."""
import sys

import numpy as np
import pytest

import oracle as O
import philip_amd as P

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def perturbed_tgv(o, seed):
    rng = np.random.default_rng(seed)
    u = o.ic_tgv().reshape(o.ncells, o.ns, o.Np)
    return (u * (1.0 + 1e-2 * rng.standard_normal(u.shape))).ravel()


def test_static_integrated_quantities_match_the_allocating_entry_point():
    from hipmem import DeviceBuffer
    lib = P.load()
    out = DeviceBuffer.from_array(np.full(4, -1.0))
    for k, (p, n, overint) in enumerate([(3, 8, 10), (2, 4, 4)]):  # second call: smaller input
        kw = dict(dim=3, poly_degree=p, n_cells_1d=n, pde=P.NAVIER_STOKES, reynolds_inf=1600.0)
        o = O.Oracle(**kw)
        d = P.DG(**kw)
        u = perturbed_tgv(o, seed=k)
        d.set_solution(u)
        rc = lib.dgx_integrated_quantities_static(d.h, None, overint, out.value)
        assert rc == 0, lib.dgx_last_error(d.h)
        d.synchronize()
        got = out.download()
        ref = d.integrated(overint)
        ref = np.array([ref["kinetic_energy"], ref["enstrophy"], ref["incompressible_kinetic_energy"],
                        ref["vorticity_dissipation"]])
        log(f"call {k}: p={p} {n}^3 overint={overint}: static {got}, allocating {ref}")
        assert np.array_equal(got, ref)
        # and the caller's own device buffer as input
        ud = DeviceBuffer.from_array(u)
        assert lib.dgx_integrated_quantities_static(d.h, ud.value, overint, out.value) == 0
        d.synchronize()
        assert np.array_equal(out.download(), ref)
        d.close()


def test_static_integrated_quantities_refuse_sizes_beyond_the_scratch():
    from hipmem import DeviceBuffer
    lib = P.load()
    out = DeviceBuffer(4)
    d = P.DG(dim=3, poly_degree=1, n_cells_1d=65)            # 274,625 cells > 262,144
    assert lib.dgx_integrated_quantities_static(d.h, None, 10, out.value) == -1
    d.close()
    d = P.DG(dim=3, poly_degree=6, n_cells_1d=2)             # p + 1 + 18 = 25 > 24 points
    assert lib.dgx_integrated_quantities_static(d.h, None, 18, out.value) == -1
    assert lib.dgx_integrated_quantities_static(d.h, None, 17, out.value) == 0
    d.synchronize()
    d.close()
