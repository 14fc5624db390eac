"""Host-side serving bookkeeping (no GPU): the HBM footprint used for admission
and the DoF count of a request, against closed forms."""
import philip_amd as P


def test_request_dofs_and_bytes_closed_form():
    for dim, p, n in [(2, 3, 8), (2, 6, 5), (3, 2, 4), (3, 5, 3)]:
        r = P.request(1.0, 0.1, dim=dim, poly_degree=p, n_cells_1d=n)
        np_ = (p + 1) ** dim
        ncells = n ** dim
        ndofs = ncells * np_ * (dim + 2)
        assert P.load().dgx_request_n_dofs(r) == ndofs
        # solution + residual + RK4 registers (accumulator, two stage buffers, four k slots) + connectivity
        assert P.request_bytes(r, "rk4") == 8 * ndofs * (2 + 3 + 4) + 4 * ncells * 2 * dim
        assert P.request_bytes(r, "heun") == 8 * ndofs * (2 + 3 + 2) + 4 * ncells * 2 * dim
    ns = P.request(1.0, 0.1, dim=3, poly_degree=2, n_cells_1d=4, pde=P.NAVIER_STOKES)
    nd = 4 ** 3 * 27 * 5
    assert P.request_bytes(ns, "rk4") == 8 * nd * (2 + 7 + 3) + 4 * 64 * 6
