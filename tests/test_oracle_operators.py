"""Oracle 1D operators, sum factorisation and Hadamard sparsity, pinned by
closed forms and brute force (the checks of the reference's
tests/unit_tests/operator_tests/volume_operators_test.cpp:135-149,
sum_factorization_test.cpp:140-200 and operators.cpp:949-1133 restated)."""
import numpy as np
import pytest

import oracle as O


@pytest.mark.parametrize("n", range(2, 12))
def test_gauss_legendre_matches_numpy(n):
    x, w = O.gauss_legendre(n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert np.allclose(x, 0.5 * (xr + 1.0), atol=1e-15, rtol=0)
    assert np.allclose(w, 0.5 * wr, atol=1e-15, rtol=0)


def test_gauss_lobatto_closed_forms():
    x, w = O.gauss_lobatto(3)
    assert np.allclose(x, [0, 0.5, 1], atol=1e-16) and np.allclose(w, [1 / 6, 2 / 3, 1 / 6], atol=1e-16)
    x, w = O.gauss_lobatto(4)
    a = 0.5 * (1 - 1 / np.sqrt(5))
    assert np.allclose(x, [0, a, 1 - a, 1], atol=1e-15) and np.allclose(w, [1 / 12, 5 / 12, 5 / 12, 1 / 12], atol=1e-15)
    x, w = O.gauss_lobatto(5)
    b = 0.5 * (1 - np.sqrt(3 / 7))
    assert np.allclose(x, [0, b, 0.5, 1 - b, 1], atol=1e-15)
    assert np.allclose(w, [1 / 20, 49 / 180, 32 / 90, 49 / 180, 1 / 20], atol=1e-15)


@pytest.mark.parametrize("n", range(2, 13))
def test_gauss_lobatto_exactness(n):
    x, w = O.gauss_lobatto(n)
    assert abs(w.sum() - 1) < 1e-14
    for k in range(2 * n - 2):  # exact up to degree 2n-3
        assert abs(np.dot(w, x ** k) - 1.0 / (k + 1)) < 1e-13


@pytest.mark.parametrize("nodes", [O.GLL, O.GL])
@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 6])
def test_operator_identities(nodes, p):
    o = O.Oracle(dim=3, poly_degree=p, flux_nodes=nodes, n_cells_1d=2)
    ops = o.operators()
    x, w, D = ops["nodes"], ops["weights"], ops["flux_grad"]
    n = p + 1
    assert np.abs(D @ np.ones(n)).max() < 1e-12          # D 1 = 0
    assert np.abs(D @ x - 1.0).max() < 1e-12              # D x = 1
    Q = np.diag(w) @ D
    B = np.zeros((n, n))
    # SBP: Q + Q^T = B (boundary operator ell ell^T evaluated at the ends)
    lf = ops["flux_surf"]
    B = np.outer(lf[1], lf[1]) - np.outer(lf[0], lf[0])
    assert np.abs(Q + Q.T - B).max() < 1e-12
    assert np.abs(ops["skew"] - (Q - Q.T)).max() < 1e-12
    # projection is a left inverse of interpolation, mass inverse matches chi^T W chi
    chi = ops["chi"]
    assert np.abs(ops["proj"] @ chi - np.eye(n)).max() < 1e-12
    M = chi.T @ np.diag(w) @ chi
    assert np.abs(ops["minv"] @ M - np.eye(n)).max() < 1e-11
    if nodes == O.GLL:
        assert np.abs(chi - np.eye(n)).max() < 1e-15      # collocated


@pytest.mark.parametrize("dim", [2, 3])
def test_sum_factorization_vs_kron(dim):
    rng = np.random.default_rng(7)
    for (rx, cx, ry, cy, rz, cz) in [(3, 4, 2, 4, 5, 4), (4, 4, 4, 4, 4, 4), (1, 5, 5, 5, 5, 5)]:
        Ax, Ay, Az = rng.standard_normal((rx, cx)), rng.standard_normal((ry, cy)), rng.standard_normal((rz, cz))
        if dim == 2:
            x = rng.standard_normal(cx * cy)
            ref = np.kron(Ay, Ax) @ x
            got = O.tensor_apply(2, Ax, Ay, None, x)
        else:
            x = rng.standard_normal(cx * cy * cz)
            ref = np.kron(Az, np.kron(Ay, Ax)) @ x
            got = O.tensor_apply(3, Ax, Ay, Az, x)
        assert np.abs(got - ref).max() < 1e-11 * max(1, np.abs(ref).max())


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("n", [2, 3, 4, 6])
def test_hadamard_volume_pattern_bit_exact(dim, n):
    rows, cols = O.hadamard_volume_pattern(dim, n)
    # independent construction: row = node, col(dir) = node with coordinate dir replaced by l
    er, ec = [], []
    for node in range(n ** dim):
        idx = list(np.unravel_index(node, (n,) * dim)[::-1])  # x fastest
        for l in range(n):
            er.append([node] * dim)
            cc = []
            for d in range(dim):
                j = list(idx)
                j[d] = l
                cc.append(int(np.ravel_multi_index(j[::-1], (n,) * dim)))
            ec.append(cc)
    assert np.array_equal(rows, np.array(er, dtype=np.uint32))
    assert np.array_equal(cols, np.array(ec, dtype=np.uint32))


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("n", [2, 3, 5])
def test_hadamard_surface_pattern_bit_exact(dim, n):
    for dnz in range(dim):
        rows, cols = O.hadamard_surface_pattern(dim, n, dnz)
        tang = [d for d in range(dim) if d != dnz]
        er, ec = [], []
        for f in range(n ** (dim - 1)):
            t = list(np.unravel_index(f, (n,) * (dim - 1))[::-1])
            for l in range(n):
                idx = [0] * dim
                for k, d in enumerate(tang):
                    idx[d] = t[k]
                idx[dnz] = l
                er.append(f)
                ec.append(int(np.ravel_multi_index(idx[::-1], (n,) * dim)))
        assert np.array_equal(rows, np.array(er, dtype=np.uint32))
        assert np.array_equal(cols, np.array(ec, dtype=np.uint32))


@pytest.mark.parametrize("dim,N", [(3, 4), (3, 8), (2, 8), (3, 3), (2, 6)])
def test_mesh_connectivity(dim, N):
    o = O.Oracle(dim=dim, poly_degree=1, n_cells_1d=N)
    ijk = o.cell_ijk()
    nbr = o.neighbors()
    # every cell appears once
    lin = ijk[:, 0] + N * ijk[:, 1] + (N * N * ijk[:, 2] if dim == 3 else 0)
    assert np.array_equal(np.sort(lin), np.arange(N ** dim))
    if N & (N - 1) == 0:  # refine_global order: Morton, x bit lowest
        levels = int(np.log2(N))
        for e in range(N ** dim):
            code = 0
            for lvl in range(levels):
                for k in range(dim):
                    code |= ((int(ijk[e, k]) >> lvl) & 1) << (lvl * dim + k)
            assert code == e
    else:  # subdivided_hyper_cube: lexicographic
        assert np.array_equal(lin, np.arange(N ** dim))
    for e in range(N ** dim):
        for f in range(2 * dim):
            k = f // 2
            j = ijk[e].copy()
            j[k] = (j[k] + (1 if f % 2 else -1)) % N
            assert np.array_equal(ijk[nbr[e, f]], j)
            assert nbr[nbr[e, f], f ^ 1] == e
    faces = o.face_list()
    assert len(faces) == dim * N ** dim
    assert np.all(faces[:, 0] < faces[:, 2])
