"""Pins of the oracle's general-quadrilateral path (2D), the restatement of
GridGenerator::general_cell + refine_global (include/base/diffusion_problem_basis.tpp:
92-99), FE_Q(1) with the bilinear mapping and QGauss(2) (:161-242) and BasisQ1
by monomial inversion at the mapped nodes (include/coefficients/basis_q1.tpp:26-96).

What fixes each part without the reference build:
  * a box given as four vertices reproduces the box path bit for bit up to
    rounding (the box path is pinned by closed forms in test_oracle.py);
  * affine cells: the stiffness on a parallelogram with tensor A equals the
    stiffness on the unit square with det(J) J^-1 A J^-T (change of variables;
    a transposed J fails it, the test uses a shear);
  * bilinear cells: the patch test -- with a constant coefficient the discrete
    harmonic extension of a linear function is that function, so
    sum_i x_i phi_i = x and sum_i y_i phi_i = y at every fine node (BasisQ1
    reproduces 1, x, y, xy and the isoparametric space holds linear
    functions); then w_x^T A^K w_x = alpha * area, w_x^T A^K w_y = 0 and
    sum_i F_i = f * area, sum_i x_i F_i = f * area * centroid_x (shoelace);
  * the refined mesh: boundary nodes lie on the straight edges, the central
    node is the mean of the four vertices.
"""
import numpy as np
import pytest

import msfem
import oracle as O


def shoelace(q):
    """Area and centroid of a quadrilateral in deal.II vertex order."""
    p = q[[0, 1, 3, 2]]  # counter-clockwise
    x, y = p[:, 0], p[:, 1]
    xn, yn = np.roll(x, -1), np.roll(y, -1)
    cr = x * yn - xn * y
    area = cr.sum() / 2
    return area, np.array([((x + xn) * cr).sum(), ((y + yn) * cr).sum()]) / (6 * area)


TRAPEZOID = np.array([[0.1, 0.2], [0.9, 0.1], [0.25, 0.8], [0.7, 1.05]])


def test_quad_mesh_nodes():
    nf = 8
    x = O.quad_nodes(TRAPEZOID, nf).reshape(nf + 1, nf + 1, 2)
    for j in range(nf + 1):
        t = j / nf
        assert np.allclose(x[0, j], TRAPEZOID[0] + t * (TRAPEZOID[1] - TRAPEZOID[0]), atol=1e-15)  # y = 0 edge
        assert np.allclose(x[j, 0], TRAPEZOID[0] + t * (TRAPEZOID[2] - TRAPEZOID[0]), atol=1e-15)  # x = 0 edge
        assert np.allclose(x[nf, j], TRAPEZOID[2] + t * (TRAPEZOID[3] - TRAPEZOID[2]), atol=1e-15)
        assert np.allclose(x[j, nf], TRAPEZOID[1] + t * (TRAPEZOID[3] - TRAPEZOID[1]), atol=1e-15)
    assert np.allclose(x[nf // 2, nf // 2], TRAPEZOID.mean(axis=0), atol=1e-15)


@pytest.mark.parametrize("kind", [O.COEF_SCALAR_CELL, O.COEF_TENSOR_CELL, O.COEF_TENSOR_QP])
def test_box_as_quad_matches_box_path(kind):
    nf = 8
    box = np.array([[0.25, 0.5], [0.5, 0.625]])
    quad = O.box_vertices(2, box)
    rng = np.random.default_rng(1)
    stride = O.coef_stride(2, kind)
    tile = rng.uniform(0.5, 2.0, nf * nf * stride)
    if kind != O.COEF_SCALAR_CELL:  # keep the tensors SPD: small off-diagonals
        tile.reshape(-1, 3)[:, 1] *= 0.1
    K0, b0 = O.local_assemble(2, nf, box, kind, tile, O.RHS_SINSIN)
    K1, b1 = O.local_assemble_quad(nf, quad, kind, tile, O.RHS_SINSIN)
    assert np.abs(K1 - K0).max() < 1e-13 * np.abs(K0).max()
    assert np.abs(b1 - b0).max() < 1e-13 * np.abs(b0).max()
    for i in range(4):
        assert np.abs(O.dirichlet_data_quad(nf, quad, i) - O.dirichlet_data(2, nf, box, i)).max() < 1e-13
    r0 = O.batch_build(2, box[None], [nf], kind, tile, want_bases=True)
    r1 = O.batch_build_quad(quad[None], [nf], kind, tile, want_bases=True)
    assert np.abs(r1["bases"] - r0["bases"]).max() < 1e-12
    assert np.abs(r1["Ae"] - r0["Ae"]).max() < 1e-12 * np.abs(r0["Ae"]).max()


def test_parallelogram_change_of_variables():
    """Shear + stretch: K_quad(A) = K_unit(det J J^-1 A J^-T), b_quad = det J b_unit."""
    nf = 8
    v0, a, b = np.array([0.2, 0.1]), np.array([0.7, 0.15]), np.array([0.3, 0.9])
    quad = np.array([v0, v0 + a, v0 + b, v0 + a + b])
    J = np.column_stack([a, b])
    Ji, dJ = np.linalg.inv(J), np.linalg.det(J)
    rng = np.random.default_rng(2)
    A = np.zeros((nf * nf, 3))
    Ahat = np.zeros((nf * nf, 3))
    for f in range(nf * nf):
        th, l1, l2 = rng.uniform(0, np.pi), rng.uniform(0.5, 2), rng.uniform(0.5, 2)
        R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
        M = R @ np.diag([l1, l2]) @ R.T
        Mh = dJ * Ji @ M @ Ji.T
        A[f] = M[0, 0], M[0, 1], M[1, 1]
        Ahat[f] = Mh[0, 0], Mh[0, 1], Mh[1, 1]
    Kq, bq = O.local_assemble_quad(nf, quad, O.COEF_TENSOR_CELL, A, O.RHS_CONST, 1.0)
    Ku, bu = O.local_assemble(2, nf, np.array([[0, 1], [0, 1]]), O.COEF_TENSOR_CELL, Ahat, O.RHS_CONST, 1.0)
    assert np.abs(Kq - Ku).max() < 1e-12 * np.abs(Ku).max()
    assert np.abs(bq - dJ * bu).max() < 1e-13


@pytest.mark.parametrize("quad", [TRAPEZOID, msfem.perturbed_quads(4, 0.2, 7)[5]])
def test_patch_test_bilinear_cell(quad):
    """Constant alpha on a non-affine cell: linear reproduction and exact energies."""
    nf, alpha, f = 16, 1.7, 2.0
    area, cen = shoelace(quad)
    tile = np.full(nf * nf, alpha)
    K, b = O.local_assemble_quad(nf, quad, O.COEF_SCALAR_CELL, tile, O.RHS_CONST, f)
    assert np.abs(K.sum(axis=1)).max() < 1e-13 * np.abs(K).max()  # constants in the kernel
    phi = O.local_bases_direct_quad(nf, quad, K)
    X = O.quad_nodes(quad, nf)
    assert np.abs(phi.sum(axis=0) - 1).max() < 1e-13
    assert np.abs(quad[:, 0] @ phi - X[:, 0]).max() < 1e-13
    assert np.abs(quad[:, 1] @ phi - X[:, 1]).max() < 1e-13
    Ae, Fe = O.element(2, nf, K, b, phi)
    wx, wy = quad[:, 0], quad[:, 1]
    assert abs(wx @ Ae @ wx - alpha * area) < 1e-12
    assert abs(wy @ Ae @ wy - alpha * area) < 1e-12
    assert abs(wx @ Ae @ wy) < 1e-12
    assert abs(Fe.sum() - f * area) < 1e-13
    assert abs(wx @ Fe - f * area * cen[0]) < 1e-13
    assert abs(wy @ Fe - f * area * cen[1]) < 1e-13
    assert np.allclose(Ae, Ae.T, atol=1e-14) and np.abs(Ae.sum(axis=1)).max() < 1e-12


def test_quad_dirichlet_is_basis_q1():
    """Boundary data = BasisQ1 (monomial inversion) at the mapped nodes; interior 0."""
    nf = 8
    X = O.quad_nodes(TRAPEZOID, nf)
    C = O.basis_q1_coeffs(2, TRAPEZOID)
    for i in range(4):
        g = O.dirichlet_data_quad(nf, TRAPEZOID, i).reshape(nf + 1, nf + 1)
        ref = np.array([O.basis_q1_value(2, C, i, p) for p in X]).reshape(nf + 1, nf + 1)
        assert np.abs(g[1:-1, 1:-1]).max() == 0
        bnd = np.ones_like(g, bool)
        bnd[1:-1, 1:-1] = False
        assert np.abs(g[bnd] - ref[bnd]).max() < 1e-15
        # N_i(v_j) = delta_ij at the four corners
        corners = [g[0, 0], g[0, nf], g[nf, 0], g[nf, nf]]
        assert np.allclose(corners, np.eye(4)[i], atol=1e-14)


def test_perturbed_mesh_generator():
    """Conforming mesh of the unit square: shared vertices agree, total area 1, boundary fixed."""
    nc = 6
    q = msfem.perturbed_quads(nc, 0.24, 3)
    assert abs(sum(shoelace(c)[0] for c in q) - 1.0) < 1e-13
    for c in range(nc * nc):
        cx, cy = c % nc, c // nc
        if cx + 1 < nc:
            assert np.array_equal(q[c][[1, 3]], q[c + 1][[0, 2]])
        if cy + 1 < nc:
            assert np.array_equal(q[c][[2, 3]], q[c + nc][[0, 1]])
    assert q.min() >= 0 and q.max() <= 1
    assert np.array_equal(msfem.perturbed_quads(nc, 0.0, 3).reshape(-1, 4, 2)[:, 0],
                          msfem.unit_boxes(2, nc)[:, :, 0])
