"""Pins of the oracle's general-hexahedron path (3D), the restatement of
GridGenerator::general_cell + refine_global (include/base/diffusion_problem_basis.tpp:
92-99), FE_Q(1) with the trilinear mapping and QGauss(2) (:161-242) and BasisQ1
by monomial inversion at the mapped nodes (include/coefficients/basis_q1.tpp:26-96).

What fixes each part without the reference build:
  * a box given as eight vertices reproduces the box path up to rounding (the
    box path is pinned by closed forms in test_oracle.py);
  * affine cells: the stiffness on a parallelepiped x = x0 + B xi with tensor A
    equals the stiffness on the unit cube with det(B) B^-1 A B^-T, and the load
    is det(B) times the unit cube's (change of variables; B is a non-symmetric
    shear, so a transposed B or J fails it);
  * trilinear (non-affine) cells: the patch test -- with a constant coefficient
    the discrete harmonic extension of a linear function is that function
    (BasisQ1 reproduces 1, x, y, z and the isoparametric space holds linear
    functions), so sum_i x_k(v_i) phi_i = x_k at every fine node; then
    w_k^T A^K w_l = alpha vol delta_kl, sum_i F_i = f vol and
    sum_i x_k(v_i) F_i = f vol centroid_k, with QGauss(2) exact for all of these
    (det J has degree <= 2 per parametric variable).  Volume and centroid come
    from a 4-point Gauss-Legendre rule on det J written here in numpy;
  * the refined mesh: nodes on the straight edges, face centres the mean of the
    face's four vertices, the cell centre the mean of all eight.
"""
import numpy as np
import pytest

import msfem
import oracle as O

# A hexahedron with a twisted top face and a non-planar side: every face but the
# bottom is a doubly curved bilinear surface.
TWISTED = np.array([[0.00, 0.00, 0.00], [1.00, 0.05, 0.00], [0.05, 0.90, 0.00], [1.10, 1.00, 0.10],
                    [0.15, -0.10, 1.00], [1.05, 0.15, 0.95], [-0.05, 0.95, 1.10], [0.90, 1.05, 1.00]])


def trilinear(h, xi):
    x = np.zeros(3)
    for v in range(8):
        w = np.prod([xi[m] if (v >> m) & 1 else 1 - xi[m] for m in range(3)])
        x += w * h[v]
    return x


def jacobian(h, xi):
    J = np.zeros((3, 3))
    for v in range(8):
        for l in range(3):
            d = 1.0 if (v >> l) & 1 else -1.0
            for m in range(3):
                if m != l:
                    d *= xi[m] if (v >> m) & 1 else 1 - xi[m]
            J[:, l] += d * h[v]
    return J


def volume_centroid(h):
    """Volume and centroid of a trilinear hexahedron by 4-point Gauss-Legendre per axis."""
    t, w = np.polynomial.legendre.leggauss(4)
    t, w = (t + 1) / 2, w / 2
    vol, mom = 0.0, np.zeros(3)
    for a in range(4):
        for b in range(4):
            for c in range(4):
                xi = (t[a], t[b], t[c])
                dv = w[a] * w[b] * w[c] * np.linalg.det(jacobian(h, xi))
                vol += dv
                mom += dv * trilinear(h, xi)
    return vol, mom / vol


def test_hex_mesh_nodes():
    nf = 4
    X = O.hex_nodes(TWISTED, nf).reshape(nf + 1, nf + 1, nf + 1, 3)  # [z][y][x]
    for j in range(nf + 1):
        t = j / nf
        assert np.allclose(X[0, 0, j], TWISTED[0] + t * (TWISTED[1] - TWISTED[0]), atol=1e-15)
        assert np.allclose(X[0, j, 0], TWISTED[0] + t * (TWISTED[2] - TWISTED[0]), atol=1e-15)
        assert np.allclose(X[j, 0, 0], TWISTED[0] + t * (TWISTED[4] - TWISTED[0]), atol=1e-15)
        assert np.allclose(X[nf, nf, j], TWISTED[6] + t * (TWISTED[7] - TWISTED[6]), atol=1e-15)
    h = nf // 2
    assert np.allclose(X[h, h, h], TWISTED.mean(axis=0), atol=1e-15)
    assert np.allclose(X[0, h, h], TWISTED[[0, 1, 2, 3]].mean(axis=0), atol=1e-15)  # z = 0 face
    assert np.allclose(X[nf, h, h], TWISTED[[4, 5, 6, 7]].mean(axis=0), atol=1e-15)  # z = 1 face
    assert np.allclose(X[h, h, 0], TWISTED[[0, 2, 4, 6]].mean(axis=0), atol=1e-15)  # x = 0 face


@pytest.mark.parametrize("kind", [O.COEF_SCALAR_CELL, O.COEF_TENSOR_CELL, O.COEF_SCALAR_QP, O.COEF_TENSOR_QP])
def test_box_as_hex_matches_box_path(kind):
    nf = 4
    box = np.array([[0.25, 0.5], [0.5, 0.625], [0.0, 0.375]])
    hexa = O.box_vertices(3, box)
    rng = np.random.default_rng(1)
    stride = O.coef_stride(3, kind)
    tile = rng.uniform(0.5, 2.0, nf ** 3 * stride)
    if kind in (O.COEF_TENSOR_CELL, O.COEF_TENSOR_QP):  # SPD: small off-diagonals
        t6 = tile.reshape(-1, 6)
        t6[:, [1, 2, 4]] *= 0.1
    K0, b0 = O.local_assemble(3, nf, box, kind, tile, O.RHS_SINSIN)
    K1, b1 = O.local_assemble_hex(nf, hexa, kind, tile, O.RHS_SINSIN)
    assert np.abs(K1 - K0).max() < 1e-13 * np.abs(K0).max()
    assert np.abs(b1 - b0).max() < 1e-13 * np.abs(b0).max()
    for i in range(8):
        assert np.abs(O.dirichlet_data_hex(nf, hexa, i) - O.dirichlet_data(3, nf, box, i)).max() < 1e-13
    r0 = O.batch_build(3, box[None], [nf], kind, tile, want_bases=True)
    r1 = O.batch_build_hex(hexa[None], [nf], kind, tile, want_bases=True)
    assert np.abs(r1["bases"] - r0["bases"]).max() < 1e-12
    assert np.abs(r1["Ae"] - r0["Ae"]).max() < 1e-12 * np.abs(r0["Ae"]).max()
    assert np.abs(r1["Fe"] - r0["Fe"]).max() < 1e-12 * np.abs(r0["Fe"]).max()


def test_parallelepiped_change_of_variables():
    """K_hex(A) = K_unit(det B B^-1 A B^-T), b_hex = det B b_unit for x = x0 + B xi."""
    nf = 4
    x0 = np.array([0.2, 0.1, -0.3])
    B = np.array([[0.7, 0.3, 0.1], [0.15, 0.9, -0.2], [0.05, 0.25, 0.8]])
    hexa = np.array([x0 + B @ np.array([v & 1, (v >> 1) & 1, (v >> 2) & 1], float) for v in range(8)])
    Bi, dB = np.linalg.inv(B), np.linalg.det(B)
    rng = np.random.default_rng(2)
    iu = np.triu_indices(3)
    A, Ahat = np.zeros((nf ** 3, 6)), np.zeros((nf ** 3, 6))
    for f in range(nf ** 3):
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        M = Q @ np.diag(rng.uniform(0.5, 2, 3)) @ Q.T
        A[f] = M[iu]
        Ahat[f] = (dB * Bi @ M @ Bi.T)[iu]
    Kh, bh = O.local_assemble_hex(nf, hexa, O.COEF_TENSOR_CELL, A, O.RHS_CONST, 1.0)
    Ku, bu = O.local_assemble(3, nf, np.array([[0, 1], [0, 1], [0, 1]]), O.COEF_TENSOR_CELL, Ahat, O.RHS_CONST, 1.0)
    assert np.abs(Kh - Ku).max() < 1e-12 * np.abs(Ku).max()
    assert np.abs(bh - dB * bu).max() < 1e-13
    # a transposed B would change the stiffness
    full = lambda a: np.array([[a[0], a[1], a[2]], [a[1], a[3], a[4]], [a[2], a[4], a[5]]])
    At = np.array([(dB * Bi.T @ full(a) @ Bi)[iu] for a in A])
    Kt, _ = O.local_assemble(3, nf, np.array([[0, 1], [0, 1], [0, 1]]), O.COEF_TENSOR_CELL, At, O.RHS_CONST, 1.0)
    assert np.abs(Kh - Kt).max() > 1e-3 * np.abs(Ku).max()


@pytest.mark.parametrize("which", ["twisted", "perturbed"])
def test_patch_test_trilinear_cell(which):
    """Constant alpha on a non-affine cell: linear reproduction and exact energies and moments."""
    hexa = TWISTED if which == "twisted" else msfem.perturbed_hexes(3, 0.16, 7)[13]
    nf, alpha, f = 6, 1.7, 2.0
    vol, cen = volume_centroid(hexa)
    tile = np.full(nf ** 3, alpha)
    K, b = O.local_assemble_hex(nf, hexa, O.COEF_SCALAR_CELL, tile, O.RHS_CONST, f)
    assert np.abs(K.sum(axis=1)).max() < 1e-13 * np.abs(K).max()  # constants in the kernel
    phi = O.local_bases_direct_hex(nf, hexa, K)
    X = O.hex_nodes(hexa, nf)
    assert np.abs(phi.sum(axis=0) - 1).max() < 1e-13
    for k in range(3):
        assert np.abs(hexa[:, k] @ phi - X[:, k]).max() < 1e-13
    Ae, Fe = O.element(3, nf, K, b, phi)
    for k in range(3):
        for l in range(3):
            want = alpha * vol if k == l else 0.0
            assert abs(hexa[:, k] @ Ae @ hexa[:, l] - want) < 1e-12
    assert abs(Fe.sum() - f * vol) < 1e-13
    for k in range(3):
        assert abs(hexa[:, k] @ Fe - f * vol * cen[k]) < 1e-13
    assert np.allclose(Ae, Ae.T, atol=1e-14) and np.abs(Ae.sum(axis=1)).max() < 1e-12


def test_hex_dirichlet_is_basis_q1():
    """Boundary data = BasisQ1 (monomial inversion) at the mapped nodes; interior 0."""
    nf = 4
    X = O.hex_nodes(TWISTED, nf)
    C = O.basis_q1_coeffs(3, TWISTED)
    bnd = np.ones((nf + 1,) * 3, bool)
    bnd[1:-1, 1:-1, 1:-1] = False
    bnd = bnd.reshape(-1)
    for i in range(8):
        g = O.dirichlet_data_hex(nf, TWISTED, i)
        ref = np.array([O.basis_q1_value(3, C, i, p) for p in X])
        assert np.abs(g[~bnd]).max() == 0
        assert np.abs(g[bnd] - ref[bnd]).max() < 1e-14
        corners = g.reshape((nf + 1,) * 3)[np.ix_([0, nf], [0, nf], [0, nf])].reshape(-1)  # [z][y][x]
        assert np.allclose(corners, np.eye(8)[i], atol=1e-13)  # N_i(v_j) = delta_ij, v = x + 2y + 4z


def test_perturbed_hex_mesh_generator():
    """Conforming mesh of the unit cube: shared vertices agree, total volume 1, boundary fixed."""
    nc = 3
    h = msfem.perturbed_hexes(nc, 0.16, 3)
    assert abs(sum(volume_centroid(c)[0] for c in h) - 1.0) < 1e-13
    for c in range(nc ** 3):
        cx, cy, cz = c % nc, (c // nc) % nc, c // (nc * nc)
        if cx + 1 < nc:
            assert np.array_equal(h[c][[1, 3, 5, 7]], h[c + 1][[0, 2, 4, 6]])
        if cy + 1 < nc:
            assert np.array_equal(h[c][[2, 3, 6, 7]], h[c + nc][[0, 1, 4, 5]])
        if cz + 1 < nc:
            assert np.array_equal(h[c][[4, 5, 6, 7]], h[c + nc * nc][[0, 1, 2, 3]])
    assert h.min() >= 0 and h.max() <= 1
    assert not np.array_equal(h, msfem.perturbed_hexes(nc, 0.0, 3))
    box = msfem.unit_boxes(3, nc)
    assert np.array_equal(msfem.perturbed_hexes(nc, 0.0, 3)[:, 0], box[:, :, 0])
    with pytest.raises(msfem.MsfemError):
        msfem.perturbed_hexes(nc, 1.0 / 6.0, 3)
