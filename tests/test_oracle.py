"""Pins the CPU oracle (oracle/msfem_oracle.c) before the GPU tests compare against it.

Each test fixes a piece of the reference's local pipeline by something other than
the oracle's own formula: golden vectors from an independent numpy/scipy
restatement (tests/golden), closed forms, invariants and hand computations.
"""
import glob
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "*.npz"))))
def test_golden_fixtures(path):
    g = np.load(path)
    dim, nf = int(g["dim"]), int(g["nf"])
    K, b = O.local_assemble(dim, nf, g["box"], int(g["coef_kind"]), g["coef"], int(g["rhs_kind"]),
                            float(g["rhs_value"]))
    phi = O.local_bases_direct(dim, nf, g["box"], K)
    Ae, Fe = O.element(dim, nf, K, b, phi)
    assert np.abs(phi - g["bases"]).max() < 1e-12
    assert rel(Ae, g["Ae"]) < 1e-12
    assert rel(Fe, g["Fe"]) < 1e-12


@pytest.mark.parametrize("dim", [2, 3])
def test_basis_q1_monomial_inversion_is_tensor_hat(dim):
    rng = np.random.default_rng(dim)
    for _ in range(5):
        lo = rng.uniform(-1, 1, dim)
        hi = lo + rng.uniform(0.1, 2, dim)
        box = np.stack([lo, hi], axis=1)
        V = O.box_vertices(dim, box)
        C = O.basis_q1_coeffs(dim, V)
        for j in range(1 << dim):  # N_i(v_j) = delta_ij
            for i in range(1 << dim):
                assert abs(O.basis_q1_value(dim, C, i, V[j]) - (i == j)) < 1e-12
        for _ in range(10):
            p = lo + rng.uniform(0, 1, dim) * (hi - lo)
            t = (p - lo) / (hi - lo)
            vals = [O.basis_q1_value(dim, C, i, p) for i in range(1 << dim)]
            hat = [np.prod([t[k] if (i >> k) & 1 else 1 - t[k] for k in range(dim)]) for i in range(1 << dim)]
            assert np.allclose(vals, hat, atol=1e-12)
            assert abs(sum(vals) - 1) < 1e-12


def test_fine_element_2d_closed_form():
    """QGauss(2) element matrix of a constant tensor on an hx x hy box equals
    a*rho*Kxx + c/rho*Kyy + b*Kxy (rho = hy/hx), Kxy = (s t^T + t s^T)/4."""
    Kxx = np.array([[2, -2, 1, -1], [-2, 2, -1, 1], [1, -1, 2, -2], [-1, 1, -2, 2]]) / 6.0
    P = [0, 2, 1, 3]
    Kyy = Kxx[np.ix_(P, P)]
    s = np.array([-1, 1, -1, 1.0])
    t = np.array([-1, -1, 1, 1.0])
    Kxy = (np.outer(s, t) + np.outer(t, s)) / 4
    for (a, b, c, hx, hy) in [(1, 0, 1, 0.1, 0.1), (2.5, 0.7, 1.3, 0.2, 0.05), (0.3, -0.2, 4.0, 1.0, 3.0)]:
        Ke, be = O.fine_element(2, [0.3, 0.1], [hx, hy], O.COEF_TENSOR_CELL, [a, b, c], O.RHS_CONST, 2.0)
        rho = hy / hx
        assert np.allclose(Ke, a * rho * Kxx + c / rho * Kyy + b * Kxy, rtol=1e-13, atol=1e-14)
        assert np.allclose(be, 2.0 * hx * hy / 4)


def test_fine_element_3d_cube_closed_form():
    """Trilinear Q1 on a cube of side h: diagonal h/3, edges 0, face and body
    diagonals -h/12."""
    h = 0.125
    Ke, _ = O.fine_element(3, [0, 0, 0], [h, h, h], O.COEF_SCALAR_CELL, [1.0])
    for a in range(8):
        for b in range(8):
            d = bin(a ^ b).count("1")
            want = {0: h / 3, 1: 0.0, 2: -h / 12, 3: -h / 12}[d]
            assert abs(Ke[a, b] - want) < 1e-15


@pytest.mark.parametrize("dim,nf", [(2, 8), (3, 4)])
def test_constant_coefficient_gives_q1_bases(dim, nf):
    """Constant diagonal A: the multilinear N_i is discretely harmonic, so phi_i = N_i."""
    box = np.array([[0.25, 0.5], [0.5, 0.75], [0.0, 0.25]][:dim])
    tile = np.tile([3.0, 0.0, 0.5] if dim == 2 else [3.0, 0, 0, 0.5, 0, 2.0], nf ** dim)
    K, b = O.local_assemble(dim, nf, box, O.COEF_TENSOR_CELL, tile)
    phi = O.local_bases_direct(dim, nf, box, K)
    np1 = nf + 1
    I = np.array(np.unravel_index(np.arange(np1 ** dim), [np1] * dim, order="F")).T / nf
    for i in range(1 << dim):
        N = np.prod([I[:, k] if (i >> k) & 1 else 1 - I[:, k] for k in range(dim)], axis=0)
        assert np.abs(phi[i] - N).max() < 1e-13
    # and A^K is the standard Q1 element matrix of the coarse box
    Ae, _ = O.element(dim, nf, K, b, phi)
    Kc, _ = O.fine_element(dim, box[:, 0], box[:, 1] - box[:, 0], O.COEF_TENSOR_CELL, tile[:3 if dim == 2 else 6])
    assert np.allclose(Ae, Kc, rtol=1e-12, atol=1e-13)


def test_single_interior_node_by_hand():
    """nf = 2: one interior node, phi_i(center) = -sum_j K_cj g_j / K_cc."""
    box = np.array([[0.0, 1.0], [0.0, 1.0]])
    tile = np.array([1.0, 2.0, 3.0, 4.0])
    K, _ = O.local_assemble(2, 2, box, O.COEF_SCALAR_CELL, tile)
    A = O.stencil_to_dense(2, 2, K)
    phi = O.local_bases_direct(2, 2, box, K)
    c = 4
    for i in range(4):
        g = O.dirichlet_data(2, 2, box, i)
        want = -sum(A[c, j] * g[j] for j in range(9) if j != c) / A[c, c]
        assert abs(phi[i][c] - want) < 1e-15
    # hand value for basis 0: alpha weights the four cells
    assert 0 < phi[0][c] < 0.25 + 1e-15


def test_lognormal_cell_invariants():
    rng = np.random.default_rng(7)
    nf = 32
    box = np.array([[0.0, 0.125], [0.375, 0.5]])
    tile = np.exp(np.clip(rng.standard_normal(nf * nf), -3.45, 3.45))
    K, b = O.local_assemble(2, nf, box, O.COEF_SCALAR_CELL, tile, O.RHS_CONST, 1.0)
    phi = O.local_bases_direct(2, nf, box, K)
    assert np.abs(phi.sum(0) - 1).max() < 1e-12                  # partition of unity
    assert phi.min() > -1e-12 and phi.max() < 1 + 1e-12           # discrete maximum principle (M-matrix)
    Ae, Fe = O.element(2, nf, K, b, phi)
    assert np.abs(Ae - Ae.T).max() < 1e-14 * np.abs(Ae).max()      # symmetric
    assert np.abs(Ae.sum(1)).max() < 1e-12                          # constants in the kernel
    ev = np.linalg.eigvalsh(Ae)
    assert ev[0] > -1e-12 and (ev[1:] > 1e-6).all()                 # PSD, rank nb - 1
    assert abs(Fe.sum() - 0.125 * 0.125) < 1e-14                     # sum_i F_i = int_K f
    A = O.stencil_to_dense(2, nf, K)
    for i in range(4):                                               # energy minimality
        N = O.dirichlet_data(2, nf, box, i)
        np1 = nf + 1
        X, Y = np.arange(np1 ** 2) % np1 / nf, np.arange(np1 ** 2) // np1 / nf
        N = ((X if i & 1 else 1 - X) * (Y if i & 2 else 1 - Y))
        assert Ae[i, i] <= N @ A @ N + 1e-14


def test_ssor_pcg_mirror_matches_direct():
    rng = np.random.default_rng(3)
    nf = 32
    box = np.array([[0.5, 0.625], [0.5, 0.625]])
    tile = np.exp(np.clip(rng.standard_normal(nf * nf), -3.45, 3.45))
    K, _ = O.local_assemble(2, nf, box, O.COEF_SCALAR_CELL, tile)
    phi = O.local_bases_direct(2, nf, box, K)
    phi2, its, fails = O.local_bases_ssor_pcg(2, nf, box, K, 1.6, 1e-12, 1000)
    assert fails == 0
    assert (its > 5).all() and (its < 1000).all()
    assert np.abs(phi2 - phi).max() < 1e-9
    # iteration cap: the reference would throw SolverControl::NoConvergence
    _, its3, fails3 = O.local_bases_ssor_pcg(2, nf, box, K, 1.6, 1e-12, 3)
    assert fails3 == 4 and (its3 == 3).all()


def q1_global(nc, alpha, f):
    """Textbook coarse Q1 FEM on the unit square, constant alpha, f const, u=0 on the boundary."""
    box = np.array([[0.0, 1.0], [0.0, 1.0]])
    K, b = O.local_assemble(2, nc, box, O.COEF_SCALAR_CELL, np.full(nc * nc, alpha), O.RHS_CONST, f)
    A = O.stencil_to_dense(2, nc, K)
    np1 = nc + 1
    X, Y = np.arange(np1 ** 2) % np1, np.arange(np1 ** 2) // np1
    free = (X > 0) & (X < nc) & (Y > 0) & (Y < nc)
    u = np.zeros(np1 ** 2)
    u[free] = np.linalg.solve(A[np.ix_(free, free)], b[free])
    return u


def test_coarse_msfem_constant_coefficient_equals_q1():
    nc, nf = 4, 8
    box = np.array([[[i / nc, (i + 1) / nc], [j / nc, (j + 1) / nc]] for j in range(nc) for i in range(nc)])
    coef = np.full(nc * nc * nf * nf, 2.5)
    r = O.batch_build(2, box, np.full(nc * nc, nf), O.COEF_SCALAR_CELL, coef, O.RHS_CONST, 1.0)
    Kc, Fc, cm, cv = O.coarse_assemble(2, nc, r["Ae"], r["Fe"], 0)
    u = O.coarse_solve(2, nc, Kc, Fc, cm, cv)
    assert np.abs(u - q1_global(nc, 2.5, 1.0)).max() < 1e-13


def test_coarse_msfem_manufactured_convergence():
    """alpha = 1, f = sin(pi x) sin(pi y): u = f / (2 pi^2); nodal error O(H^2)."""
    errs = []
    for nc in (4, 8, 16):
        nf = 16
        box = np.array([[[i / nc, (i + 1) / nc], [j / nc, (j + 1) / nc]] for j in range(nc) for i in range(nc)])
        r = O.batch_build(2, box, np.full(nc * nc, nf), O.COEF_SCALAR_CELL, np.ones(nc * nc * nf * nf),
                          O.RHS_SINSIN, 0.0)
        Kc, Fc, cm, cv = O.coarse_assemble(2, nc, r["Ae"], r["Fe"], 0)
        u = O.coarse_solve(2, nc, Kc, Fc, cm, cv)
        np1 = nc + 1
        X, Y = np.arange(np1 ** 2) % np1 / nc, np.arange(np1 ** 2) // np1 / nc
        errs.append(np.abs(u - np.sin(np.pi * X) * np.sin(np.pi * Y) / (2 * np.pi ** 2)).max())
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert (rates > 1.8).all(), (errs, rates)


def test_coarse_reference_bcs_dirichlet_and_neumann():
    """Reference default boundary data: u = g on x=0, y=0; with alpha = 1 and no
    load, the coarse problem is the standard Q1 problem with those data."""
    nc, nf = 4, 8
    box = np.array([[[i / nc, (i + 1) / nc], [j / nc, (j + 1) / nc]] for j in range(nc) for i in range(nc)])
    r = O.batch_build(2, box, np.full(nc * nc, nf), O.COEF_SCALAR_CELL, np.ones(nc * nc * nf * nf), O.RHS_CONST, 0.0)
    Kc, Fc, cm, cv = O.coarse_assemble(2, nc, r["Ae"], r["Fe"], 1)
    np1 = nc + 1
    X, Y = np.arange(np1 ** 2) % np1 / nc, np.arange(np1 ** 2) // np1 / nc
    assert (cm == ((X == 0) | (Y == 0))).all()
    assert np.allclose(cv[cm == 1], ((X - 0.5) ** 2 + (Y - 0.5) ** 2)[cm == 1])
    # Neumann load on x = 1: int h N_i over the face, h = cos(2 PI_D x) cos(2 PI_D y)
    pid = 3.14592653509793218403
    gs = np.array([0.5 - 0.5 / np.sqrt(3), 0.5 + 0.5 / np.sqrt(3)])
    node = nc + np1 * 2  # (x=1, y=2/nc): shared by faces of cells (nc-1,1) and (nc-1,2)
    want = 0.0
    for cy in (1, 2):
        for q in gs:
            y = (cy + q) / nc
            psi = q if cy == 1 else 1 - q
            want += np.cos(2 * pid) * np.cos(2 * pid * y) * psi * 0.5 / nc
    assert abs(Fc[node] - want) < 1e-15
    u = O.coarse_solve(2, nc, Kc, Fc, cm, cv)
    assert np.allclose(u[cm == 1], cv[cm == 1])


def unit_boxes3(nc):
    return np.array([[[i / nc, (i + 1) / nc], [j / nc, (j + 1) / nc], [k / nc, (k + 1) / nc]]
                     for k in range(nc) for j in range(nc) for i in range(nc)])


def test_coarse_3d_constant_coefficient_is_textbook_trilinear_stencil():
    """3D coarse assembly (diffusion_problem_ms.tpp:161-255 at dim 3) with a
    constant coefficient: phi_i = N_i, so the coarse matrix is the textbook
    trilinear Q1 Laplacian of the (nc+1)^3 grid -- at an interior node the
    27-point stencil alpha H (8/3 centre, 0 face, -1/6 edge, -1/12 corner
    neighbours) -- and the load is f H^3 / 8 per adjacent cell.  Zero
    Dirichlet marks every boundary node."""
    nc, nf, alpha = 3, 4, 2.5
    box = unit_boxes3(nc)
    r = O.batch_build(3, box, np.full(nc ** 3, nf), O.COEF_SCALAR_CELL, np.full(nc ** 3 * nf ** 3, alpha),
                      O.RHS_CONST, 1.0)
    Kc, Fc, cm, cv = O.coarse_assemble(3, nc, r["Ae"], r["Fe"], 0)
    H = 1.0 / nc
    np1 = nc + 1
    for node in range(np1 ** 3):
        I = np.array([node % np1, (node // np1) % np1, node // (np1 * np1)])
        for t in range(27):
            o = np.array([t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1])
            J = I + o
            if (J < 0).any() or (J > nc).any():
                assert Kc[node, t] == 0
                continue
            # cells shared by node I and node J: per axis, offset 0 -> the cells
            # on either side that exist, offset +-1 -> the one between them
            n_shared = np.prod([(int(I[m] > 0) + int(I[m] < nc)) if o[m] == 0 else 1 for m in range(3)])
            per_cell = {0: H / 3, 1: 0.0, 2: -H / 12, 3: -H / 12}[int(np.abs(o).sum())]
            assert abs(Kc[node, t] - alpha * n_shared * per_cell) < 1e-14, (node, t)
        n_adj = np.prod([int(I[m] > 0) + int(I[m] < nc) for m in range(3)])
        assert abs(Fc[node] - n_adj * H ** 3 / 8) < 1e-15
        assert cm[node] == int((I == 0).any() or (I == nc).any())
    interior = 1 + np1 + np1 * np1
    assert np.allclose(Kc[interior].reshape(3, 3, 3)[1], alpha * H * np.array([[-1 / 6, 0, -1 / 6], [0, 8 / 3, 0],
                                                                                  [-1 / 6, 0, -1 / 6]]), atol=1e-15)
    assert abs(Kc[interior].sum()) < 1e-14


def test_coarse_3d_manufactured_convergence():
    """alpha = 1, f = sin(pi x) sin(pi y) sin(pi z), zero Dirichlet: u = f / (3 pi^2);
    the 3D MsFEM coarse solution's nodal error falls at O(H^2)."""
    errs = []
    for nc in (2, 4, 8):
        nf = 4
        r = O.batch_build(3, unit_boxes3(nc), np.full(nc ** 3, nf), O.COEF_SCALAR_CELL, np.ones(nc ** 3 * nf ** 3),
                          O.RHS_SINSIN, 0.0)
        Kc, Fc, cm, cv = O.coarse_assemble(3, nc, r["Ae"], r["Fe"], 0)
        u = O.coarse_solve(3, nc, Kc, Fc, cm, cv)
        np1 = nc + 1
        idx = np.arange(np1 ** 3)
        X, Y, Z = idx % np1 / nc, (idx // np1) % np1 / nc, idx // (np1 * np1) / nc
        exact = np.sin(np.pi * X) * np.sin(np.pi * Y) * np.sin(np.pi * Z) / (3 * np.pi ** 2)
        errs.append(np.abs(u - exact).max())
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert (rates > 1.8).all(), (errs, rates)


def test_coarse_3d_reference_bcs():
    """Reference boundary data in 3D (diffusion_problem_ms.tpp:128-137, :206-235;
    dirichlet_bc.tpp:17-43, neumann_bc.tpp:17-41 ignore z): Dirichlet g =
    (x-1/2)^2 + (y-1/2)^2 on x = 0, y = 0, z = 0; Neumann h = cos(2 PI_D x)
    cos(2 PI_D y) on x = 1, y = 1, z = 1, integrated with QGauss(2) on the face
    against the standard trilinear face values (hand sum below)."""
    nc, nf = 2, 4
    r = O.batch_build(3, unit_boxes3(nc), np.full(nc ** 3, nf), O.COEF_SCALAR_CELL, np.ones(nc ** 3 * nf ** 3),
                      O.RHS_CONST, 0.0)
    Kc, Fc, cm, cv = O.coarse_assemble(3, nc, r["Ae"], r["Fe"], 1)
    np1 = nc + 1
    idx = np.arange(np1 ** 3)
    X, Y, Z = idx % np1 / nc, (idx // np1) % np1 / nc, idx // (np1 * np1) / nc
    assert (cm == ((X == 0) | (Y == 0) | (Z == 0))).all()
    assert np.allclose(cv[cm == 1], ((X - 0.5) ** 2 + (Y - 0.5) ** 2)[cm == 1], atol=1e-15)
    pid = 3.14592653509793218403
    gs = np.array([0.5 - 0.5 / np.sqrt(3), 0.5 + 0.5 / np.sqrt(3)])
    h = lambda x, y: np.cos(2 * pid * x) * np.cos(2 * pid * y)
    # node (1, 1, 1/2) (x = y = 1, z = 1/2): on face x = 1 of cells (1,1,0) and
    # (1,1,1) and on face y = 1 of the same cells; the node is the face vertex
    # at (y or x) = 1 and z = 1/2
    node = 2 + np1 * 2 + np1 * np1 * 1
    H = 0.5
    want = 0.0
    for cz in (0, 1):
        for qa in gs:          # along the free in-face axis (y on x = 1, x on y = 1)
            for qz in gs:
                zq = (cz + qz) * H
                pz = qz if cz == 0 else 1 - qz      # trilinear factor along z (node at z = 1/2)
                w = (0.5 * H) ** 2
                a = (1 + qa) * H                      # coordinate along the free axis in cell 1
                want += h(1.0, a) * qa * pz * w       # face x = 1: factor along y is qa (node at y = 1)
                want += h(a, 1.0) * qa * pz * w       # face y = 1: factor along x is qa (node at x = 1)
    assert abs(Fc[node] - want) < 1e-15, (Fc[node], want)
    # a node on z = 1 only: (1/2, 1/2, 1) sees face z = 1 of the four cells around it
    node = 1 + np1 * 1 + np1 * np1 * 2
    want = 0.0
    for cx in (0, 1):
        for cy in (0, 1):
            for qx in gs:
                for qy in gs:
                    px = qx if cx == 0 else 1 - qx
                    py = qy if cy == 0 else 1 - qy
                    want += h((cx + qx) * H, (cy + qy) * H) * px * py * (0.5 * H) ** 2
    assert abs(Fc[node] - want) < 1e-15, (Fc[node], want)
    u = O.coarse_solve(3, nc, Kc, Fc, cm, cv)
    assert np.allclose(u[cm == 1], cv[cm == 1])
