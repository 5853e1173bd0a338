"""Pins the oracle's a posteriori energy-norm bound of a local solve
(oracle.energy_bound_terms / energy_bound) by the mathematics it rests on:
the closed-form smallest eigenvalue of the unit Q1 stiffness against a dense
eigensolver, the complementary-energy lower bound against exact solves (and
its equality for a constant coefficient), and the bound against the true error
of perturbed solutions.  No counterpart in the reference, whose local solves
stop at an absolute residual (include/base/diffusion_problem_basis.tpp:297)."""
import numpy as np
import pytest

import oracle as O


def interior_mask(dim, nf):
    np1 = nf + 1
    idx = np.arange(np1 ** dim)
    I = [(idx // np1 ** k) % np1 for k in range(dim)]
    return np.all([(v > 0) & (v < nf) for v in I], axis=0)


def dense_KII(dim, nf, box, kind, tile):
    K, b = O.local_assemble(dim, nf, box, kind, tile)
    A = O.stencil_to_dense(dim, nf, K)
    m = interior_mask(dim, nf)
    return K, b, A, m


@pytest.mark.parametrize("dim,nf,H", [(2, 4, (1, 1)), (2, 8, (0.25, 1)), (2, 8, (3, 0.5)), (2, 16, (1, 1)),
                                      (3, 4, (1, 1, 1)), (3, 4, (0.5, 1, 2)), (3, 6, (1, 1, 1))])
def test_unit_stiffness_lambda_min_closed_form(dim, nf, H):
    box = np.array([[0.0, h] for h in H])
    _, _, A, m = dense_KII(dim, nf, box, O.COEF_SCALAR_CELL, np.ones(nf ** dim))
    lam = np.linalg.eigvalsh(A[np.ix_(m, m)])[0]
    assert abs(O.unit_stiffness_lambda_min(dim, nf, H) - lam) < 1e-12 * max(1.0, lam)


def cases():
    rng = np.random.default_rng(11)
    out = []
    nf = 8
    out.append((2, nf, np.array([[0.1, 0.35], [0.2, 0.45]]), O.COEF_SCALAR_CELL,
                np.exp(np.clip(rng.standard_normal(nf * nf) * 1.5, -3.45, 3.45))))
    th = rng.uniform(0, np.pi, nf * nf)
    l1, l2 = np.exp(rng.normal(0, 1.5, nf * nf)), np.exp(rng.normal(0, 1.5, nf * nf))
    c, s = np.cos(th), np.sin(th)
    out.append((2, nf, np.array([[0.0, 0.5], [0.0, 0.125]]), O.COEF_TENSOR_CELL,
                np.stack([c * c * l1 + s * s * l2, c * s * (l1 - l2), s * s * l1 + c * c * l2], 1).reshape(-1)))
    out.append((2, nf, np.array([[0.0, 0.125], [0.0, 0.125]]), O.COEF_SCALAR_QP,
                np.exp(rng.standard_normal(nf * nf * 4))))
    refc = np.concatenate([O.matrix_coeff_ref(2, [0.3 + x / 64, 0.6 + y / 64])
                           for y in range(nf) for x in range(nf) for _ in range(4)])
    out.append((2, nf, np.array([[0.3, 0.3 + nf / 64], [0.6, 0.6 + nf / 64]]), O.COEF_TENSOR_QP, refc))
    out.append((3, 4, np.array([[0.0, 0.25], [0.5, 0.75], [0.25, 0.5]]), O.COEF_SCALAR_CELL,
                np.exp(np.clip(rng.standard_normal(64), -3.45, 3.45))))
    return out


@pytest.mark.parametrize("case", range(5))
def test_lower_bounds_hold_on_exact_solves(case):
    """lam_lb <= lambda_min(K_II) and e_lb[i] <= ||phi_i*||_K^2 = A^K_ii of the
    exact (direct) solution, for scalar, tensor, scalar-QP and reference-QP
    tiles and a 3D tile; the energy bound is within a factor 10 here."""
    dim, nf, box, kind, tile = cases()[case]
    K, b, A, m = dense_KII(dim, nf, box, kind, tile)
    lam_lb, e_lb = O.energy_bound_terms(dim, nf, box, kind, tile)
    lam = np.linalg.eigvalsh(A[np.ix_(m, m)])[0]
    assert 0 < lam_lb <= lam * (1 + 1e-12)
    phi = O.local_bases_direct(dim, nf, box, K)
    Ae, _ = O.element(dim, nf, K, b, phi)
    for i in range(1 << dim):
        assert 0 < e_lb[i] <= Ae[i, i] * (1 + 1e-12), (i, e_lb[i], Ae[i, i])
        assert e_lb[i] > 0.1 * Ae[i, i]


@pytest.mark.parametrize("dim", [2, 3])
def test_constant_coefficient_bound_is_exact(dim):
    """Constant alpha: phi_i* = N_i and the complementary bound is attained,
    e_lb = alpha e_unit = A^K_ii; and lam_lb = alpha mu_min exactly."""
    nf = 4
    box = np.array([[0.0, 0.5], [0.25, 0.5], [0.0, 1.0]][:dim])
    tile = np.full(nf ** dim, 2.5)
    K, b, A, m = dense_KII(dim, nf, box, O.COEF_SCALAR_CELL, tile)
    lam_lb, e_lb = O.energy_bound_terms(dim, nf, box, O.COEF_SCALAR_CELL, tile)
    phi = O.local_bases_direct(dim, nf, box, K)
    Ae, _ = O.element(dim, nf, K, b, phi)
    assert np.allclose(e_lb, np.diag(Ae), rtol=1e-13)
    assert abs(lam_lb - np.linalg.eigvalsh(A[np.ix_(m, m)])[0]) < 1e-13 * lam_lb


@pytest.mark.parametrize("case", range(5))
def test_bound_dominates_true_error(case):
    """phi = phi* + delta (interior, random, three sizes): the bound from the
    residual norm and ||phi||_K^2 is >= the true relative energy error, and
    within the factor sqrt(lambda_min / lam_lb) of it for smooth errors."""
    dim, nf, box, kind, tile = cases()[case]
    K, b, A, m = dense_KII(dim, nf, box, kind, tile)
    lam_lb, _ = O.energy_bound_terms(dim, nf, box, kind, tile)
    phi = O.local_bases_direct(dim, nf, box, K)
    rng = np.random.default_rng(case)
    nn = (nf + 1) ** dim
    for i in range(1 << dim):
        ref = phi[i]
        for scale in (1e-3, 1e-6, 1e-9):
            d = np.zeros(nn)
            d[m] = scale * rng.standard_normal(m.sum())
            x = ref + d
            r = -(A @ x)[m]
            eta = np.linalg.norm(r)
            e_ref = x @ A @ x
            true = np.sqrt(d @ A @ d / (ref @ A @ ref))
            bound = O.energy_bound(lam_lb, e_ref, eta)
            assert bound >= true * (1 - 1e-9), (i, scale, bound, true)
