"""Writes the golden fixtures of the local MsFEM basis problem (tests/golden/*.npz).

An independent restatement, dense numpy assembly + scipy sparse direct solve,
of DiffusionProblemBasis<dim>::run() (include/base/diffusion_problem_basis.tpp:
make_grid :92-99, Dirichlet data N_i :119-135, QGauss(2) assembly :161-242,
basis problems :450-465, element matrix/load :247-285).  The reference itself
needs deal.II and cannot be built here, so these vectors pin the oracle and the
GPU path from a third implementation; each case stores its inputs and outputs.

Run:  python tests/golden/make_golden.py
"""
import itertools
import os

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl

HERE = os.path.dirname(os.path.abspath(__file__))
GP = np.array([0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0)])


def qp_tensors(dim, kind, tile, f):
    """Tensor at each of the 2^d QGauss(2) points of fine cell f -> [nq][d][d]."""
    nq = 1 << dim
    ns = 3 if dim == 2 else 6
    idx = [(0, 0), (0, 1), (1, 1)] if dim == 2 else [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    A = np.zeros((nq, dim, dim))
    for q in range(nq):
        if kind == 0:
            A[q] = np.eye(dim) * tile[f]
        elif kind == 1:
            s = tile[f * ns:(f + 1) * ns]
            for k, (i, j) in enumerate(idx):
                A[q, i, j] = A[q, j, i] = s[k]
        elif kind == 2:
            A[q] = np.eye(dim) * tile[f * nq + q]
        else:
            s = tile[(f * nq + q) * ns:(f * nq + q + 1) * ns]
            for k, (i, j) in enumerate(idx):
                A[q, i, j] = A[q, j, i] = s[k]
    return A


def assemble(dim, nf, box, kind, tile, rhs):
    """Dense K and b on the (nf+1)^dim lexicographic nodes."""
    np1 = nf + 1
    nn = np1 ** dim
    h = np.array([(box[k][1] - box[k][0]) / nf for k in range(dim)])
    K = np.zeros((nn, nn))
    b = np.zeros(nn)
    verts = list(itertools.product([0, 1], repeat=dim))
    verts = [tuple(v[::-1]) for v in verts]  # x fastest
    for f, ci in enumerate(itertools.product(range(nf), repeat=dim)):
        ci = ci[::-1]  # x fastest
        A = qp_tensors(dim, kind, tile, f)
        Ke = np.zeros((1 << dim, 1 << dim))
        be = np.zeros(1 << dim)
        for q, qi in enumerate(itertools.product([0, 1], repeat=dim)):
            qi = qi[::-1]
            xi = GP[list(qi)]
            x = np.array([box[k][0] + h[k] * (ci[k] + xi[k]) for k in range(dim)])
            w = np.prod(h) / (1 << dim)
            val = np.array([np.prod([xi[k] if v[k] else 1 - xi[k] for k in range(dim)]) for v in verts])
            grad = np.zeros((1 << dim, dim))
            for a, v in enumerate(verts):
                for k in range(dim):
                    g = 1.0 if v[k] else -1.0
                    for m in range(dim):
                        if m != k:
                            g *= xi[m] if v[m] else 1 - xi[m]
                    grad[a, k] = g / h[k]
            Ke += w * grad @ A[q] @ grad.T
            fv = rhs(x)
            be += w * fv * val
        nodes = [sum((ci[k] + v[k]) * np1 ** k for k in range(dim)) for v in verts]
        K[np.ix_(nodes, nodes)] += Ke
        b[nodes] += be
    return K, b


def solve_bases(dim, nf, box, K):
    np1 = nf + 1
    nn = np1 ** dim
    I = np.array(np.unravel_index(np.arange(nn), [np1] * dim, order="F")).T  # x fastest
    bnd = np.any((I == 0) | (I == nf), axis=1)
    Ii, Bi = np.where(~bnd)[0], np.where(bnd)[0]
    t = I / nf
    phi = np.zeros((1 << dim, nn))
    KII = sp.csc_matrix(K[np.ix_(Ii, Ii)])
    for i in range(1 << dim):
        g = np.prod([t[:, k] if (i >> k) & 1 else 1 - t[:, k] for k in range(dim)], axis=0)
        x = spl.spsolve(KII, -K[np.ix_(Ii, Bi)] @ g[Bi])
        phi[i, Bi] = g[Bi]
        phi[i, Ii] = x
    return phi


def case(name, dim, nf, box, kind, tile, rhs_kind, rhs_value):
    if rhs_kind == 0:
        rhs = lambda x: rhs_value
    else:
        rhs = lambda x: float(np.prod(np.sin(np.pi * x)))
    K, b = assemble(dim, nf, box, kind, tile, rhs)
    phi = solve_bases(dim, nf, box, K)
    Ae = phi @ K @ phi.T
    Fe = phi @ b
    np.savez(os.path.join(HERE, name + ".npz"), dim=dim, nf=nf, box=np.array(box, dtype=np.float64),
             coef_kind=kind, coef=tile, rhs_kind=rhs_kind, rhs_value=rhs_value, bases=phi, Ae=Ae, Fe=Fe)
    print(name, "ok", Ae.shape)


def main():
    rng = np.random.default_rng(20260)
    # 2D lognormal scalar cell, f = 1
    case("local2d_scalar_nf8", 2, 8, [[0.25, 0.375], [0.5, 0.625]], 0,
         np.exp(np.clip(rng.standard_normal(64), -3.45, 3.45)), 0, 1.0)
    # 2D SPD tensor per fine cell on a non-square box, f = sin(pi x) sin(pi y)
    th = rng.uniform(0, np.pi, 16)
    l1, l2 = np.exp(1.5 * rng.standard_normal(16)), np.exp(1.5 * rng.standard_normal(16))
    c, s = np.cos(th), np.sin(th)
    tens = np.stack([l1 * c * c + l2 * s * s, (l1 - l2) * c * s, l1 * s * s + l2 * c * c], axis=1).reshape(-1)
    case("local2d_tensor_nf4_rect", 2, 4, [[0.0, 0.5], [0.25, 0.5]], 1, tens, 1, 0.0)
    # 2D tensor at the quadrature points (smooth rotated coefficient), f = 2
    vals = []
    for f in range(16):
        for q in range(4):
            a = 1.0 + 0.5 * np.sin(f + q)
            vals += [a, 0.3 * np.cos(f * q), 2.0 - 0.5 * np.cos(f)]
    case("local2d_tensorqp_nf4", 2, 4, [[0.5, 0.75], [0.0, 0.25]], 3, np.array(vals), 0, 2.0)
    # 3D lognormal scalar cell, f = 1
    case("local3d_scalar_nf4", 3, 4, [[0.0, 0.25], [0.25, 0.5], [0.5, 0.75]], 0,
         np.exp(np.clip(rng.standard_normal(64), -3.45, 3.45)), 0, 1.0)
    # sizes the GPU kernels take (2D: 16..128 fine cells per side, 3D: 8, 16)
    case("local2d_scalar_nf16", 2, 16, [[0.5, 0.625], [0.125, 0.25]], 0,
         np.exp(np.clip(rng.standard_normal(256), -3.45, 3.45)), 0, 1.0)
    th = rng.uniform(0, np.pi, 256)
    l1, l2 = np.exp(1.5 * rng.standard_normal(256)), np.exp(1.5 * rng.standard_normal(256))
    c, s = np.cos(th), np.sin(th)
    tens = np.stack([l1 * c * c + l2 * s * s, (l1 - l2) * c * s, l1 * s * s + l2 * c * c], axis=1).reshape(-1)
    case("local2d_tensor_nf16", 2, 16, [[0.25, 0.5], [0.5, 0.75]], 1, tens, 1, 0.0)
    case("local3d_scalar_nf8", 3, 8, [[0.5, 0.625], [0.0, 0.125], [0.25, 0.375]], 0,
         np.exp(np.clip(rng.standard_normal(512), -3.45, 3.45)), 0, 1.0)


if __name__ == "__main__":
    main()
