"""Fine-scale standard FEM on the GPU (msfem_fine_solve; the reference's
DiffusionProblem<2>, include/base/diffusion_problem.tpp:40-279) against the
oracle, and the MsFEM against it.

Bars: the GPU fine solution equals the oracle's direct solve of the same
assembled system (QGauss(2) fine-cell element matrices and loads, the
reference boundary data or zero Dirichlet) within 1e-10 relative.  The MsFEM
reconstruction on the fine mesh is closer to the fine solution than the
standard Q1 FEM on the coarse mesh is (the point of the method), for the
oscillatory periodic coefficient.
"""
import sys

import numpy as np
import pytest

import msfem
import msfem_cases as C
import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = msfem.Context(0)
    yield c
    c.close()


def oracle_fine(n, kind, coef, rhs_kind, rhs_value, bc_kind):
    """Direct solve of the fine system assembled by the oracle: one oracle
    fine element per fine cell, the oracle's coarse assembly (boundary data)
    on the n x n mesh, its direct coarse solve."""
    h = 1.0 / n
    stride = msfem.coef_stride(2, kind)
    Ae, Fe = np.zeros((n * n, 4, 4)), np.zeros((n * n, 4))
    for f in range(n * n):
        x0 = np.array([(f % n) * h, (f // n) * h])
        Ke, be = O.fine_element(2, x0, np.array([h, h]), kind, coef[f * stride:(f + 1) * stride], rhs_kind, rhs_value)
        Ae[f], Fe[f] = Ke.reshape(4, 4), be
    Kc, Fc, cm, cv = O.coarse_assemble(2, n, Ae, Fe, bc_kind)
    return O.coarse_solve(2, n, Kc, Fc, cm, cv)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_fine_solve_reference_problem(ctx):
    """The reference's own fine problem: MatrixCoeff (PI_D) at the quadrature points,
    f = 2, Dirichlet g on x=0, y=0 and Neumann h on x=1, y=1; 32 x 32 fine cells."""
    say("fine solve, reference problem, n=32")
    n = 32
    unit = np.array([[[0.0, 1.0], [0.0, 1.0]]])
    coef = msfem.coef_reference(2, np.array([n], np.int32), unit)
    u, it = ctx.fine_solve(n, msfem.COEF_TENSOR_QP, coef, msfem.RHS_CONST, 2.0, msfem.BC_REFERENCE, rel_tol=1e-14)
    ref = oracle_fine(n, msfem.COEF_TENSOR_QP, coef, msfem.RHS_CONST, 2.0, msfem.BC_REFERENCE)
    say(f"  {it} CG iterations, rel. max difference {rel(u, ref):.2e}")
    assert rel(u, ref) < 1e-10


def test_fine_solve_lognormal_zero_dirichlet(ctx):
    """Scalar lognormal coefficient per fine cell, f = 1, u = 0 on the boundary; 64 x 64."""
    say("fine solve, lognormal, n=64")
    n = 64
    coef = C.lognormal_cells(dim=2, nc=1, nf=n, seed=11)["coef"]
    u, it = ctx.fine_solve(n, msfem.COEF_SCALAR_CELL, coef, msfem.RHS_CONST, 1.0, msfem.BC_ZERO_DIRICHLET,
                           rel_tol=1e-14)
    ref = oracle_fine(n, msfem.COEF_SCALAR_CELL, coef, msfem.RHS_CONST, 1.0, msfem.BC_ZERO_DIRICHLET)
    say(f"  {it} CG iterations, rel. max difference {rel(u, ref):.2e}")
    assert rel(u, ref) < 1e-10


def test_fine_solve_iteration_cap(ctx):
    """A capped solve reports non-convergence (MSFEM_ERR_NOCONV)."""
    n = 64
    coef = C.lognormal_cells(dim=2, nc=1, nf=n, seed=11)["coef"]
    with pytest.raises(msfem.MsfemError) as e:
        ctx.fine_solve(n, msfem.COEF_SCALAR_CELL, coef, max_iter=5)
    assert e.value.code == msfem.ERR_NOCONV


def test_msfem_beats_coarse_fem(ctx):
    """Periodic alpha = 2 + sin(2 pi x/eps) cos(2 pi y/eps), eps = 1/64, f = 1, zero
    Dirichlet: on the 256^2 fine mesh, the MsFEM (8 x 8 coarse cells x 32^2 fine)
    is closer to the fine solution than standard Q1 on the 8 x 8 mesh."""
    say("MsFEM vs fine vs coarse Q1, periodic eps=1/64")
    nc, nf = 8, 32
    n = nc * nf
    unit = np.array([[[0.0, 1.0], [0.0, 1.0]]])
    fine_coef = msfem.coef_periodic(np.array([n], np.int32), unit, 1 / 64)
    u_fine, it = ctx.fine_solve(n, msfem.COEF_SCALAR_QP, fine_coef, msfem.RHS_CONST, 1.0, msfem.BC_ZERO_DIRICHLET,
                                rel_tol=1e-12)
    u_fine = u_fine.reshape(n + 1, n + 1)
    # MsFEM: batched bases, coarse system, reconstruction on every cell
    case = C.periodic_cells(nc=nc, nf=nf)
    out = C.gpu_build(ctx, case)
    Kc, Fc, cm, cv = ctx.coarse_assemble(2, nc, out["Ae"], out["Fe"], msfem.BC_ZERO_DIRICHLET)
    uc, _ = ctx.coarse_solve(2, nc, Kc, Fc, cm, cv)
    w = msfem.gather_weights(2, nc, uc)
    bases = C.split_bases(case, out["bases"])
    u_ms = np.zeros((n + 1, n + 1))
    for c in range(nc * nc):
        cx, cy = c % nc, c // nc
        u_ms[cy * nf:(cy + 1) * nf + 1, cx * nf:(cx + 1) * nf + 1] = \
            np.tensordot(w[c], np.asarray(bases[c]), axes=1).reshape(nf + 1, nf + 1)
    # standard Q1 on the coarse mesh, interpolated bilinearly to the fine nodes
    q1_coef = msfem.coef_periodic(np.array([nc], np.int32), unit, 1 / 64)
    u_q1, _ = ctx.fine_solve(nc, msfem.COEF_SCALAR_QP, q1_coef, msfem.RHS_CONST, 1.0, msfem.BC_ZERO_DIRICHLET,
                             rel_tol=1e-12)
    u_q1 = u_q1.reshape(nc + 1, nc + 1)
    t = np.linspace(0, nc, n + 1)
    i0 = np.minimum(t.astype(int), nc - 1)
    s = t - i0
    ux = u_q1[:, i0] * (1 - s) + u_q1[:, i0 + 1] * s
    u_q1f = ux[i0, :] * (1 - s)[:, None] + ux[i0 + 1, :] * s[:, None]
    err_ms = np.linalg.norm(u_ms - u_fine) / np.linalg.norm(u_fine)
    err_q1 = np.linalg.norm(u_q1f - u_fine) / np.linalg.norm(u_fine)
    say(f"  fine CG {it} iterations; relative l2 error: MsFEM {err_ms:.3e}, coarse Q1 {err_q1:.3e}")
    assert err_ms < err_q1
    assert err_ms < 0.1
