"""General (non-box) quadrilateral coarse cells on the GPU against the oracle.

The reference builds each cell's fine mesh with GridGenerator::general_cell from
the cell's four vertices and refines it (include/base/diffusion_problem_basis.tpp:
92-99); its Dirichlet data is BasisQ1 by monomial inversion
(include/coefficients/basis_q1.tpp:26-96).  Cells here come from the unit
square's mesh with moved interior vertices (msfem.perturbed_quads).

Bars: plain FP64 within 1e-10 relative of the oracle's direct solves on basis
nodal values, A^K and F^K, for every fine resolution and coefficient kind;
FP32 + FP64 refinement at rel_tol 1e-8 within 1e-6 relative energy norm; the
coarse system and solution on the moved mesh within 1e-10; the patch test
(linear reproduction) on the GPU bases to 1e-10.
"""
import sys

import numpy as np
import pytest

import msfem
import msfem_cases as C
import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-10


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = msfem.Context(0)
    yield c
    c.close()


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def check(out, ref, tol=TOL):
    assert (out["status"] == 0).all()
    eb, ea, ef = rel(out["bases"], ref["bases"]), rel(out["Ae"], ref["Ae"]), rel(out["Fe"], ref["Fe"])
    say(f"  bases {eb:.2e}  A^K {ea:.2e}  F^K {ef:.2e}")
    assert eb < tol and ea < tol and ef < tol


@pytest.mark.parametrize("nf", [16, 32, 64, 128])
def test_quad_lognormal_fp64(ctx, nf):
    """Scalar lognormal per fine cell (contrast 1e3), f = 1: one launch mode per
    resolution (per-fine-cell conductances; node couplings at 128)."""
    say(f"quad lognormal nf={nf} fp64")
    case = C.quad_cells(nc=8, nf=nf, which=[9, 18, 27, 45, 54] if nf < 128 else [9, 27])
    check(C.gpu_build(ctx, case), C.oracle_build(case))


def test_quad_reference_coefficient(ctx):
    """MatrixCoeff at the mapped quadrature points, f = 2 (the reference's own problem on moved cells)."""
    say("quad reference coefficient nf=32")
    case = C.quad_cells(nc=8, nf=32, kind="reference", which=[10, 21, 42])
    check(C.gpu_build(ctx, case), C.oracle_build(case))


@pytest.mark.parametrize("nf", [64, 128])
def test_quad_tensor_sinsin_both_precisions(ctx, nf):
    """SPD tensor per fine cell, f = sin(pi x) sin(pi y) at the mapped points."""
    say(f"quad tensor nf={nf}")
    case = C.quad_cells(nc=8, nf=nf, kind="tensor", which=[18, 35])
    ref = C.oracle_build(case)
    check(C.gpu_build(ctx, case), ref)
    out = C.gpu_build(ctx, case, precision=msfem.PREC_FP32_REFINE, rel_tol=1e-8)
    err = C.energy_error(case, out["bases"], ref["bases"])
    say(f"  fp32+refinement energy error {err:.2e}")
    assert (out["status"] == 0).all() and err < 1e-6


def test_quad_intermediate_levels(ctx):
    """64^2 general cells (streamed form): the intermediate hat levels of spacing 2
    and 4 (coarse_space=1, the default), the two-level form alone (2) and Jacobi (0)
    reach the oracle's values (FP64 1e-10; FP32 + refinement at 1e-8 within 1e-6
    energy norm), and the levels take at most 0.75x the two-level form's FP64
    iterations."""
    say("quad intermediate levels nf=64")
    case = C.quad_cells(nc=8, nf=64, which=[9, 27, 45])
    ref = C.oracle_build(case)
    its = {}
    for cs in (0, 2, 1):
        out = C.gpu_build(ctx, case, coarse_space=cs)
        check(out, ref)
        its[cs] = out["iters"].mean()
        for prec in (msfem.PREC_FP64, msfem.PREC_FP32_REFINE):
            o = msfem.default_opts(coarse_space=cs, precision=prec)
            assert ctx.uses_coarse_space(2, 64, o, case["coef_kind"], True) == {0: 0, 2: 1, 1: 2}[cs]
        out = C.gpu_build(ctx, case, coarse_space=cs, precision=msfem.PREC_FP32_REFINE, rel_tol=1e-8)
        assert (out["status"] == 0).all() and C.energy_error(case, out["bases"], ref["bases"]) < 1e-6
    say(f"  mean FP64 iterations: Jacobi {its[0]:.1f}, two-level {its[2]:.1f}, with the levels {its[1]:.1f}")
    assert its[2] < 0.6 * its[0] and its[1] < 0.75 * its[2]


def test_quad_uneven_batch_graph_and_pou(ctx):
    """Mixed resolutions in one batch; graph replay and partition of unity agree with the oracle."""
    say("quad uneven batch, graph and partition of unity")
    case = C.quad_cells(nc=8, which=[9, 10, 17, 18], n_fine=[16, 32, 64, 128])
    ref = C.oracle_build(case)
    check(C.gpu_build(ctx, case), ref)
    for _ in range(2):  # capture, then replay
        check(C.gpu_build(ctx, case, graph=1), ref)
    check(C.gpu_build(ctx, case, partition_of_unity=1), ref)


def test_quad_patch_test(ctx):
    """Constant coefficient: sum_i x_i phi_i = x and sum_i y_i phi_i = y at every fine node."""
    say("quad patch test")
    q = msfem.perturbed_quads(6, 0.24, 5)[[7, 14, 21]]
    nf = np.full(3, 64, np.int32)
    out = ctx.batch_build(2, nf, None, msfem.COEF_SCALAR_CELL, np.full(3 * 64 * 64, 2.5), vertices=q)
    phi = out["bases"].reshape(3, 4, -1)
    for c in range(3):
        X = O.quad_nodes(q[c], 64)
        assert np.abs(q[c][:, 0] @ phi[c] - X[:, 0]).max() < TOL
        assert np.abs(q[c][:, 1] @ phi[c] - X[:, 1]).max() < TOL


@pytest.mark.parametrize("bc", [msfem.BC_ZERO_DIRICHLET, msfem.BC_REFERENCE])
def test_quad_coarse_system(ctx, bc):
    """The whole MsFEM on the moved 8 x 8 mesh (32^2 fine): the mesh keeps the
    unit square's boundary, so the coarse boundary data is the uniform mesh's."""
    say(f"quad coarse system bc={bc}")
    case = C.quad_cells(nc=8, nf=32)
    out = C.gpu_build(ctx, case)
    ref = C.oracle_build(case, want_bases=False)
    Kc, Fc, cm, cv = ctx.coarse_assemble(2, 8, out["Ae"], out["Fe"], bc)
    Kr, Fr, cmr, cvr = O.coarse_assemble(2, 8, ref["Ae"], ref["Fe"], bc)
    assert rel(Kc, Kr) < TOL and rel(Fc, Fr) < TOL and (cm == cmr).all()
    u, _ = ctx.coarse_solve(2, 8, Kc, Fc, cm, cv, rel_tol=1e-14)
    assert rel(u, O.coarse_solve(2, 8, Kr, Fr, cmr, cvr)) < TOL


def test_quad_rejects_bad_input(ctx):
    q = msfem.perturbed_quads(2, 0.1, 1)[:1].copy()
    nf = np.array([16], np.int32)
    coef = np.ones(256)
    bad = q.copy()
    bad[0, [1, 2]] = bad[0, [2, 1]]  # swapped vertices: negative orientation
    with pytest.raises(msfem.MsfemError) as e:
        ctx.batch_build(2, nf, None, msfem.COEF_SCALAR_CELL, coef, vertices=bad)
    assert e.value.code == msfem.ERR_ARG
    with pytest.raises(msfem.MsfemError) as e:
        ctx.batch_build(3, nf, None, msfem.COEF_SCALAR_CELL, np.ones(16 ** 3), vertices=q)
    assert e.value.code == msfem.ERR_ARG
