"""General (non-box) hexahedral coarse cells on the GPU against the oracle.

The reference builds each cell's fine mesh with GridGenerator::general_cell from
the cell's eight vertices and refines it (include/base/diffusion_problem_basis.tpp:
92-99), integrates with the trilinear mapping and QGauss(2) (:161-242), and takes
BasisQ1 by monomial inversion as Dirichlet data (include/coefficients/basis_q1.tpp:
26-96).  Cells here come from the unit cube's mesh with moved interior vertices
(msfem.perturbed_hexes).  They run the generic local solve (FP64 for both
precision settings).

Bars: 1e-10 relative of the oracle's direct solves on basis nodal values, A^K
and F^K for every coefficient kind and several resolutions; the coarse system
and solution on the moved mesh within 1e-10; the patch test (linear
reproduction) on the GPU bases to 1e-10; orientation and energy-rule inputs
rejected.
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
    assert (out["status"] == 0).all(), out["status"]
    eb, ea, ef = rel(out["bases"], ref["bases"]), rel(out["Ae"], ref["Ae"]), rel(out["Fe"], ref["Fe"])
    say(f"  bases {eb:.2e}  A^K {ea:.2e}  F^K {ef:.2e}")
    assert eb < tol and ea < tol and ef < tol


@pytest.mark.parametrize("prec", [msfem.PREC_FP64, msfem.PREC_FP32_REFINE])
@pytest.mark.parametrize("kind", ["lognormal", "reference", "tensor"])
def test_hex_coefficient_kinds(ctx, kind, prec):
    """Scalar lognormal per fine cell (f = 1), MatrixCoeff<3> at the mapped
    quadrature points (f = 2) and SPD tensors per fine cell (f = sin sin sin)."""
    say(f"hex {kind} nf=8 prec={prec}")
    case = C.hex_cells(nc=3, nf=8, kind=kind, which=[13, 4, 22])
    check(C.gpu_build(ctx, case, precision=prec), C.oracle_build(case))


@pytest.mark.parametrize("nf", [1, 2, 4, 16])
def test_hex_resolutions(ctx, nf):
    """Other n_refine_local: one fine cell up to 16^3 per coarse cell."""
    say(f"hex lognormal nf={nf}")
    case = C.hex_cells(nc=3, nf=nf, which=[13, 0] if nf == 16 else [13, 0, 26])
    check(C.gpu_build(ctx, case), C.oracle_build(case))


def test_hex_uneven_batch_graph_and_pou(ctx):
    """Mixed resolutions in one batch, the HIP-graph replay and the partition-of-unity option."""
    say("hex uneven batch, graph, pou")
    case = C.hex_cells(nc=3, kind="reference", which=[13, 4, 10, 22], n_fine=[4, 8, 2, 8])
    ref = C.oracle_build(case)
    check(C.gpu_build(ctx, case), ref)
    for _ in range(2):  # capture, then replay
        check(C.gpu_build(ctx, case, graph=1), ref)
    check(C.gpu_build(ctx, case, partition_of_unity=1), ref)


def test_hex_twisted_cell(ctx):
    """A single strongly non-affine cell (twisted top face) beside a perturbed one."""
    say("hex twisted cell")
    twisted = np.array([[0.00, 0.00, 0.00], [1.00, 0.05, 0.00], [0.05, 0.90, 0.00], [1.10, 1.00, 0.10],
                        [0.15, -0.10, 1.00], [1.05, 0.15, 0.95], [-0.05, 0.95, 1.10], [0.90, 1.05, 1.00]])
    h = np.stack([twisted, msfem.perturbed_hexes(3, 0.16, 4)[13]])
    nfa = np.array([8, 8], np.int32)
    case = dict(dim=3, n_fine=nfa, box=None, vertices=h, coef_kind=msfem.COEF_TENSOR_QP,
                coef=msfem.coef_reference_hex(nfa, h), rhs_kind=msfem.RHS_SINSIN, rhs_value=0.0)
    check(C.gpu_build(ctx, case), C.oracle_build(case))


def test_hex_patch_test(ctx):
    """Constant coefficient: sum_i x_k(v_i) phi_i = x_k at every fine node."""
    say("hex patch test")
    h = msfem.perturbed_hexes(3, 0.16, 5)[[13, 4]]
    nf = np.full(2, 16, np.int32)
    out = ctx.batch_build(3, nf, None, msfem.COEF_SCALAR_CELL, np.full(2 * 16 ** 3, 2.5), vertices=h)
    assert (out["status"] == 0).all()
    phi = out["bases"].reshape(2, 8, -1)
    for c in range(2):
        X = O.hex_nodes(h[c], 16)
        for k in range(3):
            assert np.abs(h[c][:, k] @ phi[c] - X[:, k]).max() < TOL


@pytest.mark.parametrize("bc", [msfem.BC_ZERO_DIRICHLET, msfem.BC_REFERENCE])
def test_hex_coarse_system(ctx, bc):
    """The whole MsFEM on the moved 3^3 mesh (4^3 fine): the mesh keeps the unit
    cube's boundary, so the coarse boundary data is the uniform mesh's."""
    say(f"hex coarse system bc={bc}")
    case = C.hex_cells(nc=3, nf=4, kind="reference")
    out = C.gpu_build(ctx, case)
    ref = C.oracle_build(case, want_bases=False)
    Kc, Fc, cm, cv = ctx.coarse_assemble(3, 3, out["Ae"], out["Fe"], bc)
    Kr, Fr, cmr, cvr = O.coarse_assemble(3, 3, ref["Ae"], ref["Fe"], bc)
    assert rel(Kc, Kr) < TOL and rel(Fc, Fr) < TOL and (cm == cmr).all()
    u, _ = ctx.coarse_solve(3, 3, Kc, Fc, cm, cv, rel_tol=1e-14)
    assert rel(u, O.coarse_solve(3, 3, Kr, Fr, cmr, cvr)) < TOL


def test_hex_rejects_bad_input(ctx):
    h = msfem.perturbed_hexes(2, 0.1, 1)[:1].copy()
    nf = np.array([4], np.int32)
    coef = np.ones(64)
    bad = h.copy()
    bad[0, [1, 2]] = bad[0, [2, 1]]  # swapped vertices: negative orientation
    with pytest.raises(msfem.MsfemError) as e:
        ctx.batch_build(3, nf, None, msfem.COEF_SCALAR_CELL, coef, vertices=bad)
    assert e.value.code == msfem.ERR_ARG
    flat = h.copy()
    flat[0, 4:, 2] = flat[0, :4, 2]  # top face on the bottom face: degenerate
    with pytest.raises(msfem.MsfemError) as e:
        ctx.batch_build(3, nf, None, msfem.COEF_SCALAR_CELL, coef, vertices=flat)
    assert e.value.code == msfem.ERR_ARG
    with pytest.raises(msfem.MsfemError) as e:  # the energy rule's eigenvalue bound is for boxes
        ctx.batch_build(3, nf, None, msfem.COEF_SCALAR_CELL, coef, vertices=h,
                        opts=msfem.default_opts(energy_tol=1e-6))
    assert e.value.code == msfem.ERR_ARG
