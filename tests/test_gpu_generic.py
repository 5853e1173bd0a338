"""The generic local solve (csrc/generic.hip) against the oracle's exact solves:
any number of local refinements (the reference refines n_refine_local times,
include/base/diffusion_problem_basis.tpp:92-99) and, in 3D, every coefficient
kind -- including the reference's MatrixCoeff<3> at the QGauss(2) points
(include/coefficients/matrix_coeff.tpp:28-41,66-91).  Bars as for the tuned
kernels: 1e-10 relative on basis nodal values, A^K and F^K; the energy rule's
guarantee on the generic classes too; a mixed batch (tuned and generic
classes) gives every cell what it gives alone."""
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


def check(case, out, ref, tol=TOL):
    assert (out["status"] == 0).all(), out["status"]
    assert rel(out["bases"], ref["bases"]) < tol
    assert rel(out["Ae"], ref["Ae"]) < tol
    assert rel(out["Fe"], ref["Fe"]) < tol


def tensor3(n, seed):
    """SPD 3x3 tensors R diag(l) R^T, log l ~ N(0, 1) -> [n][6] (a11,a12,a13,a22,a23,a33)."""
    rng = np.random.default_rng(seed)
    out = np.zeros((n, 6))
    for k in range(n):
        Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        A = Q @ np.diag(np.exp(rng.standard_normal(3))) @ Q.T
        out[k] = [A[0, 0], A[0, 1], A[0, 2], A[1, 1], A[1, 2], A[2, 2]]
    return out.reshape(-1)


@pytest.mark.parametrize("prec", [msfem.PREC_FP64, msfem.PREC_FP32_REFINE])
@pytest.mark.parametrize("nf", [2, 4, 8, 256])
def test_any_refinement_2d(ctx, nf, prec):
    """2D boxes with nf = 2^n_refine_local outside the tuned set: lognormal
    scalar cells (f = 1) and SPD tensor cells (f = sin sin)."""
    say(f"generic 2D nf={nf} prec={prec}")
    which = np.array([0, 5]) if nf == 256 else np.array([0, 5, 17, 40])
    for case in (C.lognormal_cells(nc=8, nf=nf, seed=3, which=which),
                 C.tensor_cells(nc=128, which=which, n_fine=[nf] * len(which))):
        out = C.gpu_build(ctx, case, precision=prec)
        check(case, out, C.oracle_build(case, n_threads=8))


def test_any_refinement_3d_and_every_kind(ctx):
    """3D: nf = 4 and 32 scalar (lognormal), nf = 8 with an SPD tensor per fine
    cell, a scalar at the quadrature points, and the reference MatrixCoeff<3>
    at the quadrature points (f = 2)."""
    cases = {
        "scalar nf4": C.lognormal_cells(dim=3, nc=4, nf=4, seed=5, which=np.array([0, 9, 63])),
        "scalar nf32": C.lognormal_cells(dim=3, nc=2, nf=32, seed=5, which=np.array([3])),
    }
    box = msfem.unit_boxes(3, 4)[np.array([1, 22])]
    nfa = np.array([8, 8], np.int32)
    cases["tensor nf8"] = dict(dim=3, n_fine=nfa, box=box, coef_kind=msfem.COEF_TENSOR_CELL,
                               coef=tensor3(2 * 512, 1), rhs_kind=msfem.RHS_SINSIN, rhs_value=0.0)
    cases["scalar-qp nf8"] = dict(dim=3, n_fine=nfa, box=box, coef_kind=msfem.COEF_SCALAR_QP,
                                  coef=np.exp(np.random.default_rng(2).standard_normal(2 * 512 * 8)),
                                  rhs_kind=msfem.RHS_CONST, rhs_value=1.0)
    cases["reference MatrixCoeff<3> nf8"] = dict(dim=3, n_fine=nfa, box=box, coef_kind=msfem.COEF_TENSOR_QP,
                                                 coef=msfem.coef_reference(3, nfa, box), rhs_kind=msfem.RHS_CONST,
                                                 rhs_value=2.0)
    for name, case in cases.items():
        say(f"generic 3D {name}")
        ref = C.oracle_build(case)
        for prec in (msfem.PREC_FP64, msfem.PREC_FP32_REFINE):
            out = C.gpu_build(ctx, case, precision=prec)
            check(case, out, ref)


def test_general_quadrilaterals_any_refinement(ctx):
    say("generic 2D general quadrilaterals nf=8, reference coefficient")
    case = C.quad_cells(nc=8, nf=8, kind="reference", which=np.array([0, 9, 35]))
    for q1 in (0, 1):
        out = C.gpu_build(ctx, case, q1_start=q1)
        check(case, out, C.oracle_build(case))


@pytest.mark.parametrize("name", ["2d nf8", "3d tensor-qp nf4"])
def test_energy_rule_on_generic_classes(ctx, name):
    import test_gpu_energy as E
    tau = 1e-6
    if name == "2d nf8":
        case = C.lognormal_cells(nc=8, nf=8, seed=3, which=np.array([1, 2, 3]))
    else:
        box = msfem.unit_boxes(3, 2)[:2]
        nfa = np.array([4, 4], np.int32)
        case = dict(dim=3, n_fine=nfa, box=box, coef_kind=msfem.COEF_TENSOR_QP, coef=msfem.coef_reference(3, nfa, box),
                    rhs_kind=msfem.RHS_CONST, rhs_value=2.0)
    say(f"energy rule on a generic class: {name}")
    out = ctx.batch_build(case["dim"], case["n_fine"], case["box"], case["coef_kind"], case["coef"],
                          case["rhs_kind"], case["rhs_value"], energy_bound=True,
                          opts=msfem.default_opts(energy_tol=tau, rel_tol=0.0))
    lam, elb = E.oracle_terms(case)
    nb = 1 << case["dim"]
    diag = out["Ae"][:, np.arange(nb), np.arange(nb)]
    assert np.abs(out["energy_bound"] - O.energy_bound(lam[:, None], diag, out["resid"])).max() <= \
        1e-8 * max(out["energy_bound"].max(), 1e-300)
    assert out["energy_bound"].max() <= tau
    true = E.true_errors(case, out["bases"], C.oracle_build(case)["bases"])
    assert true.max() <= tau


def test_mixed_tuned_and_generic_classes(ctx):
    """One batch with cells of nf = 8 (generic), 16, 64 (tuned) and 256 (generic):
    every cell bitwise what a batch of that cell alone gives, and the oracle's values."""
    say("mixed tuned and generic classes")
    case = C.tensor_cells(nc=128, which=np.array([3, 77, 901, 4000]), n_fine=[8, 16, 64, 256])
    whole = C.gpu_build(ctx, case)
    check(case, whole, C.oracle_build(case))
    boff = np.concatenate([[0], np.cumsum(4 * (case["n_fine"] + 1) ** 2)])
    coff = np.concatenate([[0], np.cumsum(3 * case["n_fine"] ** 2)])
    for i in range(4):
        one = dict(case, n_fine=case["n_fine"][i:i + 1], box=case["box"][i:i + 1], coef=case["coef"][coff[i]:coff[i + 1]])
        alone = C.gpu_build(ctx, one)
        assert np.array_equal(whole["bases"][boff[i]:boff[i + 1]], alone["bases"])
        assert np.array_equal(whole["Ae"][i], alone["Ae"][0])


def test_resolution_limits(ctx):
    case = C.lognormal_cells(nc=8, nf=8, which=np.array([0]))
    for dim, nf in ((2, 2048), (3, 128), (2, 0)):
        with pytest.raises(msfem.MsfemError) as e:
            ctx.batch_build(dim, [nf], msfem.unit_boxes(dim, 1), msfem.COEF_SCALAR_CELL,
                            np.ones(max(nf, 1) ** dim))
        assert e.value.code == msfem.ERR_ARG
    assert case is not None
