"""GPU parity: the HIP engine against the CPU oracle and the golden fixtures.

Bars (stated per test): plain FP64 path within 1e-10 relative of the exact
(direct-solve) values on basis nodal values, A^K and F^K; FP32+FP64-refinement
path to the same bar when refined to the same tolerance, and to <= 1e-6
relative energy norm at its loose setting.
"""
import glob
import os
import sys

import numpy as np
import pytest

import msfem
import msfem_cases as C
import oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
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


def check_against_oracle(case, out, ref, tol=TOL):
    assert (out["status"] == 0).all()
    assert rel(out["bases"], ref["bases"]) < tol
    assert rel(out["Ae"], ref["Ae"]) < tol
    assert rel(out["Fe"], ref["Fe"]) < tol


# every fixture: resolutions outside the tuned set run the generic solve
GPU_GOLDEN = sorted(glob.glob(os.path.join(GOLDEN, "*.npz")))


@pytest.mark.parametrize("path", GPU_GOLDEN, ids=[os.path.basename(p) for p in GPU_GOLDEN])
@pytest.mark.parametrize("prec", [msfem.PREC_FP64, msfem.PREC_FP32_REFINE])
def test_golden(ctx, path, prec):
    g = np.load(path)
    say(f"golden {os.path.basename(path)} prec={prec}")
    dim = int(g["dim"])
    out = ctx.batch_build(dim, [int(g["nf"])], g["box"][None], int(g["coef_kind"]), g["coef"], int(g["rhs_kind"]),
                          float(g["rhs_value"]), opts=msfem.default_opts(precision=prec))
    assert rel(out["bases"], g["bases"].reshape(-1)) < TOL
    assert rel(out["Ae"][0], g["Ae"]) < TOL
    assert rel(out["Fe"][0], g["Fe"]) < TOL


def test_periodic_8x8_nf32_and_coarse(ctx):
    """8x8 coarse, 32^2 fine, alpha = 2 + sin(2 pi x/eps) cos(2 pi y/eps), eps = 1/64,
    f = 1, zero Dirichlet: bases, A^K, F^K, coarse system and coarse solution."""
    say("periodic 8x8 nf32")
    case = C.periodic_cells()
    out = C.gpu_build(ctx, case)
    ref = C.oracle_build(case)
    check_against_oracle(case, out, ref)
    Kc, Fc, cm, cv = ctx.coarse_assemble(2, 8, out["Ae"], out["Fe"], msfem.BC_ZERO_DIRICHLET)
    Kr, Fr, cmr, cvr = O.coarse_assemble(2, 8, ref["Ae"], ref["Fe"], 0)
    assert rel(Kc, Kr) < TOL and rel(Fc, Fr) < TOL and (cm == cmr).all()
    u, it = ctx.coarse_solve(2, 8, Kc, Fc, cm, cv, rel_tol=1e-14)
    ur = O.coarse_solve(2, 8, Kr, Fr, cmr, cvr)
    assert rel(u, ur) < TOL, it


@pytest.mark.parametrize("prec", [msfem.PREC_FP64, msfem.PREC_FP32_REFINE])
def test_lognormal_nf64(ctx, prec):
    """Cells of the 64x64-coarse lognormal batch (64^2 fine, contrast 1e3, f = 1)."""
    say(f"lognormal nf64 prec={prec}")
    case = C.lognormal_cells(which=np.array([0, 1, 63, 64, 2080, 4095, 1234, 777, 3000, 17, 4031, 99]))
    out = C.gpu_build(ctx, case, precision=prec)
    ref = C.oracle_build(case)
    check_against_oracle(case, out, ref)
    assert (out["iters"] > 50).all()


def test_lognormal_mixed_loose_energy_error(ctx):
    """FP32 CG + FP64 refinement stopped at rel. residual 1e-8: energy-norm error <= 1e-6."""
    say("lognormal nf64 mixed loose")
    case = C.lognormal_cells(which=np.arange(0, 4096, 512))
    out = C.gpu_build(ctx, case, precision=msfem.PREC_FP32_REFINE, rel_tol=1e-8)
    ref = C.oracle_build(case)
    assert (out["status"] == 0).all()
    err = C.energy_error(case, out["bases"], ref["bases"])
    say(f"energy error {err:.3e}")
    assert err <= 1e-6
    assert rel(out["Ae"], ref["Ae"]) < 1e-6


@pytest.mark.parametrize("prec", [msfem.PREC_FP64, msfem.PREC_FP32_REFINE])
def test_graph_build_matches_stream_build(ctx, prec):
    """The HIP-graph form (one concurrent branch per size class, instantiated once,
    then replayed) gives bit-identical outputs to the stream form, on every
    replay, with eager builds of other batches in between."""
    say(f"graph build prec={prec}")
    case = C.tensor_cells(nc=128, which=np.arange(0, 16384, 1024), n_fine=[16, 32, 64, 128] * 4)
    other = C.lognormal_cells(dim=2, nc=64, nf=32, seed=2, which=np.arange(5))
    ref = C.gpu_build(ctx, case, precision=prec)
    for rep in range(4):
        if rep == 3:
            C.gpu_build(ctx, other, precision=prec)  # another layout: the graph is captured anew
        out = C.gpu_build(ctx, case, precision=prec, graph=1)
        for k in ("Ae", "Fe", "bases", "iters", "status"):
            assert np.array_equal(out[k], ref[k]), (rep, k)
        C.gpu_build(ctx, case, precision=prec)  # eager work (same layout) between replays


@pytest.mark.parametrize("prec", [msfem.PREC_FP64, msfem.PREC_FP32_REFINE])
def test_tensor_uneven_classes(ctx, prec):
    """Uneven batch: SPD tensor per fine cell, every resolution class 16..128, f = sin sin."""
    say(f"tensor uneven prec={prec}")
    case = C.tensor_cells(nc=128, which=np.arange(0, 16384, 1024), n_fine=[16, 32, 64, 128] * 4)
    out = C.gpu_build(ctx, case, precision=prec)
    ref = C.oracle_build(case)
    check_against_oracle(case, out, ref)


@pytest.mark.parametrize("prec", [msfem.PREC_FP64, msfem.PREC_FP32_REFINE])
def test_lognormal_3d(ctx, prec):
    """3D: 16^3 coarse, 16^3 fine per cell, lognormal, 8 bases per cell."""
    say(f"lognormal 3d nf16 prec={prec}")
    case = C.lognormal_cells(dim=3, nc=16, nf=16, seed=5, which=np.array([0, 15, 273, 4095]))
    out = C.gpu_build(ctx, case, precision=prec)
    ref = C.oracle_build(case)
    check_against_oracle(case, out, ref)


def test_reference_default_cells(ctx):
    """Reference default problem: MatrixCoeff with PI_D at the quadrature points,
    f = 2, 128^2 fine per cell (source/main.cxx:23)."""
    say("reference default nf128")
    case = C.reference_cells(which=np.array([0, 27]))
    out = C.gpu_build(ctx, case)
    ref = C.oracle_build(case)
    check_against_oracle(case, out, ref)


def test_reference_default_cells_fp32_levels(ctx):
    """The reference default cells (MatrixCoeff at the quadrature points, 128^2)
    with FP32 + refinement: the intermediate hat levels of spacing 4 and 8 (the
    default) and the two-level form alone reach the oracle's values within 1e-6
    relative energy norm at rel. residual 1e-8, the levels in at most 0.75x the
    iterations."""
    say("reference default nf128, fp32 with and without the levels")
    case = C.reference_cells(which=np.array([0, 27, 45]))
    ref = C.oracle_build(case)
    its = {}
    for cs in (2, 1):
        out = C.gpu_build(ctx, case, coarse_space=cs, precision=msfem.PREC_FP32_REFINE, rel_tol=1e-8)
        err = C.energy_error(case, out["bases"], ref["bases"])
        say(f"  coarse_space={cs}: energy error {err:.2e}, mean iterations {out['iters'].mean():.1f}")
        assert (out["status"] == 0).all() and err < 1e-6
        its[cs] = out["iters"].mean()
    assert its[1] < 0.75 * its[2]


def test_partition_of_unity_option(ctx):
    say("partition of unity")
    case = C.lognormal_cells(which=np.array([5, 6, 7]))
    full = C.gpu_build(ctx, case)
    pou = C.gpu_build(ctx, case, partition_of_unity=1)
    assert rel(pou["bases"], full["bases"]) < TOL
    assert rel(pou["Ae"], full["Ae"]) < TOL
    assert (pou["iters"][:, 3] == 0).all() and (pou["iters"][:, :3] > 0).all()


def test_initial_guess_shortens_solve(ctx):
    say("initial guess")
    case = C.lognormal_cells(which=np.array([10, 11]))
    ref = C.oracle_build(case)
    rng = np.random.default_rng(0)
    x0 = ref["bases"] + 1e-6 * rng.standard_normal(ref["bases"].shape)
    cold = C.gpu_build(ctx, case)
    o = msfem.default_opts()
    warm = ctx.batch_build(2, case["n_fine"], case["box"], case["coef_kind"], case["coef"], x0=x0, opts=o)
    check_against_oracle(case, warm, ref)
    assert warm["iters"].sum() < cold["iters"].sum()


@pytest.mark.parametrize("prec", [msfem.PREC_FP64, msfem.PREC_FP32_REFINE])
def test_q1_start(ctx, prec):
    """The default start (interior = the Q1 interpolant of g) and the zero start
    meet the same test ||r|| <= rel_tol ||K_IB g||, so both match the oracle; the
    Q1 start takes fewer iterations."""
    say(f"q1 start prec={prec}")
    case = C.lognormal_cells(which=np.array([20, 21, 22, 23]))
    ref = C.oracle_build(case)
    q1 = C.gpu_build(ctx, case, precision=prec)
    zero = C.gpu_build(ctx, case, precision=prec, q1_start=0)
    check_against_oracle(case, q1, ref)
    check_against_oracle(case, zero, ref)
    assert q1["iters"].sum() < zero["iters"].sum(), (q1["iters"], zero["iters"])


def test_q1_start_constant_coefficient_is_exact(ctx):
    """Constant scalar coefficient on box cells: the bilinear nodal function is the
    discrete solution, so the Q1 start meets the tolerance in 0 iterations and the
    bases are the closed-form hats (32^2 cells: Jacobi; 64^2: two-level)."""
    say("q1 start constant coefficient")
    box = C.unit_cells(2, 4)
    nf = np.array([32] * 8 + [64] * 8, np.int32)
    coef = np.full(int((nf.astype(np.int64) ** 2).sum()), 3.0)
    out = ctx.batch_build(2, nf, box, msfem.COEF_SCALAR_CELL, coef)
    zero = ctx.batch_build(2, nf, box, msfem.COEF_SCALAR_CELL, coef, opts=msfem.default_opts(q1_start=0))
    assert (out["iters"] == 0).all() and (zero["iters"] > 0).all()
    assert rel(out["bases"], zero["bases"]) < TOL and rel(out["Ae"], zero["Ae"]) < TOL
    off = 0
    for n in nf:
        t = np.arange(n + 1) / n
        hats = [np.outer(ty, tx).ravel() for ty in (1 - t, t) for tx in (1 - t, t)]
        for i in range(4):
            assert np.abs(out["bases"][off:off + (n + 1) ** 2] - hats[i]).max() < 1e-14
            off += (n + 1) ** 2


def test_iteration_cap_reports_noconvergence(ctx):
    say("iteration cap")
    case = C.lognormal_cells(which=np.array([3]))
    with pytest.raises(msfem.MsfemError) as e:
        C.gpu_build(ctx, case, max_iter=5)
    assert e.value.code == msfem.ERR_NOCONV
    o = msfem.default_opts(max_iter=5)
    out = ctx.batch_build(2, case["n_fine"], case["box"], case["coef_kind"], case["coef"], opts=o, raise_noconv=False)
    assert (out["status"] == 1).all() and (out["iters"] == 5).all()


def test_invalid_arguments_are_rejected(ctx):
    say("invalid args")
    case = C.lognormal_cells(which=np.array([0]))
    with pytest.raises(msfem.MsfemError) as e:
        ctx.batch_build(2, [2048], case["box"], case["coef_kind"], np.ones(2048 * 2048))
    assert e.value.code == msfem.ERR_ARG
    with pytest.raises(msfem.MsfemError):
        ctx.batch_build(4, [16], C.unit_cells(3, 2)[:1], msfem.COEF_TENSOR_CELL, np.ones(16 ** 3 * 6))


def test_empty_and_ragged_batches(ctx):
    """An empty batch is a no-op (host and device entry points, graph form too);
    a ragged batch -- one cell of each resolution in an order that interleaves
    the size classes -- gives every cell exactly what a batch of that cell
    alone gives (class grouping and batch composition do not change a value)."""
    say("empty and ragged batches")
    for graph in (0, 1):
        out = ctx.batch_build(2, np.zeros(0, np.int32), np.zeros((0, 4)), msfem.COEF_SCALAR_CELL, np.zeros(0),
                              opts=msfem.default_opts(graph=graph))
        assert out["Ae"].shape == (0, 4, 4) and out["bases"].size == 0
    ctx.batch_build_device(2, np.zeros(0, np.int32), None, msfem.COEF_SCALAR_CELL, None, None, None,
                           stream=ctx.stream)
    ctx.synchronize()
    case = C.tensor_cells(nc=128, which=np.array([5, 900, 77, 3000, 12, 16000]), n_fine=[128, 16, 64, 32, 16, 64])
    whole = C.gpu_build(ctx, case)
    boff = np.concatenate([[0], np.cumsum(4 * (case["n_fine"] + 1) ** 2)])
    coff = np.concatenate([[0], np.cumsum(3 * case["n_fine"] ** 2)])
    for i, nf in enumerate(case["n_fine"]):
        one = dict(case, n_fine=case["n_fine"][i:i + 1], box=case["box"][i:i + 1], coef=case["coef"][coff[i]:coff[i + 1]])
        alone = C.gpu_build(ctx, one)
        assert np.array_equal(whole["bases"][boff[i]:boff[i + 1]], alone["bases"])
        assert np.array_equal(whole["Ae"][i], alone["Ae"][0]) and np.array_equal(whole["iters"][i], alone["iters"][0])


def test_full_lognormal_batch_properties(ctx):
    """The whole 64x64-coarse batch (4096 cells, 64^2 fine): size-independent
    properties on every cell, oracle parity on a sample."""
    say("full lognormal 4096 cells")
    case = C.lognormal_cells()
    out = C.gpu_build(ctx, case)
    assert (out["status"] == 0).all()
    Ae, Fe = out["Ae"], out["Fe"]
    assert np.abs(Ae - Ae.transpose(0, 2, 1)).max() < 1e-12
    assert np.abs(Ae.sum(2)).max() < 1e-10
    assert np.abs(Fe.sum(1) - (1 / 64) ** 2).max() < 1e-15
    b = out["bases"].reshape(4096, 4, 65 * 65)
    assert np.abs(b.sum(1) - 1).max() < 1e-10
    assert b.min() > -1e-10 and b.max() < 1 + 1e-10
    sample = np.array([0, 2047, 4095])
    ref = C.oracle_build(C.lognormal_cells(which=sample))
    assert rel(b[sample].reshape(-1), ref["bases"]) < TOL
    assert rel(Ae[sample], ref["Ae"]) < TOL
    # coarse system on the device path vs oracle
    Kc, Fc, cm, cv = ctx.coarse_assemble(2, 64, Ae, Fe, msfem.BC_ZERO_DIRICHLET)
    u, it = ctx.coarse_solve(2, 64, Kc, Fc, cm, cv, rel_tol=1e-14)
    ur = O.coarse_solve(2, 64, *O.coarse_assemble(2, 64, Ae, Fe, 0))
    assert rel(u, ur) < TOL, it


def test_reconstruction_device(ctx):
    say("reconstruction")
    case = C.lognormal_cells(which=np.array([0, 1, 2]))
    out = C.gpu_build(ctx, case)
    w = np.random.default_rng(1).standard_normal((3, 4))
    tb = ctx.to_device(out["bases"])
    tw = ctx.to_device(w)
    tu = ctx.alloc(3 * 65 * 65, np.float64)
    ctx.reconstruct_device(2, case["n_fine"], tb.ptr, tw.ptr, tu.ptr, stream=ctx.stream)
    ctx.synchronize()
    want = np.concatenate([O.reconstruct(2, 64, out["bases"][c * 4 * 4225:(c + 1) * 4 * 4225], w[c]) for c in range(3)])
    assert rel(tu.numpy(), want) < 1e-14


@pytest.mark.parametrize("prec", [msfem.PREC_FP64, msfem.PREC_FP32_REFINE])
def test_every_launch_shape(ctx, prec):
    """Every compiled launch shape (autotuner candidates) of every 2D class gives the exact values."""
    case = C.tensor_cells(nc=128, which=np.arange(0, 16384, 2048), n_fine=[16, 32, 64, 128] * 2)
    ref = C.oracle_build(case)
    for nf in (16, 32, 64, 128):
        for v in range(1, msfem.variant_count(2, nf) + 1):
            say(f"launch shape nf={nf} variant={v} prec={prec}")
            sel = np.where(case["n_fine"] == nf)[0]
            sub = C.tensor_cells(nc=128, which=np.arange(0, 16384, 2048)[sel], n_fine=[nf] * len(sel))
            out = C.gpu_build(ctx, sub, precision=prec, variant=v)
            r = C.oracle_build(sub)
            check_against_oracle(sub, out, r)
    for v in range(1, msfem.variant_count(3, 16) + 1):
        say(f"launch shape 3d nf=16 variant={v} prec={prec}")
        c3 = C.lognormal_cells(dim=3, nc=16, nf=16, seed=5, which=np.array([7, 100]))
        check_against_oracle(c3, C.gpu_build(ctx, c3, precision=prec, variant=v), C.oracle_build(c3))


@pytest.mark.parametrize("kind", ["lognormal", "tensor"])
def test_two_level_preconditioner(ctx, kind):
    """64^2 cells with the two-level local solve plus the intermediate hat levels
    (coarse_space=1, the default), the two-level form alone (2) and Jacobi alone
    (0): the same values (oracle, 1e-10; FP32 + refinement at 1e-8 within 1e-6
    energy norm), the two-level form at most 0.6x Jacobi's iterations and the
    intermediate levels at most 0.75x the two-level form's, on both launch shapes."""
    say(f"two-level preconditioner, {kind} 64^2")
    if kind == "lognormal":
        case = C.lognormal_cells(which=np.array([3, 700, 2222, 4000]))
    else:
        case = C.tensor_cells(nc=128, which=np.arange(8), n_fine=[64] * 8)
    ref = C.oracle_build(case)
    its = {}
    for cs in (0, 2, 1):
        for variant in (1, 4):  # both launch shapes of the two-level form
            out = C.gpu_build(ctx, case, coarse_space=cs, variant=variant)
            check_against_oracle(case, out, ref)
            its[cs, variant] = out["iters"].mean()
        assert ctx.uses_coarse_space(2, 64, msfem.default_opts(coarse_space=cs, variant=1), case["coef_kind"]) == \
            {0: 0, 2: 1, 1: 2}[cs]
        out = C.gpu_build(ctx, case, coarse_space=cs, precision=msfem.PREC_FP32_REFINE, rel_tol=1e-8)
        assert (out["status"] == 0).all() and C.energy_error(case, out["bases"], ref["bases"]) < 1e-6
    say(f"  mean iterations: Jacobi {its[0, 1]:.1f}, two-level {its[2, 1]:.1f} / {its[2, 4]:.1f}, "
        f"with the intermediate levels {its[1, 1]:.1f} / {its[1, 4]:.1f}")
    for v in (1, 4):
        assert its[2, v] < 0.6 * its[0, v] and its[1, v] < 0.75 * its[2, v]
    # shapes outside the two-level table keep Jacobi
    assert not ctx.uses_coarse_space(2, 64, msfem.default_opts(variant=2), case["coef_kind"])
    assert not ctx.uses_coarse_space(2, 32, msfem.default_opts(), case["coef_kind"])


def test_intermediate_levels_128(ctx):
    """128^2 tensor cells, FP32 + refinement (streamed form): the intermediate hat
    levels of spacing 4 and 8 (coarse_space=1, the default), the two-level form
    alone (2) and Jacobi (0) reach the oracle's values (1e-10), and the levels take
    at most 0.75x the two-level form's iterations.  FP64 keeps the two-level form."""
    say("intermediate levels, tensor 128^2 FP32")
    case = C.tensor_cells(nc=128, which=np.array([5, 3000, 9001]), n_fine=[128] * 3)
    ref = C.oracle_build(case)
    its = {}
    for cs in (0, 2, 1):
        out = C.gpu_build(ctx, case, coarse_space=cs, precision=msfem.PREC_FP32_REFINE)
        check_against_oracle(case, out, ref)
        its[cs] = out["iters"].mean()
        assert ctx.uses_coarse_space(2, 128, msfem.default_opts(coarse_space=cs,
                                                                 precision=msfem.PREC_FP32_REFINE),
                                     case["coef_kind"]) == {0: 0, 2: 1, 1: 2}[cs]
    say(f"  mean iterations: Jacobi {its[0]:.1f}, two-level {its[2]:.1f}, with the levels {its[1]:.1f}")
    assert its[2] < 0.6 * its[0] and its[1] < 0.75 * its[2]
    assert ctx.uses_coarse_space(2, 128, msfem.default_opts(), case["coef_kind"]) == 1


@pytest.mark.parametrize("bc", [msfem.BC_ZERO_DIRICHLET, msfem.BC_REFERENCE])
def test_coarse_3d_lognormal_full_batch(ctx, bc):
    """The 3D coarse system of the 16^3 x 16^3 lognormal batch
    (diffusion_problem_ms.tpp:161-296 at dim 3): every cell built on the GPU (FP32
    CG + FP64 refinement to rel. residual 1e-13), then the 27-point coarse
    assembly with the 3D boundary data and the FP64 coarse CG on the GPU, against
    the oracle's assembly and direct solve of the same A^K, F^K (1e-10); a sample
    of cells against the oracle's own build (1e-10)."""
    say(f"coarse 3d lognormal 4096 cells bc={bc}")
    case = C.lognormal_cells(dim=3, nc=16, nf=16, seed=5)
    out = C.gpu_build(ctx, case, precision=msfem.PREC_FP32_REFINE)
    assert (out["status"] == 0).all()
    Ae, Fe = out["Ae"], out["Fe"]
    assert np.abs(Ae - Ae.transpose(0, 2, 1)).max() < 1e-12 * np.abs(Ae).max()
    assert np.abs(Ae.sum(2)).max() < 1e-10 * np.abs(Ae).max()
    assert np.abs(Fe.sum(1) - (1 / 16) ** 3).max() < 1e-15
    sample = np.array([0, 1365, 4095])
    ref = C.oracle_build(C.lognormal_cells(dim=3, nc=16, nf=16, seed=5, which=sample))
    assert rel(Ae[sample], ref["Ae"]) < TOL and rel(Fe[sample], ref["Fe"]) < TOL
    Kc, Fc, cm, cv = ctx.coarse_assemble(3, 16, Ae, Fe, bc)
    Kr, Fr, cmr, cvr = O.coarse_assemble(3, 16, Ae, Fe, bc)
    assert rel(Kc, Kr) < 1e-13 and rel(Fc, Fr) < 1e-13 and (cm == cmr).all() and rel(cv, cvr) < 1e-15
    u, it = ctx.coarse_solve(3, 16, Kc, Fc, cm, cv, rel_tol=1e-14)
    ur = O.coarse_solve(3, 16, Kr, Fr, cmr, cvr)
    say(f"  coarse CG iterations {it}")
    assert rel(u, ur) < TOL, it


def test_async_builds_on_two_streams_with_layout_change(ctx):
    """Two device builds enqueued back to back with no host synchronisation: the
    first on the context stream, the second -- another batch layout, so the
    context's shared plan metadata and scratch are rewritten -- on the
    context's non-blocking second stream.  Each gives bitwise what the same
    batch gives built alone."""
    say("async builds on two streams, layout change")
    a = C.lognormal_cells(which=np.arange(0, 4096, 64))
    b = C.tensor_cells(nc=128, which=np.array([3, 99, 1000, 7000, 12000, 16000]), n_fine=[64, 128, 16, 32, 64, 128])
    want = [C.gpu_build(ctx, a), C.gpu_build(ctx, b)]
    aux = ctx.aux_stream
    bufs = []
    for case, stream in ((a, ctx.stream), (b, aux)):
        n = len(case["n_fine"])
        d = dict(box=ctx.to_device(np.asarray(case["box"]).reshape(-1)), coef=ctx.to_device(case["coef"]),
                 Ae=ctx.alloc(n * 16, np.float64), Fe=ctx.alloc(n * 4, np.float64),
                 bases=ctx.alloc(msfem.bases_size(2, case["n_fine"]), np.float64),
                 it=ctx.alloc(n * 4, np.int32), st=ctx.alloc(n * 4, np.int32))
        bufs.append(d)
        ctx.batch_build_device(2, case["n_fine"], d["box"].ptr, case["coef_kind"], d["coef"].ptr, d["Ae"].ptr,
                               d["Fe"].ptr, d["bases"].ptr, d["it"].ptr, d["st"].ptr, rhs_kind=case["rhs_kind"],
                               rhs_value=case["rhs_value"], stream=stream)
    ctx.synchronize()
    for d, w in zip(bufs, want):
        assert np.array_equal(d["Ae"].numpy().reshape(w["Ae"].shape), w["Ae"])
        assert np.array_equal(d["bases"].numpy(), w["bases"])
        assert np.array_equal(d["it"].numpy().reshape(w["iters"].shape), w["iters"])
        for arr in d.values():
            arr.free()


@pytest.mark.parametrize("nc,bc", [(128, msfem.BC_ZERO_DIRICHLET), (128, msfem.BC_REFERENCE), (12, msfem.BC_REFERENCE),
                                   (25, msfem.BC_ZERO_DIRICHLET)])
def test_coarse_solve_sizes(ctx, nc, bc):
    """The coarse CG (multigrid-preconditioned where the grid coarsens to a small
    dense level: 128 -> 64 ... 2, 12 -> 6 -> 3; Jacobi-preconditioned for 25)
    against the oracle's direct solve of the same coarse system (1e-10), on
    the A^K, F^K of nc x nc lognormal cells of 16^2 fine cells."""
    say(f"coarse solve nc={nc} bc={bc}")
    case = C.lognormal_cells(dim=2, nc=nc, nf=16, seed=7)
    out = C.gpu_build(ctx, case)
    Kc, Fc, cm, cv = ctx.coarse_assemble(2, nc, out["Ae"], out["Fe"], bc)
    u, it = ctx.coarse_solve(2, nc, Kc, Fc, cm, cv, rel_tol=1e-14)
    ur = O.coarse_solve(2, nc, Kc, Fc, cm, cv)
    say(f"  coarse CG iterations {it}")
    assert rel(u, ur) < TOL, it
    if nc != 25:
        assert it <= 30  # multigrid preconditioning
