"""Benchmark of the MsFEM basis build on one MI355X.

Default workload (surrogate128, the one the headline metric is quoted on): the
128 x 128 coarse Q1 cells of the unit square, each refined to a 64 x 64 fine Q1
grid (4225 nodes, 4 basis problems per cell, 65 536 local solves),
alpha = exp(g) with g a Gaussian random field (correlation length 4 fine cells,
clipped to contrast 1e3, seed 41), f = 1, zero Dirichlet.  One step = every
cell's basis solves + A^K, F^K + the coarse assembly, inputs resident in HBM,
timed three ways:
  fp64            the plain FP64 path (CG to relative residual 1e-13, the
                  bases within 1e-10 of the exact ones)
  fp32            FP32 CG with FP64 refinement under the energy stopping rule
                  (every basis guaranteed within 5e-7 relative energy norm of
                  the exact discrete solution: a 2x margin on the 1e-6 bar)
  fp32_surrogate  the same, each basis solve started from the trained
                  surrogate's prediction (its BF16 inference inside the step,
                  each 4096-cell chunk's inference overlapping the builds of the
                  previous chunk); the build's energy test is the gate: a
                  prediction that meets it takes no CG iteration, any other
                  falls back to the CG from the prediction
`value` is the fp32_surrogate leg (the metric's "batched FP32+surrogate"), with
the plain FP64 leg, the ratio, the fallback rate and, over all 16 384 cells,
the largest a posteriori bound and the largest measured energy-norm error
against the FP64 leg beside it.

Other workloads (--workload): lognormal64 (64^2 cells x 64^2 fine, lognormal,
FP64), tensor128 (128^2 cells, fine 16..128 by contrast, SPD tensor, FP32),
lognormal3d (16^3 x 16^3), periodic8 (8^2 x 32^2), reference (8^2 x 128^2, the
reference's default MatrixCoeff problem), quad64 (64^2 general quadrilaterals).

Algorithmic work per CG iteration of a problem with N_I interior nodes:
F_iter = (2*3^d - 1 + 10) N_I flops (stencil + CG vector updates), plus
8 N_I + 2*49^2 for the two-level correction where it runs, plus f_ml(nf)
(75 496 at 64^2) for the intermediate hat levels where they run.  The CPU
baseline is the oracle's restatement of the reference's own recipe (CG +
SSOR(1.6), absolute tolerance 1e-12, 1000 iterations;
include/base/diffusion_problem_basis.tpp:295-317) on a sample of the same
cells, one MPI rank per host core (mpiexec; one OpenMP thread per core
without MPICH).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "mpi-parallel-multiscale-diffusion-fem_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

PEAK_FP64_TFLOPS = 78.6       # MI355X dense FP64 (vector FMA issue rate equals it)
PEAK_FP32_TFLOPS = 157.3      # MI355X published FP32 peak
PEAK_BF16_MFMA_TFLOPS = 2500.0  # dense
# HBM bytes per launch of the dominant kernel, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE passes scaled by a plain copy of known size (--copy-calibration):
# profiles/local_solve_2d_lognormal64_hbm_traffic.md (FP64, 4096 cells) and
# profiles/local_solve_2d_fp32_grf128_hbm_traffic.md (FP32 energy rule, 16384 cells)
TRAFFIC_BYTES = {("lognormal64", 0): 2.946802e9, ("surrogate128", 1): 9.696504e9}
METRIC = "coarse-cell bases built/sec: plain FP64 vs batched FP32+surrogate @1e-6 err"
TAU = 5e-7  # energy stopping rule of the FP32 legs: a 2x margin on the 1e-6 relative energy-norm bar
WEIGHTS = os.path.join(PKG, "surrogate_weights.npz")


def f_ml(nf):
    """Algorithmic flops per iteration of the intermediate hat levels (spacing
    nf/32 and nf/16: 2 and 4 at 64^2, 4 and 8 at 128^2) of a 2D problem of nf^2
    fine cells: separable restriction (row weights over the fine columns, then
    column weights), diagonal scaling, separable bilinear prolongation and the
    addition to z."""
    if nf < 64:  # no intermediate levels below 64^2
        return 0
    n = nf - 1
    f = 2 * n * n  # z += the two levels' values
    for s in (nf // 32, nf // 16):
        m = nf // s - 1  # interior coarse nodes per side
        f += m * n * 2 * (2 * s - 1) + m * m * 2 * (2 * s - 1)  # restriction
        f += m * m                                               # D_s^-1
        f += m * n * 3 + n * n * 3                               # prolongation
    return f


def f_iter_2d(nf, coarse):
    """Algorithmic flops of one CG iteration of one 2D local problem; coarse as
    msfem_uses_coarse_space (0 Jacobi, 1 two-level, 2 with the intermediate levels)."""
    f = (2 * 9 - 1 + 10) * (nf - 1) ** 2
    if coarse:
        f += 8 * (nf - 1) ** 2 + 2 * 49 * 49
    return f + (f_ml(nf) if coarse == 2 else 0)


def workload(name):
    """-> (case, nc, main precision, bc kind, description)."""
    import msfem
    import workloads as C
    if name == "lognormal64":
        return (C.lognormal_cells(dim=2, nc=64, nf=64, seed=2), 64, msfem.PREC_FP64, 0,
                "64^2 coarse Q1 cells x 64^2 fine Q1 per cell, lognormal alpha (contrast 1e3), f=1, 16384 local solves")
    if name == "tensor128":
        return (C.tensor_cells(nc=128, seed=3, sigma=1.5), 128, msfem.PREC_FP32_REFINE, 0,
                "128^2 coarse cells, fine 16/32/64/128 per cell by probe contrast, SPD tensor per fine cell, "
                "f=sin(pi x)sin(pi y), 65536 local solves")
    if name == "lognormal3d":
        return (C.lognormal_cells(dim=3, nc=16, nf=16, seed=5), 16, msfem.PREC_FP32_REFINE, 0,
                "16^3 coarse Q1 cells x 16^3 fine Q1 per cell, lognormal alpha, f=1, 32768 local solves")
    if name == "periodic8":
        return (C.periodic_cells(nc=8, nf=32), 8, msfem.PREC_FP64, 0,
                "8^2 coarse cells x 32^2 fine, alpha=2+sin(2 pi x/eps)cos(2 pi y/eps) eps=1/64, f=1, 256 local solves")
    if name == "quad64":
        return (C.quad_cells(nc=64, nf=64, amp=0.24, seed=9), 64, msfem.PREC_FP64, 0,
                "64^2 general quadrilateral coarse cells (unit-square mesh, interior vertices moved up to 0.24 H) "
                "x 64^2 fine Q1 per cell with the bilinear mapping, lognormal alpha, f=1, 16384 local solves")
    if name == "reference":
        return (C.reference_cells(nc=8, nf=128), 8, msfem.PREC_FP64, 1,
                "reference default: 8^2 coarse cells x 128^2 fine, MatrixCoeff (PI_D) at QGauss(2) points, f=2, "
                "Dirichlet/Neumann coarse data")
    raise SystemExit(f"unknown workload {name}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="surrogate128",
                    choices=["surrogate128", "lognormal64", "tensor128", "lognormal3d", "periodic8", "reference",
                             "quad64"])
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="cells in the CPU-baseline sample (default 2048; 1024 for tensor128, whose 128^2 cells "
                         "keep 2048 beyond half a minute)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (3 times): the reference kernel that scales the "
                         "FETCH_SIZE / WRITE_SIZE counters to bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the local solves of each coarse cell are independent and "
              "the multi-GPU split (cells sharded, one coarse-system reduction) is not built.", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.workload == "surrogate128":
        return bench_headline(args)
    return bench_workload(args)


def timed(ctx, fn, steps, warmup):
    """Wall-clock and device ms per step of fn over `steps` steps after `warmup`."""
    for _ in range(warmup):
        fn()
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.mark(0)
    for _ in range(steps):
        fn()
    ctx.mark(1)
    ctx.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps, ctx.elapsed(0, 1) / steps


def bench_headline(args):
    """surrogate128: plain FP64 vs batched FP32 (+ surrogate) on the 128^2-coarse GRF batch."""
    import msfem
    import surrogate as S
    ctx = msfem.Context(0)
    tuned = ctx.load_variant_table()
    net = S.Surrogate.load(ctx, WEIGHTS)
    data = S.TileSet(ctx, 128 * 128, 41)
    n, nf = data.n, data.nf
    per = 4 * (nf + 1) ** 2
    x0, ref = ctx.alloc(n * per, np.float64), ctx.alloc(n * per, np.float64)
    bound, err = ctx.alloc(n * 4, np.float64), ctx.alloc(n * 4, np.float64)
    ncn = 129 ** 2
    Kc, Fc = ctx.alloc(ncn * 9, np.float64), ctx.alloc(ncn, np.float64)
    cm, cv = ctx.alloc(ncn, np.uint8), ctx.alloc(ncn, np.float64)
    u, cit = ctx.alloc(ncn, np.float64), ctx.alloc(1, np.int32)
    o64 = msfem.default_opts(precision=msfem.PREC_FP64, graph=1)
    o32 = msfem.default_opts(precision=msfem.PREC_FP32_REFINE, energy_tol=TAU, rel_tol=0.0, graph=1)

    def coarse():
        ctx.coarse_assemble_device(2, 128, data.Ae.ptr, data.Fe.ptr, msfem.BC_ZERO_DIRICHLET, Kc.ptr, Fc.ptr, cm.ptr,
                                   cv.ptr, stream=ctx.stream)

    legs = {
        "fp64": lambda: (data.solve(o64), coarse()),
        "fp32": lambda: (data.solve(o32, bound_ptr=bound.ptr), coarse()),
        "fp32_surrogate": lambda: (S.serve_pipelined(net, data, x0, o32, bound_ptr=bound.ptr), coarse()),
    }
    res = {}
    for name, fn in legs.items():
        ms, dev_ms = timed(ctx, fn, args.steps, args.warmup)
        if data.status.numpy().any():
            raise RuntimeError(f"{name}: local solves did not converge")
        its = data.iters.numpy().astype(np.int64)
        r = {"cells_per_s": round(n / (ms * 1e-3), 1), "ms_per_step": round(ms, 3), "device_ms": round(dev_ms, 3),
             "iters_mean": round(float(its.mean()), 2), "iters_p99": int(np.percentile(its, 99))}
        if name == "fp64":
            ref.upload(data.bases.numpy())  # the reference bases of the energy check (outside the timed region)
        else:
            ctx.energy_error_device(2, data.n_fine, data.box.ptr, msfem.COEF_SCALAR_CELL, data.coef.ptr,
                                    data.bases.ptr, ref.ptr, err.ptr, stream=ctx.stream)
            ctx.synchronize()
            e, b = err.numpy(), bound.numpy()
            r.update({"energy_err_max": float(e.max()), "energy_err_median": float(np.median(e)),
                      "energy_bound_max": float(b.max()), "energy_checked_bases": int(e.size)})
            if not (e.max() <= 1e-6 and b.max() <= TAU):
                raise RuntimeError(f"{name}: energy-norm error {e.max():.3e} / bound {b.max():.3e} above the bar")
            if name == "fp32_surrogate":
                r["fallback_rate"] = round(float((its > 0).mean()), 4)
        res[name] = r

    # device time of the dominant pieces: the FP32 local solve (stream form with
    # HIP events on its stream) and the inference alone (marks on its stream)
    o32s = msfem.default_opts(precision=msfem.PREC_FP32_REFINE, energy_tol=TAU, rel_tol=0.0)
    ctx.set_profiling(True)
    solve_ms, solve_its = [], []
    for k in range(args.warmup + args.steps):
        data.solve(o32s)
        t = ctx.profile()[0]
        if k >= args.warmup:
            solve_ms.append(t)
            solve_its.append(int(data.iters.numpy().astype(np.int64).sum()))
    ctx.set_profiling(False)
    solve_ms = float(np.mean(solve_ms))
    for _ in range(args.warmup):
        net.predict_bases(data.x, n, x0.ptr)
    ctx.mark(2)
    for _ in range(args.steps):
        net.predict_bases(data.x, n, x0.ptr)
    ctx.mark(3)
    infer_ms = ctx.elapsed(2, 3) / args.steps
    # coarse solve of the step's coarse system (multigrid-preconditioned FP64 CG)
    data.solve(o64)
    coarse()
    ctx.coarse_solve_device(2, 128, Kc.ptr, Fc.ptr, cm.ptr, cv.ptr, 1e-12, 1000, u.ptr, cit.ptr, stream=ctx.stream)
    ctx.synchronize()  # (warm: scratch allocated)
    ctx.mark(4)
    for _ in range(3):
        ctx.coarse_solve_device(2, 128, Kc.ptr, Fc.ptr, cm.ptr, cv.ptr, 1e-12, 1000, u.ptr, cit.ptr,
                                stream=ctx.stream)
    ctx.mark(5)
    ctx.synchronize()
    coarse_ms = ctx.elapsed(4, 5) / 3  # device time (HIP events on the context stream)

    coarse = ctx.uses_coarse_space(2, nf, o32s, msfem.COEF_SCALAR_CELL)
    flops_solve = float(np.mean(solve_its)) * f_iter_2d(nf, coarse)
    solve_tf = flops_solve / (solve_ms * 1e-3) / 1e12
    flops_inf = float(net.flops_per_sample()) * n
    inf_tf = flops_inf / (infer_ms * 1e-3) / 1e12
    # what the prediction buys: CG iterations saved, their solve time, the break-even inference
    ms_per_it = solve_ms / float(np.mean(solve_its))
    saved_its = res["fp32"]["iters_mean"] - res["fp32_surrogate"]["iters_mean"]
    saved_ms = saved_its * 4 * n * ms_per_it
    econ = {"network_flops_per_tile": net.flops_per_sample(), "network_params": net.n_params(),
            "inference_ms": round(infer_ms, 3), "inference_tflops": round(inf_tf, 2),
            "cg_iterations_saved_per_solve": round(saved_its, 2), "solve_ms_saved": round(saved_ms, 3),
            "break_even_flops_per_tile_at_peak": round(saved_ms * 1e-3 * PEAK_BF16_MFMA_TFLOPS * 1e12 / n),
            "pays_for_itself": bool(infer_ms < saved_ms)}
    rl_solve = {"bound": "mfma", "engine": "valu", "kernel": "local_solve_2d<float, 64, ...> (FP32 CG + FP64 "
                "refinement, energy rule)", "achieved": round(solve_tf, 3), "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s", "frac": round(solve_tf / PEAK_FP32_TFLOPS, 4),
                "traffic": TRAFFIC_BYTES[("surrogate128", msfem.PREC_FP32_REFINE)],
                "kernel_ms": round(solve_ms, 4), "flops_per_launch": flops_solve,
                "hbm_GBps_counted": round(TRAFFIC_BYTES[("surrogate128", msfem.PREC_FP32_REFINE)] / (solve_ms * 1e-3) /
                                          1e9, 1)}
    rl_inf = {"bound": "mfma", "engine": "mfma bf16", "kernel": "surrogate inference (direct 3x3 BF16-MFMA "
              "convolutions + glue, all 16384 tiles)", "achieved": round(inf_tf, 3), "peak": PEAK_BF16_MFMA_TFLOPS,
              "unit": "TFLOP/s", "frac": round(inf_tf / PEAK_BF16_MFMA_TFLOPS, 4), "traffic": None,
              "kernel_ms": round(infer_ms, 4), "flops_per_launch": flops_inf}
    # the step's dominant kernel is the one function with the most device time:
    # the FP32 local solve (one launch per chunk); the inference is ~25 kernels
    dominant, other = rl_solve, rl_inf
    main_leg = res["fp32_surrogate"]
    result = {
        "metric": METRIC,
        "value": main_leg["cells_per_s"],
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_leg["ms_per_step"],
        "local_solves_per_s": round(main_leg["cells_per_s"] * 4, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 CG with f64 refinement + bf16 network (f32 accumulate)",
        "data": "synthetic (Philox-4x32-10 Gaussian random fields generated on the host, resident in HBM); network "
                "trained on the GPU from plain-solver bases of 5000 other cells (seed 40)",
        "config": {"workload": "surrogate128: 128^2 coarse cells x 64^2 fine, alpha=exp(g), g Gaussian random field "
                               "(corr. length 4 fine cells, contrast <= 1e3, seed 41), f=1, 65536 local solves; step = "
                               "(inference +) batched basis build + coarse assembly",
                   "coarse_cells": n, "dim": 2, "local_solves": 4 * n, "energy_tol": TAU,
                   "path": "FP32 CG + FP64 refinement under the energy stopping rule, started from the surrogate's "
                           "prediction (4 chunks, inference on a second stream)", "autotuned_classes": tuned},
        "plain_fp64_cells_per_s": res["fp64"]["cells_per_s"],
        "ratio_vs_plain_fp64": round(main_leg["cells_per_s"] / res["fp64"]["cells_per_s"], 3),
        "fp32_without_surrogate_ratio_vs_plain_fp64": round(res["fp32"]["cells_per_s"] / res["fp64"]["cells_per_s"],
                                                             3),
        "fallback_rate": main_leg["fallback_rate"],
        "energy_err_max_vs_fp64_all_cells": main_leg["energy_err_max"],
        "energy_bound_max_all_cells": main_leg["energy_bound_max"],
        "roofline": dominant,
        "roofline_other": other,
        "legs": res,
        "surrogate_economics": econ,
        "coarse_solve": {"ms": round(coarse_ms, 3), "iterations": int(cit.numpy()[0]), "nodes": ncn},
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(2, data.box_host, np.full(n, nf, np.int64), msfem.COEF_SCALAR_CELL,
                                              data.coef_host, msfem.RHS_CONST, 1.0, n, args.cpu_sample or 2048)
    for arr in (x0, ref, bound, err, Kc, Fc, cm, cv, u, cit):
        arr.free()
    data.free()
    ctx.close()
    print(json.dumps(result))


def bench_workload(args):
    """The other workloads: the main precision's build as a replayed HIP graph."""
    import msfem
    case, nc, main_prec, bc_kind, desc = workload(args.workload)
    dim = case["dim"]
    nf_arr = np.asarray(case["n_fine"]).astype(np.int64)
    n = len(nf_arr)
    nb = 1 << dim
    n_bases = msfem.bases_size(dim, nf_arr)

    ctx = msfem.Context(0)
    quads = case.get("vertices")
    t_box = None if quads is not None else ctx.to_device(np.asarray(case["box"]).reshape(-1))
    t_vert = None if quads is None else ctx.to_device(np.asarray(quads).reshape(-1))
    t_coef = ctx.to_device(case["coef"])
    t_Ae = ctx.alloc(n * nb * nb, np.float64)
    t_Fe = ctx.alloc(n * nb, np.float64)
    t_bases = ctx.alloc(n_bases, np.float64)
    t_it = ctx.alloc(n * nb, np.int32)
    t_st = ctx.alloc(n * nb, np.int32)
    t_bound = ctx.alloc(n * nb, np.float64)
    ncn = (nc + 1) ** dim
    t_Kc = ctx.alloc(ncn * 3 ** dim, np.float64)
    t_Fc = ctx.alloc(ncn, np.float64)
    t_cm = ctx.alloc(ncn, np.uint8)
    t_cv = ctx.alloc(ncn, np.float64)
    t_u = ctx.alloc(ncn, np.float64)
    t_cit = ctx.alloc(1, np.int32)
    tuned = ctx.load_variant_table()  # launch shapes picked by autotune.py on this GPU
    stream = ctx.stream

    def step(opts, bound=False):
        ctx.batch_build_device(dim, case["n_fine"], t_box.ptr if t_box is not None else None, case["coef_kind"],
                               t_coef.ptr, t_Ae.ptr, t_Fe.ptr, t_bases.ptr, t_it.ptr, t_st.ptr,
                               rhs_kind=case["rhs_kind"], rhs_value=case["rhs_value"], opts=opts, stream=stream,
                               vertices_ptr=t_vert.ptr if t_vert is not None else None,
                               energy_bound_ptr=t_bound.ptr if bound else None)
        ctx.coarse_assemble_device(dim, nc, t_Ae.ptr, t_Fe.ptr, bc_kind, t_Kc.ptr, t_Fc.ptr, t_cm.ptr, t_cv.ptr,
                                   stream=stream)

    # FP32 + FP64 refinement under the energy stopping rule (box cells); general
    # cells have no eigenvalue bound, so their FP32 leg stops at rel. residual 1e-9
    energy = quads is None
    o32 = (msfem.default_opts(precision=msfem.PREC_FP32_REFINE, energy_tol=TAU, rel_tol=0.0) if energy else
           msfem.default_opts(precision=msfem.PREC_FP32_REFINE, rel_tol=1e-9))
    opts = {msfem.PREC_FP64: msfem.default_opts(precision=msfem.PREC_FP64), msfem.PREC_FP32_REFINE: o32}
    pname = {msfem.PREC_FP64: "fp64",
             msfem.PREC_FP32_REFINE: "fp32_refine_energy_rule" if energy else "fp32_refine_rel1e-9"}
    ctx.set_profiling(True)
    res = {}
    for p, o in opts.items():
        solve_ms = []
        for k in range(args.warmup + args.steps):
            step(o, bound=energy)
            t = ctx.profile()[0]
            if k >= args.warmup:
                solve_ms.append(t)
        ctx.synchronize()
        assert int(t_st.numpy().sum()) == 0, "local solves did not converge"
        r = {"solve_kernel_ms": float(np.mean(solve_ms)), "its": t_it.numpy().astype(np.int64).reshape(n, nb)}
        if energy:
            r["energy_bound_max"] = float(t_bound.numpy().max())
        res[p] = r
    solve_ms, its = res[main_prec]["solve_kernel_ms"], res[main_prec]["its"]
    # the headline step: the same build as one replayed HIP graph (a concurrent
    # branch per size class), timed by the wall clock and by device marks
    gopts = msfem.default_opts(graph=1)
    for f, *_ in msfem.SolverOpts._fields_:
        if f != "graph":
            setattr(gopts, f, getattr(opts[main_prec], f))
    ms, graph_dev_ms = timed(ctx, lambda: step(gopts), args.steps, args.warmup)
    assert int(t_st.numpy().sum()) == 0, "local solves did not converge (graph)"
    # the same step with the last basis of each cell as 1 - (sum of the others)
    # (exact by linearity: K 1 = 0 and the boundary data sum to 1)
    popts = msfem.default_opts(graph=1)
    for f, *_ in msfem.SolverOpts._fields_:
        setattr(popts, f, getattr(gopts, f))
    popts.partition_of_unity = 1
    pou_ms, _ = timed(ctx, lambda: step(popts), args.steps, args.warmup)
    assert int(t_st.numpy().sum()) == 0, "local solves did not converge (partition of unity)"
    # coarse solve (timed apart: multigrid-preconditioned FP64 CG)
    step(opts[main_prec])
    ctx.coarse_solve_device(dim, nc, t_Kc.ptr, t_Fc.ptr, t_cm.ptr, t_cv.ptr, 1e-12, 100000, t_u.ptr, t_cit.ptr,
                            stream=stream)
    ctx.synchronize()  # (warm: scratch allocated)
    tc0 = time.perf_counter()
    ctx.coarse_solve_device(dim, nc, t_Kc.ptr, t_Fc.ptr, t_cm.ptr, t_cv.ptr, 1e-12, 100000, t_u.ptr, t_cit.ptr,
                            stream=stream)
    ctx.synchronize()
    coarse_ms = (time.perf_counter() - tc0) * 1e3  # one warm call, wall clock

    s = 3 ** dim
    f_iter = (2 * s - 1 + 10) * (nf_arr - 1) ** dim                      # per cell
    cs = np.array([ctx.uses_coarse_space(dim, int(k), opts[main_prec], case["coef_kind"], quads is not None)
                   for k in nf_arr])
    f_iter = f_iter + (cs > 0) * (8 * (nf_arr - 1) ** dim + 2 * 49 * 49)
    f_iter = f_iter + (cs == 2) * np.array([f_ml(int(k)) for k in nf_arr])
    w = 8 if main_prec == msfem.PREC_FP64 else 4
    stride = msfem.coef_stride(dim, case["coef_kind"])
    b_iter = 11 * (nf_arr + 1) ** dim * w + nf_arr ** dim * stride * w   # per cell
    flops = float((its.sum(axis=1) * f_iter).sum())
    peak = PEAK_FP64_TFLOPS if main_prec == msfem.PREC_FP64 else PEAK_FP32_TFLOPS
    achieved = flops / (solve_ms * 1e-3) / 1e12
    b_cell = nf_arr ** dim * stride * 8 + nb * (nf_arr + 1) ** dim * 8 + (nb * nb + nb) * 8
    cells_per_s = n / (ms * 1e-3)
    classes = {int(k): int(v) for k, v in zip(*np.unique(nf_arr, return_counts=True))}

    result = {
        "metric": METRIC,
        "value": round(cells_per_s, 1),
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "local_solves_per_s": round(cells_per_s * nb, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_prec == msfem.PREC_FP64 else "f32 (FP64 refinement)",
        "data": "synthetic (Philox-4x32-10 fields generated on the host, resident in HBM)",
        "config": {"workload": f"{args.workload}: {desc}", "coarse_cells": n, "dim": dim,
                   "fine_per_cell": classes, "local_solves": n * nb, "path": pname[main_prec],
                   "autotuned_classes": tuned, "launch": "HIP graph, one concurrent branch per size class",
                   "two_level_cells": int((cs > 0).sum()), "intermediate_level_cells": int((cs == 2).sum())},
        "roofline": {"bound": "mfma", "engine": "valu", "kernel": f"local_solve_{dim}d ({pname[main_prec]} CG)",
                     "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4), "traffic": TRAFFIC_BYTES.get((args.workload, main_prec)),
                     "kernel_ms": round(solve_ms, 4),
                     "flops_per_launch": flops,
                     "lds_model_TBps": round(float((its.sum(axis=1) * b_iter).sum()) / (solve_ms * 1e-3) / 1e12, 2),
                     "hbm_GBps": round(float(b_cell.sum()) / (ms * 1e-3) / 1e9, 2)},
        "iterations": {pname[p] + "_mean": round(float(r["its"].mean()), 1) for p, r in res.items()},
        "variants": {pname[p] + "_solve_kernel_ms": round(r["solve_kernel_ms"], 4) for p, r in res.items()},
    }
    if energy:
        result["variants"]["fp32_energy_bound_max"] = res[msfem.PREC_FP32_REFINE]["energy_bound_max"]
    result["variants"].update({pname[main_prec] + "_partition_of_unity_cells_per_s": round(n / (pou_ms * 1e-3), 1)})
    result["variants"].update({"coarse_solve_ms": round(coarse_ms, 3), "coarse_cg_iterations": int(t_cit.numpy()[0]),
                               "graph_step_device_ms": round(graph_dev_ms, 4)})
    if args.workload == "reference":
        # the reference's main also runs the single-scale Q1 solve on the fine mesh
        # (source/main.cxx:33-35): 8 * 128 = 1024 fine cells per side, same data
        nfine = nc * int(nf_arr[0])
        unit = np.array([[[0.0, 1.0], [0.0, 1.0]]])
        fcoef = msfem.coef_reference(2, np.array([nfine], np.int32), unit)
        tf0 = time.perf_counter()
        _, fit = ctx.fine_solve(nfine, msfem.COEF_TENSOR_QP, fcoef, msfem.RHS_CONST, 2.0, msfem.BC_REFERENCE,
                                rel_tol=1e-12)
        result["variants"].update({"fine_fem_%d_solve_ms_incl_upload" % nfine:
                                   round((time.perf_counter() - tf0) * 1e3, 2), "fine_fem_cg_iterations": fit})

    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(dim, case["box"], nf_arr, case["coef_kind"], case["coef"],
                                              case["rhs_kind"], case["rhs_value"], n,
                                              args.cpu_sample or (1024 if args.workload == "tensor128" else 2048),
                                              quads)
    ctx.close()
    print(json.dumps(result))


def copy_calibration(mib):
    """out = x over mib MiB of FP32 (msfem_nn_axpby with b = 0): reads and writes
    exactly mib MiB each per launch."""
    import msfem
    import surrogate as S
    ctx = msfem.Context(0)
    n = mib * (1 << 20) // 4
    x, y = ctx.alloc(n, np.float32), ctx.alloc(n, np.float32)
    for _ in range(3):
        S._ok(S._nn().msfem_nn_axpby(n, 1.0, x.ptr, 0.0, None, y.ptr, ctx.stream))
    ctx.synchronize()
    print(json.dumps({"copy_calibration": {"kernel": "axpby_kernel", "bytes_read": n * 4, "bytes_written": n * 4,
                                           "launches": 3}}))
    ctx.close()


def cpu_baseline(dim, box, nf_arr, coef_kind, coef, rhs_kind, rhs_value, n, sample_cells, quads=None):
    """The oracle's restatement of the reference recipe (CG + SSOR(1.6), abs tol
    1e-12, 1000 iterations) on an evenly spaced sample of the cells, run as the
    reference runs: `mpiexec -n P` with one single-threaded rank per host core,
    contiguous blocks of cells per rank, MPI_Barrier + MPI_Wtime around the
    build, max over ranks (oracle/cpu_baseline_mpi).  Without MPICH on the host,
    the same recipe with one OpenMP thread per core instead."""
    import shutil
    import subprocess
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "oracle"))  # the oracle, for the CPU-baseline leg only
    import msfem
    import oracle as O
    sample = np.unique(np.linspace(0, n - 1, min(sample_cells, n)).astype(np.int64))
    stride_c = msfem.coef_stride(dim, coef_kind)
    nf_arr = np.asarray(nf_arr, np.int64)
    offs = np.concatenate([[0], np.cumsum(nf_arr ** dim * stride_c)])
    sub_coef = np.concatenate([coef[offs[i]:offs[i + 1]] for i in sample])
    cores = len(os.sched_getaffinity(0))
    geo = np.asarray(quads if quads is not None else box, np.float64)[sample]
    exe = os.path.join(ROOT, "oracle", "cpu_baseline_mpi")
    mpiexec = shutil.which("mpiexec") or ("/opt/conda/bin/mpiexec" if os.path.exists("/opt/conda/bin/mpiexec")
                                          else None)
    if os.path.exists(exe) and mpiexec:
        with tempfile.NamedTemporaryFile(suffix=".bin", delete=False) as f:
            np.array([dim, len(sample), coef_kind, rhs_kind, 1 if quads is not None else 0, 1000],
                     np.int32).tofile(f)
            np.array([rhs_value, 1.6, 1e-12], np.float64).tofile(f)
            nf_arr[sample].astype(np.int32).tofile(f)
            np.ascontiguousarray(geo).tofile(f)
            np.ascontiguousarray(sub_coef, np.float64).tofile(f)
            path = f.name
        try:
            p = subprocess.run([mpiexec, "-n", str(cores), exe, path], capture_output=True, text=True, timeout=600)
        finally:
            os.unlink(path)
        if p.returncode != 0:
            raise RuntimeError(f"cpu_baseline_mpi failed: {p.stderr[-500:]}")
        r = json.loads(p.stdout.strip().splitlines()[-1])
        t = r["seconds"]
        return {"value": round(len(sample) / t, 2), "unit": "cells/s", "cores": cores, "kind": "port",
                "sample": f"{len(sample)} of the {n} cells (evenly spaced), mpiexec -n {cores} (one rank per core, "
                          f"contiguous blocks of cells), reference recipe CG+SSOR(1.6), abs tol 1e-12, max 1000 it "
                          f"(not converged: {r['fails']}, mean it {r['iters_mean']:.1f}), {t:.1f} s (max over ranks)"}
    t0 = time.perf_counter()
    if quads is not None:
        r = O.batch_build_quad(geo, nf_arr[sample], coef_kind, sub_coef, rhs_kind, rhs_value,
                               solver=O.SOLVER_SSOR_PCG, n_threads=cores)
    else:
        r = O.batch_build(dim, geo, nf_arr[sample], coef_kind, sub_coef, rhs_kind, rhs_value,
                          solver=O.SOLVER_SSOR_PCG, n_threads=cores)
    t = time.perf_counter() - t0
    return {"value": round(len(sample) / t, 2), "unit": "cells/s", "cores": cores, "kind": "port",
            "sample": f"{len(sample)} of the {n} cells (evenly spaced), {cores} OpenMP threads (no MPICH on this "
                      f"host), reference recipe CG+SSOR(1.6), abs tol 1e-12, max 1000 it (not converged: "
                      f"{r['fails']}, mean it {r['iters'].mean():.1f}), {t:.1f} s wall"}


if __name__ == "__main__":
    main()
