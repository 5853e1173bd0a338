"""Benchmark of the MsFEM basis build on one MI355X.

Default workload (the one the headline line is quoted on): the batched plain
basis build on the unit square, 64x64 coarse Q1 cells, each refined to a 64x64
fine Q1 grid (4225 nodes, 4 basis problems per cell, 16384 local solves),
alpha = exp(g) with g ~ N(0,1) i.i.d. per fine cell (Philox, seed 2) clipped
to contrast 1e3, f = 1, zero Dirichlet.  Other workloads (--workload):
  tensor128    128x128 coarse cells, fine 16..128 per cell by probe contrast,
               SPD tensor per fine cell, f = sin(pi x) sin(pi y)   (uneven batch)
  lognormal3d  16^3 coarse cells x 16^3 fine, lognormal alpha, f = 1
  periodic8    8x8 coarse cells x 32^2 fine, alpha = 2 + sin(2 pi x/eps) cos(2 pi y/eps)
  reference    the reference default: 8x8 x 128^2, MatrixCoeff (PI_D), f = 2
  surrogate128 128x128 coarse cells x 64^2 fine, alpha = exp(g), g a Gaussian random
               field (correlation length 4 fine cells, contrast <= 1e3, seed 41);
               plain FP64 vs FP32+refinement, each with and without the trained
               surrogate's prediction as initial guess (inference inside the step)

One step = every cell's basis solves + A^K, F^K + the coarse assembly, with all
inputs already resident in HBM.  The dominant kernel (the local solves) is timed
with HIP events on the stream it runs on.  Algorithmic work per CG iteration of
a problem with N_I interior nodes: F_iter = (2*3^d - 1 + 10) N_I flops (stencil +
CG vector updates), B_iter = 11 N w + n^d w_c bytes (LDS byte model).

The CPU baseline is the oracle's restatement of the reference's own recipe
(CG + SSOR(1.6), absolute tolerance 1e-12, 1000 iterations;
include/base/diffusion_problem_basis.tpp:295-317) on a sample of the same cells,
one OpenMP thread per available core.
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

PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_VALU_TFLOPS = 78.6  # non-packed FP32 FMA issue rate equals FP64 (packed FP32 is disabled)
PEAK_BF16_MFMA_TFLOPS = 2500.0  # dense
# HBM bytes per launch of the dominant kernel, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE passes scaled by a plain copy of known size (--copy-calibration):
# profiles/local_solve_2d_lognormal64_hbm_traffic.md
TRAFFIC_BYTES = {("lognormal64", 0): 2.168743e9}
METRIC = "coarse-cell bases built/sec: plain FP64 vs batched FP32+surrogate @1e-6 err"


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
    ap.add_argument("--workload", default="lognormal64",
                    choices=["lognormal64", "tensor128", "lognormal3d", "periodic8", "reference", "quad64",
                             "surrogate128"])
    ap.add_argument("--cpu-sample", type=int, default=2048, help="cells in the CPU-baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (3 times): the reference kernel that scales the "
                         "FETCH_SIZE / WRITE_SIZE counters to bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the local solves of each coarse cell are independent and "
              "the multi-GPU split (cells sharded, one coarse-system reduction) is not built.", file=sys.stderr)
        sys.exit(2)

    import msfem

    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.workload == "surrogate128":
        return bench_surrogate(args)
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
    ncn = (nc + 1) ** dim
    t_Kc = ctx.alloc(ncn * 3 ** dim, np.float64)
    t_Fc = ctx.alloc(ncn, np.float64)
    t_cm = ctx.alloc(ncn, np.uint8)
    t_cv = ctx.alloc(ncn, np.float64)
    t_u = ctx.alloc(ncn, np.float64)
    t_cit = ctx.alloc(1, np.int32)
    tuned = ctx.load_variant_table()  # launch shapes picked by autotune.py on this GPU
    ctx.set_profiling(True)
    stream = ctx.stream

    def step(opts):
        ctx.batch_build_device(dim, case["n_fine"], t_box.ptr if t_box is not None else None, case["coef_kind"], t_coef.ptr,
                               t_Ae.ptr, t_Fe.ptr, t_bases.ptr, t_it.ptr, t_st.ptr, rhs_kind=case["rhs_kind"],
                               rhs_value=case["rhs_value"], opts=opts, stream=stream,
                               vertices_ptr=t_vert.ptr if t_vert is not None else None)
        ctx.coarse_assemble_device(dim, nc, t_Ae.ptr, t_Fe.ptr, bc_kind, t_Kc.ptr, t_Fc.ptr, t_cm.ptr, t_cv.ptr,
                                   stream=stream)

    def run(opts):
        for _ in range(args.warmup):
            step(opts)
        ctx.synchronize()
        solve_ms = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step(opts)
            solve_ms.append(ctx.profile()[0])
        ctx.synchronize()
        t1 = time.perf_counter()
        assert int(t_st.numpy().sum()) == 0, "local solves did not converge"
        return (t1 - t0) * 1e3 / args.steps, float(np.mean(solve_ms)), t_it.numpy().astype(np.int64).reshape(n, nb)

    opts = {msfem.PREC_FP64: msfem.default_opts(precision=msfem.PREC_FP64),
            # FP32 CG + FP64 refinement stopped where the energy-norm error is <= 1e-6
            msfem.PREC_FP32_REFINE: msfem.default_opts(precision=msfem.PREC_FP32_REFINE, rel_tol=1e-8)}
    res = {p: run(o) for p, o in opts.items()}
    stream_ms, solve_ms, its = res[main_prec]
    # the headline step: the same build as one replayed HIP graph (a concurrent
    # branch per size class), timed by the wall clock and by device marks
    gopts = msfem.default_opts(precision=main_prec, rel_tol=opts[main_prec].rel_tol, graph=1)
    for _ in range(args.warmup):
        step(gopts)
    ctx.synchronize()
    tg0 = time.perf_counter()
    ctx.mark(0)
    for _ in range(args.steps):
        step(gopts)
    ctx.mark(1)
    ctx.synchronize()
    ms = (time.perf_counter() - tg0) * 1e3 / args.steps
    graph_dev_ms = ctx.elapsed(0, 1) / args.steps
    assert int(t_st.numpy().sum()) == 0, "local solves did not converge (graph)"
    # the same step with the last basis of each cell as 1 - (sum of the others)
    # (exact by linearity: K 1 = 0 and the boundary data sum to 1); reported
    # beside the headline, which solves every basis as the reference does
    popts = msfem.default_opts(precision=main_prec, rel_tol=opts[main_prec].rel_tol, graph=1, partition_of_unity=1)
    for _ in range(args.warmup):
        step(popts)
    ctx.synchronize()
    tp0 = time.perf_counter()
    for _ in range(args.steps):
        step(popts)
    ctx.synchronize()
    pou_ms = (time.perf_counter() - tp0) * 1e3 / args.steps
    assert int(t_st.numpy().sum()) == 0, "local solves did not converge (partition of unity)"
    # coarse solve (timed apart: one single-workgroup FP64 CG)
    step(opts[main_prec])
    ctx.synchronize()
    tc0 = time.perf_counter()
    ctx.coarse_solve_device(dim, nc, t_Kc.ptr, t_Fc.ptr, t_cm.ptr, t_cv.ptr, 1e-12, 100000, t_u.ptr, t_cit.ptr,
                            stream=stream)
    ctx.synchronize()
    coarse_ms = (time.perf_counter() - tc0) * 1e3

    s = 3 ** dim
    f_iter = (2 * s - 1 + 10) * (nf_arr - 1) ** dim                      # per cell
    # two-level classes: restriction and prolongation (2 FMA each per node) and
    # the 49 x 49 coarse product per iteration
    cs = np.array([ctx.uses_coarse_space(dim, int(k), opts[main_prec], case["coef_kind"], quads is not None)
                   for k in nf_arr])
    f_iter = f_iter + cs * (8 * (nf_arr - 1) ** dim + 2 * 49 * 49)
    w = 8 if main_prec == msfem.PREC_FP64 else 4
    stride = msfem.coef_stride(dim, case["coef_kind"])
    b_iter = 11 * (nf_arr + 1) ** dim * w + nf_arr ** dim * stride * w   # per cell
    flops = float((its.sum(axis=1) * f_iter).sum())
    peak = PEAK_FP64_TFLOPS if main_prec == msfem.PREC_FP64 else PEAK_FP32_VALU_TFLOPS
    achieved = flops / (solve_ms * 1e-3) / 1e12
    b_cell = nf_arr ** dim * stride * 8 + nb * (nf_arr + 1) ** dim * 8 + (nb * nb + nb) * 8
    cells_per_s = n / (ms * 1e-3)
    pname = {msfem.PREC_FP64: "fp64", msfem.PREC_FP32_REFINE: "fp32_refine_1e-6_energy"}
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
                   "two_level_cells": int(cs.sum())},
        # compute-bound (the flops ceiling): FP64 vector instructions, whose MI355X
        # peak equals the dense FP64 MFMA figure, 78.6 TFLOP/s
        "roofline": {"bound": "mfma", "engine": "valu", "kernel": f"local_solve_{dim}d ({pname[main_prec]} CG)",
                     "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4), "traffic": TRAFFIC_BYTES.get((args.workload, main_prec)),
                     "kernel_ms": round(solve_ms, 4),
                     "flops_per_launch": flops,
                     "lds_model_TBps": round(float((its.sum(axis=1) * b_iter).sum()) / (solve_ms * 1e-3) / 1e12, 2),
                     "hbm_GBps": round(float(b_cell.sum()) / (ms * 1e-3) / 1e9, 2)},
        "iterations": {p + "_mean": round(float(r[2].mean()), 1) for p, r in
                       ((pname[k], v) for k, v in res.items())},
        "variants": {pname[p] + "_cells_per_s": round(n / (r[0] * 1e-3), 1) for p, r in res.items()},
    }
    result["variants"].update({pname[p] + "_solve_kernel_ms": round(r[1], 4) for p, r in res.items()})
    result["variants"].update({pname[main_prec] + "_partition_of_unity_cells_per_s": round(n / (pou_ms * 1e-3), 1)})
    result["variants"].update({"coarse_solve_ms": round(coarse_ms, 3), "coarse_cg_iterations": int(t_cit.numpy()[0]),
                               "graph_step_device_ms": round(graph_dev_ms, 4),
                               "stream_step_ms": round(stream_ms, 4)})
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
                                              case["rhs_kind"], case["rhs_value"], n, args.cpu_sample, quads)
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
    1e-12, 1000 iterations) on an evenly spaced sample of the cells."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))  # the oracle, for the CPU-baseline leg only
    import msfem
    import oracle as O
    sample = np.unique(np.linspace(0, n - 1, min(sample_cells, n)).astype(np.int64))
    stride_c = msfem.coef_stride(dim, coef_kind)
    offs = np.concatenate([[0], np.cumsum(nf_arr ** dim * stride_c)])
    sub_coef = np.concatenate([coef[offs[i]:offs[i + 1]] for i in sample])
    cores = len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    if quads is not None:
        r = O.batch_build_quad(np.asarray(quads)[sample], nf_arr[sample], coef_kind, sub_coef, rhs_kind, rhs_value,
                               solver=O.SOLVER_SSOR_PCG, n_threads=cores)
    else:
        r = O.batch_build(dim, np.asarray(box)[sample], nf_arr[sample], coef_kind, sub_coef, rhs_kind, rhs_value,
                          solver=O.SOLVER_SSOR_PCG, n_threads=cores)
    t = time.perf_counter() - t0
    return {"value": round(len(sample) / t, 2), "unit": "cells/s", "cores": cores, "kind": "port",
            "sample": f"{len(sample)} of the {n} cells (evenly spaced), reference recipe CG+SSOR(1.6), abs tol "
                      f"1e-12, max 1000 it (not converged: {r['fails']}, mean it {r['iters'].mean():.1f}), "
                      f"{t:.1f} s wall"}


def bench_surrogate(args):
    """surrogate128: the headline comparison on the learned-surrogate workload."""
    import msfem
    import surrogate as S
    ctx = msfem.Context(0)
    ctx.load_variant_table()
    net = S.Surrogate.load(ctx, os.path.join(PKG, "surrogate_weights.npz"))
    data = S.TileSet(ctx, 128 * 128, 41)
    rep = S.serve_report(ctx, net, n_side=128, reps=args.steps, warmup=args.warmup, data=data)
    main_leg = rep["fp32_refine_surrogate_pipelined"]
    pr = rep["prediction"]
    flops = net.flops_per_sample() * data.n
    achieved = flops / (pr["infer_ms"] * 1e-3) / 1e12
    result = {
        "metric": METRIC,
        "value": main_leg["cells_per_s"],
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_leg["ms"],
        "local_solves_per_s": round(main_leg["cells_per_s"] * 4, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 network (f32 accumulate) + f32 CG (FP64 refinement)",
        "data": "synthetic (Philox-4x32-10 Gaussian random fields generated on the host, resident in HBM); "
                "network trained on the GPU from plain-solver bases of 5000 other cells (seed 40)",
        "config": {"workload": "surrogate128: 128^2 coarse cells x 64^2 fine, alpha=exp(g), g Gaussian random field "
                               "(corr. length 4 fine cells, contrast <= 1e3), f=1, 65536 local solves; step = "
                               "inference + batched basis build + coarse assembly",
                   "coarse_cells": data.n, "dim": 2, "local_solves": 4 * data.n,
                   "path": "fp32_refine_1e-6_energy + surrogate initial guess, inference of each 4096-cell chunk "
                           "overlapping the builds of the previous chunk (two streams)",
                   "network_params": net.n_params()},
        "roofline": {"bound": "mfma", "kernel": "surrogate inference (direct 3x3 BF16-MFMA convolutions + glue)",
                     "achieved": round(achieved, 3), "peak": PEAK_BF16_MFMA_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_BF16_MFMA_TFLOPS, 5), "traffic": None,
                     "kernel_ms": pr["infer_ms"], "flops_per_launch": float(flops)},
        "variants": {k: v for k, v in rep.items() if isinstance(v, dict)},
        "speedup_vs_plain_fp64": round(main_leg["cells_per_s"] / rep["fp64"]["cells_per_s"], 3),
        "speedup_vs_fp32_without_surrogate": round(main_leg["cells_per_s"] / rep["fp32_refine"]["cells_per_s"], 3),
    }
    if not args.no_cpu_baseline:
        nf_arr = np.full(data.n, data.nf, np.int64)
        result["cpu_baseline"] = cpu_baseline(2, data.box_host, nf_arr, msfem.COEF_SCALAR_CELL, data.coef_host,
                                              msfem.RHS_CONST, 1.0, data.n, args.cpu_sample)
    data.free()
    ctx.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
