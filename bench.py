"""Benchmark of the MsFEM basis build on one MI355X.

Workload (default): the batched plain basis build on the unit square, 64x64
coarse Q1 cells, each refined to a 64x64 fine Q1 grid (4225 nodes, 4 basis
problems per cell, 16384 local solves), alpha = exp(g) with g ~ N(0,1) i.i.d.
per fine cell (Philox, seed 2) clipped to contrast 1e3, f = 1, zero Dirichlet.

One step = every cell's 4 basis solves + A^K, F^K + the coarse assembly, with
all inputs already resident in HBM.  The dominant kernel (the local solves) is
timed with HIP events on the stream it runs on; algorithmic work per CG
iteration is F_iter = (2*9 - 1 + 10) * N_I flops (9-point stencil + CG vector
updates) and B_iter = 11 N w + n^2 w_c bytes.

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
for p in (PKG, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def workload(name):
    import msfem_cases as C
    if name == "lognormal64":
        return C.lognormal_cells(dim=2, nc=64, nf=64, seed=2), 64
    if name == "lognormal3d":
        return C.lognormal_cells(dim=3, nc=16, nf=16, seed=5), 16
    raise SystemExit(f"unknown workload {name}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="lognormal64")
    ap.add_argument("--cpu-sample", type=int, default=64, help="cells in the CPU-baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the local solves of each coarse cell are independent and "
              "the multi-GPU split (cells sharded, one coarse-system reduction) is not built.", file=sys.stderr)
        sys.exit(2)

    import msfem

    case, nc = workload(args.workload)
    dim = case["dim"]
    n = len(case["n_fine"])
    nf = int(case["n_fine"][0])
    nb = 1 << dim
    nn = (nf + 1) ** dim
    ni = (nf - 1) ** dim
    ncf = nf ** dim

    ctx = msfem.Context(0)
    t_box = ctx.to_device(case["box"].reshape(-1))
    t_coef = ctx.to_device(case["coef"])
    t_Ae = ctx.alloc(n * nb * nb, np.float64)
    t_Fe = ctx.alloc(n * nb, np.float64)
    t_bases = ctx.alloc(n * nb * nn, np.float64)
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
        ctx.batch_build_device(dim, case["n_fine"], t_box.ptr, case["coef_kind"], t_coef.ptr,
                               t_Ae.ptr, t_Fe.ptr, t_bases.ptr, t_it.ptr, t_st.ptr,
                               rhs_kind=case["rhs_kind"], rhs_value=case["rhs_value"], opts=opts, stream=stream)
        ctx.coarse_assemble_device(dim, nc, t_Ae.ptr, t_Fe.ptr, msfem.BC_ZERO_DIRICHLET, t_Kc.ptr,
                                   t_Fc.ptr, t_cm.ptr, t_cv.ptr, stream=stream)

    def run(opts, steps, warmup):
        for _ in range(warmup):
            step(opts)
        ctx.synchronize()
        solve_ms = []
        t0 = time.perf_counter()
        for _ in range(steps):
            step(opts)
            solve_ms.append(ctx.profile()[0])
        ctx.synchronize()
        t1 = time.perf_counter()
        assert int(t_st.numpy().sum()) == 0, "local solves did not converge"
        return (t1 - t0) * 1e3 / steps, float(np.mean(solve_ms)), t_it.numpy().astype(np.int64)

    # plain FP64 path (the parity path)
    o64 = msfem.default_opts(precision=msfem.PREC_FP64)
    ms64, solve64, it64 = run(o64, args.steps, args.warmup)
    # FP32 CG + FP64 refinement stopped where the energy-norm error is <= 1e-6
    o32 = msfem.default_opts(precision=msfem.PREC_FP32_REFINE, rel_tol=1e-8)
    ms32, solve32, it32 = run(o32, args.steps, args.warmup)
    # coarse solve (timed apart: one single-workgroup FP64 CG)
    step(o64)
    ctx.synchronize()
    tc0 = time.perf_counter()
    ctx.coarse_solve_device(dim, nc, t_Kc.ptr, t_Fc.ptr, t_cm.ptr, t_cv.ptr, 1e-12,
                            100000, t_u.ptr, t_cit.ptr, stream=stream)
    ctx.synchronize()
    coarse_ms = (time.perf_counter() - tc0) * 1e3

    s = 3 ** dim
    f_iter = (2 * s - 1 + 10) * ni
    b_iter = 11 * nn * 8 + ncf * 8
    flops64 = float(it64.sum()) * f_iter
    achieved = flops64 / (solve64 * 1e-3) / 1e12
    b_cell = ncf * 8 + nb * nn * 8 + (nb * nb + nb) * 8
    cells_per_s = n / (ms64 * 1e-3)

    result = {
        "metric": "coarse-cell bases built/sec: plain FP64 vs batched FP32+surrogate @1e-6 err",
        "value": round(cells_per_s, 1),
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms64, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Philox-4x32-10 lognormal field, generated on the host, resident in HBM)",
        "config": {"workload": f"{args.workload}: {nc}^{dim} coarse Q1 cells x {nf}^{dim} fine Q1 per cell, "
                               f"{n * nb} local solves, FP64, rel. tol 1e-13",
                   "coarse_cells": n, "fine_per_cell": nf, "dim": dim, "local_solves": n * nb,
                   "autotuned_classes": tuned},
        "roofline": {"bound": "valu", "kernel": "local_solve_2d (FP64 CG)" if dim == 2 else "local_solve_3d",
                     "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": None,
                     "kernel_ms": round(solve64, 4), "flops_per_launch": flops64,
                     "lds_model_TBps": round(float(it64.sum()) * b_iter / (solve64 * 1e-3) / 1e12, 2),
                     "hbm_GBps": round(n * b_cell / (ms64 * 1e-3) / 1e9, 2)},
        "iterations": {"fp64_mean": round(float(it64.mean()), 1), "fp64_p99": int(np.percentile(it64, 99)),
                       "fp32_refine_mean": round(float(it32.mean()), 1),
                       "fp32_refine_p99": int(np.percentile(it32, 99))},
        "variants": {"fp64_cells_per_s": round(cells_per_s, 1),
                     "fp32_refine_1e-6_energy_cells_per_s": round(n / (ms32 * 1e-3), 1),
                     "fp32_refine_ms_per_step": round(ms32, 4), "fp32_refine_solve_kernel_ms": round(solve32, 4),
                     "coarse_solve_ms": round(coarse_ms, 3), "coarse_cg_iterations": int(t_cit.numpy()[0])},
    }

    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))  # the checker, for the CPU-baseline leg only
        import msfem_cases as C
        sample = np.linspace(0, n - 1, min(args.cpu_sample, n)).astype(np.int64)
        sub = C.lognormal_cells(dim=dim, nc=nc, nf=nf, seed=2 if dim == 2 else 5, which=sample)
        cores = len(os.sched_getaffinity(0))
        tcpu0 = time.perf_counter()
        r = C.oracle_build(sub, want_bases=False, solver=1, n_threads=cores)
        tcpu = time.perf_counter() - tcpu0
        result["cpu_baseline"] = {
            "value": round(len(sample) / tcpu, 2), "unit": "cells/s", "cores": cores, "kind": "port",
            "sample": f"{len(sample)} cells spread over the same batch, reference recipe CG+SSOR(1.6), abs tol 1e-12,"
                      f" max 1000 it (fails: {r['fails']}, mean it {r['iters'].mean():.1f}), {tcpu:.1f} s wall"}
    ctx.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
