"""Autotuner of the local-solve launch shapes.

For every resolution class of a workload it times each compiled launch shape
(msfem_variant_count ids; rows per lane, waves per problem, occupancy target)
on a sample of that class's cells with HIP events on the solve stream, keeps
the fastest, and writes autotune_table.json next to this file.  The engine
applies the table through Context.load_variant_table() / msfem_set_variant.

    python autotune.py --workload lognormal64 --precision fp64
    python autotune.py --workload grf128 --precision fp32

Samples of a few thousand cells matter: the shapes differ in how many
workgroups share a CU, which a small sample (few workgroups per CU) hides.
"""
import argparse
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "autotune_table.json")
sys.path.insert(0, HERE)

import msfem  # noqa: E402


def _subcase(case, idx):
    dim = case["dim"]
    stride = msfem.coef_stride(dim, case["coef_kind"])
    sizes = np.array([int(n) ** dim * stride for n in case["n_fine"]])
    offs = np.concatenate([[0], np.cumsum(sizes)])
    coef = np.concatenate([case["coef"][offs[i]:offs[i + 1]] for i in idx])
    out = dict(case)
    out.update(n_fine=np.asarray(case["n_fine"])[idx], box=np.asarray(case["box"])[idx], coef=coef)
    return out


def time_class(ctx, case, precision, rel_tol, variant, repeats):
    """Best-of-`repeats` device time (ms) of the local solves of `case`."""
    dim = case["dim"]
    n = len(case["n_fine"])
    nb = 1 << dim
    d_box = ctx.to_device(np.asarray(case["box"]).reshape(-1))
    d_coef = ctx.to_device(case["coef"])
    d_Ae = ctx.alloc(n * nb * nb, np.float64)
    d_Fe = ctx.alloc(n * nb, np.float64)
    d_st = ctx.alloc(n * nb, np.int32)
    d_it = ctx.alloc(n * nb, np.int32)
    opts = msfem.default_opts(precision=precision, rel_tol=rel_tol, variant=variant)
    ctx.set_profiling(True)
    best = float("inf")
    for _ in range(repeats + 1):  # first run warms up
        ctx.batch_build_device(dim, case["n_fine"], d_box.ptr, case["coef_kind"], d_coef.ptr, d_Ae.ptr, d_Fe.ptr,
                               None, d_it.ptr, d_st.ptr, rhs_kind=case["rhs_kind"], rhs_value=case["rhs_value"],
                               opts=opts, stream=ctx.stream)
        best = min(best, ctx.profile()[0])
    ctx.synchronize()
    ok = int(d_st.numpy().sum()) == 0
    iters = float(d_it.numpy().mean())
    return best, ok, iters


def tune(ctx, case, precision, rel_tol, max_cells=1024, repeats=2, log=print):
    dim = case["dim"]
    entries = []
    nf_all = np.asarray(case["n_fine"])
    for nf in sorted(set(int(v) for v in nf_all)):
        idx = np.where(nf_all == nf)[0][:max_cells]
        sub = _subcase(case, idx)
        res = {}
        for v in range(1, msfem.variant_count(dim, nf) + 1):
            ms, ok, iters = time_class(ctx, sub, precision, rel_tol, v, repeats)
            res[v] = ms if ok else float("inf")
            log(f"  dim {dim} nf {nf:3d} prec {precision} variant {v}: {ms:9.3f} ms "
                f"({len(idx)} cells, {iters:.1f} it/solve){'' if ok else ' NOT CONVERGED'}")
        best = min(res, key=res.get)
        entries.append({"dim": dim, "nf": nf, "precision": precision, "coef_kind": int(case["coef_kind"]),
                        "variant": best, "ms": {str(k): round(v, 4) for k, v in res.items()}, "cells": len(idx)})
    return entries


def save(entries, path=TABLE):
    table = {"entries": []}
    if os.path.exists(path):
        table = json.load(open(path))
    keyf = lambda e: (e["dim"], e["nf"], e["precision"], e["coef_kind"])
    merged = {keyf(e): e for e in table["entries"]}
    for e in entries:
        merged[keyf(e)] = e
    table["entries"] = [merged[k] for k in sorted(merged)]
    json.dump(table, open(path, "w"), indent=1)
    return path


def main():
    import workloads as C
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="lognormal64", choices=["lognormal64", "grf128", "tensor128", "lognormal3d"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--max-cells", type=int, default=4096)
    ap.add_argument("--out", default=TABLE)
    args = ap.parse_args()
    prec = msfem.PREC_FP64 if args.precision == "fp64" else msfem.PREC_FP32_REFINE
    rel_tol = 1e-13 if prec == msfem.PREC_FP64 else 1e-8
    if args.workload == "lognormal64":
        case = C.lognormal_cells()
    elif args.workload == "grf128":
        case = C.grf_cells()
    elif args.workload == "tensor128":
        case = C.tensor_cells()
    else:
        case = C.lognormal_cells(dim=3, nc=16, nf=16, seed=5)
    with msfem.Context(0) as ctx:
        entries = tune(ctx, case, prec, rel_tol, args.max_cells)
    print("wrote", save(entries, args.out))
    for e in entries:
        print(json.dumps(e))


if __name__ == "__main__":
    main()
