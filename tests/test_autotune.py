"""The launch-shape autotuner (autotune.py): sub-batch slicing and table merging
on the CPU; on the GPU, one tuning pass picks the fastest converged shape of
every class, the engine applies the written table, and the tuned build gives
the same bases as the default one (a launch shape changes the schedule, never
the converged values beyond the solver tolerance)."""
import json

import numpy as np
import pytest

import autotune as A
import msfem
import workloads as W


def test_subcase_slices_mixed_resolutions():
    case = W.tensor_cells(nc=4, seed=3, sigma=1.5, n_fine=[16, 32, 16, 64] * 4)
    idx = np.array([2, 3, 0])
    sub = A._subcase(case, idx)
    stride = msfem.coef_stride(2, case["coef_kind"])
    sizes = [int(n) ** 2 * stride for n in case["n_fine"]]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    want = np.concatenate([case["coef"][offs[i]:offs[i + 1]] for i in idx])
    assert np.array_equal(sub["n_fine"], np.asarray(case["n_fine"])[idx])
    assert np.array_equal(sub["box"], np.asarray(case["box"])[idx])
    assert np.array_equal(sub["coef"], want)
    assert sub["coef_kind"] == case["coef_kind"] and sub["rhs_kind"] == case["rhs_kind"]


def test_save_merges_by_class(tmp_path):
    path = tmp_path / "table.json"
    e = lambda nf, prec, kind, v: {"dim": 2, "nf": nf, "precision": prec, "coef_kind": kind, "variant": v,
                                  "ms": {}, "cells": 1}
    A.save([e(64, 1, 0, 1), e(16, 0, 1, 3)], path)
    A.save([e(64, 1, 0, 4), e(32, 1, 0, 2)], path)  # replaces (2, 64, 1, 0), adds (2, 32, 1, 0)
    t = json.load(open(path))["entries"]
    keys = [(x["dim"], x["nf"], x["precision"], x["coef_kind"]) for x in t]
    assert keys == sorted(keys) and len(keys) == 3
    assert {k: x["variant"] for k, x in zip(keys, t)}[(2, 64, 1, 0)] == 4


def test_grf_workload_matches_the_serving_fields():
    """workloads.grf_cells draws the fields of surrogate.TileSet (same generator, seed, clip)."""
    case = W.grf_cells(nc=4, nf=16, seed=7, which=[5, 2])
    g = msfem.grf_log(16, 16, 7, 4.0, 0.5 * np.log(1e3))
    assert np.array_equal(case["coef"].reshape(2, -1), np.exp(g.astype(np.float64))[[5, 2]])
    assert case["coef"].max() / case["coef"].min() <= 1e3 * (1 + 1e-12)


@pytest.mark.gpu
def test_tune_picks_fastest_and_engine_applies_it(tmp_path):
    print("[gpu-case] autotune lognormal 2D nf=16/32, 64 cells each", flush=True)
    case = W.tensor_cells(nc=8, seed=3, sigma=1.5, n_fine=[16, 32] * 32)
    with msfem.Context(0) as ctx:
        entries = A.tune(ctx, case, msfem.PREC_FP32_REFINE, 1e-8, max_cells=64, repeats=1, log=lambda m: None)
        assert [e["nf"] for e in entries] == [16, 32]
        for e in entries:
            ms = {int(k): v for k, v in e["ms"].items()}
            assert set(ms) == set(range(1, msfem.variant_count(2, e["nf"]) + 1))
            assert e["variant"] == min(ms, key=ms.get) and np.isfinite(ms[e["variant"]])
        path = A.save(entries, tmp_path / "table.json")
        ref = W.gpu_build(ctx, case, precision=msfem.PREC_FP32_REFINE, rel_tol=1e-12)
        assert ctx.load_variant_table(str(path)) == 2
        out = W.gpu_build(ctx, case, precision=msfem.PREC_FP32_REFINE, rel_tol=1e-12)
    assert (out["status"] == 0).all()
    scale = np.abs(ref["bases"]).max()
    assert np.abs(out["bases"] - ref["bases"]).max() < 1e-9 * scale
    assert np.abs(out["Ae"] - ref["Ae"]).max() < 1e-9 * np.abs(ref["Ae"]).max()
