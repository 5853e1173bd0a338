"""The CPU baseline's MPI driver (oracle/cpu_baseline_mpi, run by bench.py as
`mpiexec -n P`): on a small batch split over two ranks it runs the same recipe
as the in-process oracle (CG + SSOR(1.6), absolute 1e-12, 1000 iterations) --
the same iteration counts for box cells and general quadrilaterals."""
import os
import shutil
import sys

import numpy as np
import pytest

import oracle as O
import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "cpu_baseline_mpi")
MPIEXEC = shutil.which("mpiexec") or ("/opt/conda/bin/mpiexec" if os.path.exists("/opt/conda/bin/mpiexec") else None)

pytestmark = pytest.mark.skipif(not (os.path.exists(EXE) and MPIEXEC), reason="no MPICH on this host")


def run(monkeypatch, case, sample, quads=None):
    sys.path.insert(0, ROOT)
    import bench
    monkeypatch.setattr(os, "sched_getaffinity", lambda pid: {0, 1})  # two ranks
    return bench.cpu_baseline(case["dim"], case["box"], case["n_fine"], case["coef_kind"], case["coef"],
                              case["rhs_kind"], case["rhs_value"], len(case["n_fine"]), sample, quads)


def test_mpi_driver_matches_oracle_boxes(monkeypatch):
    case = W.lognormal_cells(dim=2, nc=4, nf=16, seed=2)
    r = run(monkeypatch, case, 16)
    assert r["cores"] == 2 and "mpiexec -n 2" in r["sample"] and "not converged: 0" in r["sample"]
    ref = O.batch_build(2, case["box"], case["n_fine"], case["coef_kind"], case["coef"], case["rhs_kind"],
                        case["rhs_value"], solver=O.SOLVER_SSOR_PCG, n_threads=1)
    assert f"mean it {ref['iters'].mean():.1f}" in r["sample"]


def test_mpi_driver_matches_oracle_quads(monkeypatch):
    case = W.quad_cells(nc=4, nf=16)
    r = run(monkeypatch, case, 16, case["vertices"])
    ref = O.batch_build_quad(case["vertices"], case["n_fine"], case["coef_kind"], case["coef"], case["rhs_kind"],
                             case["rhs_value"], solver=O.SOLVER_SSOR_PCG, n_threads=1)
    assert f"mean it {ref['iters'].mean():.1f}" in r["sample"] and "not converged: 0" in r["sample"]
