"""The C ABI of libmsfem_hip.so without a GPU: the library loads, exports every
symbol include/msfem.h declares, its struct layouts agree with the ctypes
mirrors, and it fails cleanly when no device is present."""
import ctypes
import os
import re
import subprocess

import pytest

import msfem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "msfem.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(msfem_\w+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(msfem.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(msfem.LIB_PATH)
    for name in declared_functions():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", msfem.LIB_PATH], capture_output=True, text=True).stdout
    for name in declared_functions():
        assert re.search(rf"\bT {name}\b", out), name


def test_struct_layouts_match_ctypes(tmp_path):
    src = tmp_path / "sizes.c"
    src.write_text("""
#include <stdio.h>
#include <stddef.h>
#include "msfem.h"
int main(void) {
  printf("%zu %zu %zu\\n", sizeof(msfem_solver_opts), sizeof(msfem_batch), sizeof(msfem_batch_out));
  printf("%zu %zu %zu\\n", offsetof(msfem_solver_opts, rel_tol), offsetof(msfem_solver_opts, variant),
         offsetof(msfem_batch, rhs_value));
  printf("%zu %zu %zu\\n", offsetof(msfem_batch, x0), offsetof(msfem_batch, vertices), offsetof(msfem_batch_out, resid));
  printf("%zu %zu\\n", offsetof(msfem_solver_opts, coarse_space), offsetof(msfem_solver_opts, q1_start));
  printf("%zu %zu\\n", offsetof(msfem_solver_opts, energy_tol), offsetof(msfem_batch_out, energy_bound));
  return 0;
}
""")
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    C = ctypes
    assert vals == [C.sizeof(msfem.SolverOpts), C.sizeof(msfem.Batch), C.sizeof(msfem.BatchOut),
                    msfem.SolverOpts.rel_tol.offset, msfem.SolverOpts.variant.offset, msfem.Batch.rhs_value.offset,
                    msfem.Batch.x0.offset, msfem.Batch.vertices.offset, msfem.BatchOut.resid.offset,
                    msfem.SolverOpts.coarse_space.offset, msfem.SolverOpts.q1_start.offset,
                    msfem.SolverOpts.energy_tol.offset, msfem.BatchOut.energy_bound.offset]


def test_default_opts_and_version():
    o = msfem.default_opts()
    assert o.precision == msfem.PREC_FP64 and o.rel_tol == 1e-13 and o.max_iter > 1000
    assert o.coarse_space == 1 and o.q1_start == 1 and o.energy_tol == 0
    assert msfem.lib().msfem_version() == 3


def test_no_device_is_a_clean_error():
    try:
        msfem.Context(0).close()
        pytest.skip("a GPU is present")
    except msfem.MsfemError as e:
        assert e.code == msfem.ERR_NODEVICE


def test_gather_weights_host():
    import numpy as np
    u = np.arange(9.0)  # 3x3 coarse nodes, nc = 2
    w = msfem.gather_weights(2, 2, u)
    assert w.tolist() == [[0, 1, 3, 4], [1, 2, 4, 5], [3, 4, 6, 7], [4, 5, 7, 8]]
