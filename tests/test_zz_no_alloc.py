"""The allocation-free coarse solve (msfem_coarse_solve_device_noalloc, its
scratch in the library's fixed device array) against msfem_coarse_solve_device
(context-owned scratch), both through the C ABI: the same u bitwise and the
same iteration count, on two calls in a row with the second system smaller,
in 2D and 3D; a system whose scratch exceeds the fixed array is rejected."""
import sys

import numpy as np
import pytest

import msfem
import msfem_cases as C

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = msfem.Context(0)
    yield c
    c.close()


def coarse_system(ctx, dim, nc, nf, bc):
    case = C.lognormal_cells(dim=dim, nc=nc, nf=nf, seed=11)
    out = C.gpu_build(ctx, case)
    nn = (nc + 1) ** dim
    d = dict(Ae=ctx.to_device(out["Ae"].reshape(-1)), Fe=ctx.to_device(out["Fe"].reshape(-1)),
             Kc=ctx.alloc(nn * 3 ** dim, np.float64), Fc=ctx.alloc(nn, np.float64), cm=ctx.alloc(nn, np.uint8),
             cv=ctx.alloc(nn, np.float64), u=ctx.alloc(nn, np.float64), it=ctx.alloc(1, np.int32))
    ctx.coarse_assemble_device(dim, nc, d["Ae"].ptr, d["Fe"].ptr, bc, d["Kc"].ptr, d["Fc"].ptr, d["cm"].ptr,
                               d["cv"].ptr, stream=ctx.stream)
    return d


def solve(ctx, fn, dim, nc, d):
    msfem._check(fn(ctx.h, dim, nc, d["Kc"].ptr, d["Fc"].ptr, d["cm"].ptr, d["cv"].ptr, 1e-13, 10000, d["u"].ptr,
                    d["it"].ptr, ctx.stream), ctx.h)
    ctx.synchronize()
    return d["u"].numpy().copy(), int(d["it"].numpy()[0])


@pytest.mark.parametrize("dim,sizes", [(2, (64, 16)), (3, (16, 8))])
def test_noalloc_matches_device_entry(ctx, dim, sizes):
    L = msfem.lib()
    for k, nc in enumerate(sizes):
        bc = msfem.BC_REFERENCE if k == 0 else msfem.BC_ZERO_DIRICHLET
        say(f"no-alloc coarse solve dim={dim} nc={nc} bc={bc} (call {k + 1} of 2)")
        d = coarse_system(ctx, dim, nc, 8 if dim == 3 else 16, bc)
        u_ref, it_ref = solve(ctx, L.msfem_coarse_solve_device, dim, nc, d)
        u, it = solve(ctx, L.msfem_coarse_solve_device_noalloc, dim, nc, d)
        assert it == it_ref and it <= 30
        assert np.array_equal(u, u_ref)
        for a in d.values():
            a.free()


def test_noalloc_rejects_oversize(ctx):
    d = {k: ctx.alloc(8, np.float64) for k in ("Kc", "Fc", "cm", "cv", "u", "it")}
    for dim, nc in ((2, 512), (3, 64), (2, 100000)):
        with pytest.raises(msfem.MsfemError) as e:
            solve(ctx, msfem.lib().msfem_coarse_solve_device_noalloc, dim, nc, d)
        assert e.value.code == msfem.ERR_ARG and "scratch" in str(e.value)
    for a in d.values():
        a.free()
