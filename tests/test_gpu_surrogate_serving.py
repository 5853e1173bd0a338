"""What the trained surrogate gains, and the accept/fallback gate, on fresh cells.

Cells: 512 Gaussian-random-field tiles (64x64 fine cells, correlation length
4, contrast <= 1e3, seed 41: not the training seed 40), solved by the FP32 CG
with FP64 refinement at relative residual 1e-8 (the setting whose energy-norm
error vs plain FP64 is <= 1e-6).  Bars:
  * with the trained network's prediction as the initial guess, the mean CG
    iteration count drops at least 10% below the cold start, and below the
    count with an untrained network of the same shape (what training buys);
  * every surrogate-started solve meets the 1e-6 relative energy-norm bar
    against plain FP64;
  * the gate accepts an initial guess that already meets the tolerance (the
    FP64 bases themselves: 0 iterations, returned unchanged) and falls back
    to CG for a corrupted one (NaN entries), which still converges.
"""
import os
import sys

import numpy as np
import pytest

import msfem
import surrogate as S

pytestmark = pytest.mark.gpu
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(S.__file__)), "surrogate_weights.npz")


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = msfem.Context(0)
    yield c
    c.close()


def worst_energy_error(data, bases, ref, cells):
    nf = data.nf
    per = 4 * (nf + 1) ** 2
    worst = 0.0
    for c in cells:
        alpha = data.coef_host[c * nf * nf:(c + 1) * nf * nf].reshape(nf, nf)
        phi = bases[c * per:(c + 1) * per].reshape(4, nf + 1, nf + 1)
        r = ref[c * per:(c + 1) * per].reshape(4, nf + 1, nf + 1)
        worst = max(worst, S.energy_error(alpha, phi, r))
    return worst


def test_surrogate_gain_and_gate(ctx):
    say("surrogate serving: 512 GRF cells, seed 41")
    net = S.Surrogate.load(ctx, WEIGHTS)
    data = S.TileSet(ctx, 512, 41)
    n, per = data.n, 4 * (data.nf + 1) ** 2
    o64 = msfem.default_opts(precision=msfem.PREC_FP64)
    o32 = msfem.default_opts(precision=msfem.PREC_FP32_REFINE, rel_tol=1e-8)
    x0 = ctx.alloc(n * per, np.float64)

    def run(opts, x0_ptr=None):
        data.solve(opts, x0_ptr)
        ctx.synchronize()
        assert not data.status.numpy().any()
        return data.iters.numpy().astype(np.int64), data.bases.numpy().reshape(-1)

    _, ref = run(o64)
    cold, _ = run(o32)
    net.predict_bases(data.x, n, x0.ptr)
    warm, bases = run(o32, x0.ptr)
    untrained = S.Surrogate(ctx, seed=123)
    untrained.dscale = net.dscale
    untrained.predict_bases(data.x, n, x0.ptr)
    rnd, _ = run(o32, x0.ptr)
    say(f"  mean iterations: cold {cold.mean():.1f}  trained {warm.mean():.1f}  untrained {rnd.mean():.1f}  "
        f"fallback rate {(warm > 0).mean():.3f}")
    assert warm.mean() < 0.9 * cold.mean()
    assert warm.mean() < rnd.mean()
    sample = np.linspace(0, n - 1, 24).astype(int)
    err = worst_energy_error(data, bases, ref, sample)
    say(f"  energy-norm error vs FP64 (24 cells): {err:.2e}")
    assert err <= 1e-6

    # the gate accepts a guess that meets the tolerance ...
    x0.upload(ref)
    its, got = run(o32, x0.ptr)
    assert (its == 0).all()
    assert np.array_equal(got, ref)
    # ... and falls back to CG for a corrupted one
    bad = ref.copy().reshape(n, 4, data.nf + 1, data.nf + 1)
    bad[::7, :, 10:20, 5:9] = np.nan
    bad[1::7, 2, 30, 30] = 1e3
    x0.upload(bad.reshape(-1))
    its, got = run(o32, x0.ptr)
    hit = np.zeros(n, bool)
    hit[::7] = True
    hit[1::7] = True
    assert (its.reshape(n, 4)[hit].max(axis=1) > 0).all()
    assert worst_energy_error(data, got, ref, np.where(hit)[0][:16]) <= 1e-6
    data.free()
    x0.free()


def test_pipelined_serving_is_bitwise(ctx):
    """serve_pipelined (inference of chunk k+1 on the context's second stream
    beside the builds of chunk k) gives bitwise the bases, A^K, F^K and
    iteration counts of predict_bases + one build of all cells with the same
    inference batch, over two steps in a row (the second step's inference waits
    for the first step's builds)."""
    net = S.Surrogate.load(ctx, WEIGHTS)
    data = S.TileSet(ctx, 640, 43)
    n, per = data.n, 4 * (data.nf + 1) ** 2
    o32 = msfem.default_opts(precision=msfem.PREC_FP32_REFINE, rel_tol=1e-8)
    x0 = ctx.alloc(n * per, np.float64)
    outs = (data.bases, data.Ae, data.Fe, data.iters)
    for batch, chunks in ((512, 4), (128, 3)):
        say(f"pipelined serving: 640 GRF cells, inference batch {batch}, {chunks} chunks")
        net.predict_bases(data.x, n, x0.ptr, batch=batch)
        data.solve(o32, x0.ptr)
        ctx.synchronize()
        want = [a.numpy().reshape(-1).copy() for a in outs]
        pred = x0.numpy().reshape(-1).copy()
        for steps in (1, 2):  # one step; two steps back to back
            for a in (data.bases, x0):
                a.upload(np.zeros(n * per))
            for _ in range(steps):
                S.serve_pipelined(net, data, x0, o32, chunks=chunks, batch=batch)
            ctx.synchronize()
            assert not data.status.numpy().any()
            assert np.array_equal(x0.numpy().reshape(-1), pred)
            for w, a in zip(want, outs):
                assert np.array_equal(a.numpy().reshape(-1), w)
    data.free()
    x0.free()
