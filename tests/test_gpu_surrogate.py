"""GPU checks of the surrogate's HIP kernels against numpy / the FP64 network oracle.

Bars: BF16-rounding, pooling, upsampling, im2col/col2im and the delta<->bases
maps are bit-exact (they are exact in FP32); the MFMA GEMM is within 1e-5
relative (FP32 accumulation of exact BF16 products, vs FP64 sums); a tiny
network's output, loss and every layer's weight/bias gradient within 1e-3
relative Frobenius norm of the oracle (measured: about 1e-7; the margin covers
activations that land one BF16 step apart after re-rounding on either side).
"""
import sys

import numpy as np
import pytest

import msfem
import surrogate as S
import surrogate_ref as R

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = msfem.Context(0)
    yield c
    c.close()


def bf16_bits(a):
    """float64 values already on the BF16 grid -> raw uint16."""
    return (np.asarray(a, np.float32).view(np.uint32) >> 16).astype(np.uint16)


def frob(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30)


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(70, 100, 45), (64, 64, 32), (5, 300, 600), (130, 7, 9)])
def test_gemm_bf16(ctx, ta, tb, M, N, K):
    say(f"gemm ta={ta} tb={tb} M={M} N={N} K={K}")
    rng = np.random.default_rng(M * 7 + N + K + 13 * ta + 29 * tb)
    A = R.bf16(rng.standard_normal((M, K)))
    B = R.bf16(rng.standard_normal((K, N)))
    C0 = rng.standard_normal((M, N)).astype(np.float32)
    a_store = A.T if ta else A
    b_store = B.T if tb else B
    dA, dB = ctx.to_device(bf16_bits(np.ascontiguousarray(a_store))), ctx.to_device(bf16_bits(np.ascontiguousarray(b_store)))
    for acc in (0, 1):
        dC = ctx.to_device(C0)
        rc = S._nn().msfem_nn_gemm_bf16(ta, tb, M, N, K, dA.ptr, a_store.shape[1], dB.ptr, b_store.shape[1], dC.ptr, N,
                                        acc, ctx.stream)
        assert rc == 0
        ctx.synchronize()
        want = A @ B + (C0 if acc else 0.0)
        got = dC.numpy().reshape(-1).reshape(M, N)
        assert np.abs(got - want).max() <= 1e-5 * max(np.abs(want).max(), 1.0), (acc, np.abs(got - want).max())


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K,splits", [(32, 288, 4100, 4), (256, 2304, 700, 3), (70, 100, 45, 2)])
def test_gemm_bf16_splitk(ctx, ta, tb, M, N, K, splits):
    """Split-K form (fixed-order sum of per-slice partials) against numpy."""
    say(f"gemm split-K ta={ta} tb={tb} M={M} N={N} K={K} splits={splits}")
    rng = np.random.default_rng(M + N + K + splits + 3 * ta + 5 * tb)
    A = R.bf16(rng.standard_normal((M, K)))
    B = R.bf16(rng.standard_normal((K, N)))
    C0 = rng.standard_normal((M, N)).astype(np.float32)
    a_store = A.T if ta else A
    b_store = B.T if tb else B
    dA, dB = ctx.to_device(bf16_bits(np.ascontiguousarray(a_store))), ctx.to_device(bf16_bits(np.ascontiguousarray(b_store)))
    work = ctx.alloc(splits * M * N, np.float32)
    for acc in (0, 1):
        dC = ctx.to_device(C0)
        rc = S._nn().msfem_nn_gemm_bf16_splitk(ta, tb, M, N, K, dA.ptr, a_store.shape[1], dB.ptr, b_store.shape[1],
                                               dC.ptr, N, acc, splits, work.ptr, ctx.stream)
        assert rc == 0
        ctx.synchronize()
        want = A @ B + (C0 if acc else 0.0)
        got = dC.numpy().reshape(M, N)
        assert np.abs(got - want).max() <= 1e-5 * max(np.abs(want).max(), 1.0), (acc, np.abs(got - want).max())


def test_gemm_rejects_bad_sizes(ctx):
    d = ctx.alloc(16, np.float32)
    assert S._nn().msfem_nn_gemm_bf16(0, 0, 0, 4, 4, d.ptr, 4, d.ptr, 4, d.ptr, 4, 0, ctx.stream) == msfem.ERR_ARG
    assert S._nn().msfem_nn_gemm_bf16(0, 0, 4, 4, 4, None, 4, d.ptr, 4, d.ptr, 4, 0, ctx.stream) == msfem.ERR_ARG


def test_elementwise_primitives(ctx):
    say("nn elementwise primitives")
    L, st = S._nn(), ctx.stream
    rng = np.random.default_rng(5)
    C, B, H, W = 3, 2, 6, 8
    x = rng.standard_normal((C, B, H, W)).astype(np.float32)
    x.reshape(-1)[:4] = [1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, -0.0, 65504.0]
    dx = ctx.to_device(x)
    # BF16 rounding
    dbf = ctx.alloc(x.size, np.uint16)
    assert L.msfem_nn_to_bf16(x.size, dx.ptr, dbf.ptr, st) == 0
    assert np.array_equal(dbf.numpy().reshape(-1), bf16_bits(R.bf16(x)).reshape(-1))
    # im2col (3x3 pad 1, 2x2 pad 0) and col2im
    for k, pad in ((3, 1), (2, 0)):
        col, (Ho, Wo) = R.im2col(x.astype(np.float64), k, pad)
        dcol = ctx.alloc(col.size, np.uint16)
        assert L.msfem_nn_im2col(C, B, H, W, k, k, pad, dx.ptr, dcol.ptr, st) == 0
        assert np.array_equal(dcol.numpy().reshape(-1), bf16_bits(R.bf16(col)).reshape(-1))
        g = rng.standard_normal(col.shape).astype(np.float32)
        dg, dout = ctx.to_device(g), ctx.alloc(x.size, np.float32)
        assert L.msfem_nn_col2im(C, B, H, W, k, k, pad, dg.ptr, dout.ptr, st) == 0
        assert np.allclose(dout.numpy().reshape(-1), R.col2im(g.astype(np.float64), x.shape, k, pad).reshape(-1), rtol=1e-6, atol=1e-5)
    # pool / up, forward and backward
    out = ctx.alloc(x.size, np.float32)
    assert L.msfem_nn_pool2(C * B, H, W, dx.ptr, out.ptr, 0, st) == 0
    assert np.allclose(out.numpy().reshape(-1)[:x.size // 4], R.pool2(x.astype(np.float64)).reshape(-1), rtol=1e-6, atol=1e-6)
    small = rng.standard_normal((C, B, H // 2, W // 2)).astype(np.float32)
    ds = ctx.to_device(small)
    assert L.msfem_nn_pool2(C * B, H, W, ds.ptr, out.ptr, 1, st) == 0
    assert np.array_equal(out.numpy().reshape(-1), R.pool2_bwd(small.astype(np.float64)).astype(np.float32).reshape(-1))
    assert L.msfem_nn_up2(C * B, H // 2, W // 2, ds.ptr, out.ptr, 0, st) == 0
    assert np.array_equal(out.numpy().reshape(-1), R.up2(small).reshape(-1))
    assert L.msfem_nn_up2(C * B, H // 2, W // 2, dx.ptr, out.ptr, 1, st) == 0
    assert np.allclose(out.numpy().reshape(-1)[:x.size // 4], R.up2_bwd(x.astype(np.float64)).reshape(-1), rtol=1e-6, atol=1e-6)
    # bias + ReLU, masked backward, bias gradient
    N = B * H * W
    bias = rng.standard_normal(C).astype(np.float32)
    dbias, y = ctx.to_device(bias), ctx.to_device(x)
    assert L.msfem_nn_bias_act(C, N, dbias.ptr, y.ptr, 1, st) == 0
    yref = np.maximum(x.reshape(C, N) + bias[:, None], 0)
    assert np.array_equal(y.numpy().reshape(-1).reshape(C, N), yref)
    d = rng.standard_normal((C, N)).astype(np.float32)
    dd, dbf2 = ctx.to_device(d), ctx.alloc(C * N, np.uint16)
    assert L.msfem_nn_act_bwd(C * N, y.ptr, dd.ptr, dbf2.ptr, 1, st) == 0
    dm = np.where(yref > 0, d, 0).astype(np.float32)
    assert np.array_equal(dd.numpy().reshape(-1).reshape(C, N), dm)
    assert np.array_equal(dbf2.numpy().reshape(-1), bf16_bits(R.bf16(dm)).reshape(-1))
    gb = ctx.alloc(C, np.float32)
    assert L.msfem_nn_bias_grad(C, N, dd.ptr, gb.ptr, st) == 0
    assert np.allclose(gb.numpy().reshape(-1), dm.astype(np.float64).sum(axis=1), rtol=1e-5, atol=1e-5)
    # axpy / axpby
    a2 = ctx.to_device(x)
    assert L.msfem_nn_axpby(x.size, 2.0, dx.ptr, -1.0, dx.ptr, a2.ptr, st) == 0
    assert np.allclose(a2.numpy().reshape(-1), x.reshape(-1))
    assert L.msfem_nn_axpy(x.size, 0.5, dx.ptr, a2.ptr, st) == 0
    assert np.allclose(a2.numpy().reshape(-1), 1.5 * x.reshape(-1))
    assert L.msfem_nn_axpby(x.size, 1.0, dx.ptr, 0.0, None, a2.ptr, st) == 0
    assert np.array_equal(a2.numpy().reshape(-1), x.reshape(-1))


@pytest.mark.parametrize("C,B,H,W,k,pad", [(2, 3, 4, 4, 3, 1), (3, 2, 16, 16, 3, 1), (2, 8, 2, 2, 3, 1),
                                           (4, 2, 8, 8, 2, 0), (1, 5, 6, 10, 3, 1)])
def test_im2col_geometries(ctx, C, B, H, W, k, pad):
    """im2col bit-exact against numpy for the wide (8 columns per thread, shared
    or per-column rows) and the element-wise forms."""
    say(f"im2col C={C} B={B} H={H} W={W} k={k} pad={pad}")
    rng = np.random.default_rng(C * 100 + B * 10 + H + W + k)
    x = rng.standard_normal((C, B, H, W)).astype(np.float32)
    col, _ = R.im2col(x.astype(np.float64), k, pad)
    dx, dcol = ctx.to_device(x), ctx.alloc(col.size, np.uint16)
    assert S._nn().msfem_nn_im2col(C, B, H, W, k, k, pad, dx.ptr, dcol.ptr, ctx.stream) == 0
    assert np.array_equal(dcol.numpy().reshape(-1), bf16_bits(R.bf16(col)).reshape(-1))


@pytest.mark.parametrize("Cin,Cout,B,H,W,k,pad,splits", [
    (1, 32, 3, 16, 16, 3, 1, 1), (32, 32, 2, 64, 64, 3, 1, 1), (64, 128, 4, 16, 8, 3, 1, 1),
    (256, 256, 16, 8, 8, 3, 1, 4), (128, 64, 2, 32, 32, 3, 1, 3), (5, 7, 3, 9, 8, 3, 1, 1)])
def test_implicit_conv(ctx, Cin, Cout, B, H, W, k, pad, splits):
    """msfem_nn_conv_bf16 (column matrix gathered in the GEMM's tile loads) is
    bitwise equal to msfem_nn_im2col + the GEMM, split K included, and matches
    the oracle's convolution in FP64 within the BF16 operand rounding."""
    say(f"implicit conv Cin={Cin} Cout={Cout} B={B} H={H} W={W} splits={splits}")
    L, st = S._nn(), ctx.stream
    rng = np.random.default_rng(Cin + Cout + B + H)
    x = rng.standard_normal((Cin, B, H, W)).astype(np.float32)
    K = Cin * k * k
    w = (rng.standard_normal((Cout, K)) / np.sqrt(K)).astype(np.float32)
    Ho, Wo = H + 2 * pad - k + 1, W + 2 * pad - k + 1
    N = B * Ho * Wo
    dx, dw = ctx.to_device(x), ctx.to_device(bf16_bits(R.bf16(w)))
    y1, y2 = ctx.alloc(Cout * N, np.float32), ctx.alloc(Cout * N, np.float32)
    work = ctx.alloc(max(1, splits) * Cout * N, np.float32)
    wp = work.ptr if splits > 1 else None
    bias = rng.standard_normal(Cout).astype(np.float32)
    db = ctx.to_device(bias)
    for relu in (1, 0):  # fused bias + activation against the separate kernel
        assert L.msfem_nn_conv_bf16(Cin, B, H, W, k, k, pad, dw.ptr, Cout, dx.ptr, y1.ptr, splits, wp, db.ptr, relu,
                                    st) == 0
        col = ctx.alloc(K * N, np.uint16)
        assert L.msfem_nn_im2col(Cin, B, H, W, k, k, pad, dx.ptr, col.ptr, st) == 0
        assert L.msfem_nn_gemm_bf16_splitk(0, 0, Cout, N, K, dw.ptr, K, col.ptr, N, y2.ptr, N, 0, splits, wp, st) == 0
        assert L.msfem_nn_bias_act(Cout, N, db.ptr, y2.ptr, relu, st) == 0
        a, b = y1.numpy().reshape(-1), y2.numpy().reshape(-1)
        assert np.array_equal(a, b)
    # without bias: the plain product, against the oracle's convolution in FP64
    assert L.msfem_nn_conv_bf16(Cin, B, H, W, k, k, pad, dw.ptr, Cout, dx.ptr, y1.ptr, splits, wp, None, 0, st) == 0
    a = y1.numpy().reshape(-1)
    cref, _ = R.im2col(x.astype(np.float64), k, pad)
    ref = R.bf16(w).astype(np.float64) @ R.bf16(cref)
    assert np.abs(a - ref.reshape(-1)).max() < 1e-4 * np.abs(ref).max()
    # a width that is not a multiple of 8 is refused (the caller uses the explicit path)
    assert L.msfem_nn_conv_bf16(Cin, B, H, 6, k, k, pad, dw.ptr, Cout, dx.ptr, y1.ptr, 1, None, None, 0,
                                st) == msfem.ERR_ARG


@pytest.mark.parametrize("Cin,Cout,B,H,W", [
    (32, 32, 2, 64, 64), (64, 32, 1, 64, 64), (32, 64, 3, 32, 32), (128, 64, 1, 32, 32), (64, 128, 2, 16, 16),
    (256, 128, 1, 16, 16), (32, 96, 2, 8, 16), (64, 64, 1, 6, 64), (128, 256, 4, 8, 8), (256, 256, 2, 8, 8),
    (32, 32, 6, 8, 8), (256, 256, 16, 4, 4), (64, 32, 8, 4, 4)])
def test_conv3x3_direct(ctx, Cin, Cout, B, H, W):
    """msfem_nn_conv3x3_bf16 (input rows staged once per channel block, 9 taps
    from LDS) against the oracle's convolution in FP64 on the same BF16
    operands: every output within the recursive FP32 summation bound
    K 2^-24 sum|w x| (K = 9 Cin terms), bias + ReLU applied as bias_act; and
    against the implicit GEMM within twice that bound."""
    say(f"direct conv3x3 Cin={Cin} Cout={Cout} B={B} H={H} W={W}")
    L, st = S._nn(), ctx.stream
    rng = np.random.default_rng(Cin + 3 * Cout + B + H + W)
    x = rng.standard_normal((Cin, B, H, W)).astype(np.float32)
    K = Cin * 9
    w = (rng.standard_normal((Cout, K)) / np.sqrt(K)).astype(np.float32)
    bias = rng.standard_normal(Cout).astype(np.float32)
    N = B * H * W
    dx, dw, db = ctx.to_device(x), ctx.to_device(bf16_bits(R.bf16(w))), ctx.to_device(bias)
    dwp = ctx.alloc(Cout * K, np.uint16)
    y1, y2 = ctx.alloc(Cout * N, np.float32), ctx.alloc(Cout * N, np.float32)
    assert L.msfem_nn_conv3x3_pack(Cin, Cout, dw.ptr, dwp.ptr, st) == 0
    cref, _ = R.im2col(R.bf16(x.astype(np.float64)), 3, 1)
    wb = R.bf16(w).astype(np.float64)
    ref, aref = wb @ cref, np.abs(wb) @ np.abs(cref)
    bound = K * 2.0 ** -24 * aref * 1.0001 + 1e-30
    for relu in (0, 1):
        assert L.msfem_nn_conv3x3_bf16(Cin, B, H, W, dwp.ptr, Cout, dx.ptr, 0, y1.ptr, db.ptr, relu, None, st) == 0
        if W % 8 == 0:  # the implicit GEMM
            assert L.msfem_nn_conv_bf16(Cin, B, H, W, 3, 3, 1, dw.ptr, Cout, dx.ptr, y2.ptr, 1, None, db.ptr, relu,
                                        st) == 0
        else:  # the explicit form
            col = ctx.alloc(K * N, np.uint16)
            assert L.msfem_nn_im2col(Cin, B, H, W, 3, 3, 1, dx.ptr, col.ptr, st) == 0
            assert L.msfem_nn_gemm_bf16(0, 0, Cout, N, K, dw.ptr, K, col.ptr, N, y2.ptr, N, 0, st) == 0
            assert L.msfem_nn_bias_act(Cout, N, db.ptr, y2.ptr, relu, st) == 0
            col.free()
        a, b = y1.numpy().reshape(Cout, N).astype(np.float64), y2.numpy().reshape(Cout, N).astype(np.float64)
        r = ref + bias[:, None]
        if relu:
            r = np.maximum(r, 0.0)  # ReLU is 1-Lipschitz: the bound carries over
        assert (np.abs(a - r) <= bound).all(), np.abs(a - r).max()
        assert (np.abs(a - b) <= 2 * bound).all()
    # without bias: the plain sums
    assert L.msfem_nn_conv3x3_bf16(Cin, B, H, W, dwp.ptr, Cout, dx.ptr, 0, y1.ptr, None, 0, None, st) == 0
    assert (np.abs(y1.numpy().reshape(Cout, N) - ref) <= bound).all()
    # decoder form: nearest upsampling of a half-resolution input in the staging
    # and the skip sum in the epilogue, bitwise what up2 + conv + axpby give
    xs = ctx.to_device(rng.standard_normal((Cin, B, H // 2, W // 2)).astype(np.float32))
    skip = ctx.to_device(rng.standard_normal((Cout, N)).astype(np.float32))
    assert L.msfem_nn_up2(Cin * B, H // 2, W // 2, xs.ptr, dx.ptr, 0, st) == 0
    assert L.msfem_nn_conv3x3_bf16(Cin, B, H, W, dwp.ptr, Cout, dx.ptr, 0, y2.ptr, db.ptr, 1, None, st) == 0
    assert L.msfem_nn_axpby(Cout * N, 1.0, y2.ptr, 1.0, skip.ptr, y2.ptr, st) == 0
    assert L.msfem_nn_conv3x3_bf16(Cin, B, H, W, dwp.ptr, Cout, xs.ptr, 1, y1.ptr, db.ptr, 1, skip.ptr, st) == 0
    assert np.array_equal(y1.numpy().reshape(-1), y2.numpy().reshape(-1))
    # shapes outside the kernel's tiling are refused (the caller takes the implicit GEMM)
    bad = [(Cin, B, H + 1, W), (Cin + 16, B, H, W), (Cin, B, H, 2)]
    bad += [(Cin, B + 1, W, W), (Cin, B, 2 * W, W)] if W <= 8 else [(Cin, 2, 4, 8), (Cin, 8, 2, 4)]
    for c, bb, h, ww in bad:  # every shape here is refused before any launch
        assert L.msfem_nn_conv3x3_bf16(c, bb, h, ww, dwp.ptr, Cout, dx.ptr, 0, y1.ptr, None, 0, None, st) == msfem.ERR_ARG
    assert L.msfem_nn_conv3x3_bf16(Cin, B, H, W, dwp.ptr, Cout + 16, dx.ptr, 0, y1.ptr, None, 0, None, st) == msfem.ERR_ARG
    for a in (dx, dw, db, dwp, y1, y2):
        a.free()


@pytest.mark.parametrize("Cin,Cout,B,H,W", [
    (32, 32, 2, 64, 64), (32, 32, 64, 64, 64), (32, 64, 1, 32, 32), (64, 128, 2, 16, 16), (128, 256, 4, 8, 8),
    (256, 256, 16, 4, 4), (32, 32, 3, 32, 32), (64, 32, 2, 64, 64)])
def test_conv3x3_fused_pool(ctx, Cin, Cout, B, H, W):
    """msfem_nn_conv3x3_bf16_ex with pooling: the convolution's output is bitwise that of
    msfem_nn_conv3x3_bf16 and the pooled output bitwise msfem_nn_pool2 of it --
    the persistent kernel (32 -> 32 at 64^2, one tile and many tiles per
    workgroup) and the tile kernel (BM 32 and 64, whole rows and whole images
    per tile)."""
    say(f"fused pool conv3x3 Cin={Cin} Cout={Cout} B={B} H={H} W={W}")
    L, st = S._nn(), ctx.stream
    rng = np.random.default_rng(7 * Cin + Cout + B + H + W)
    K, N = Cin * 9, B * H * W
    x = ctx.to_device(rng.standard_normal((Cin, B, H, W)).astype(np.float32))
    dw = ctx.to_device(bf16_bits(R.bf16((rng.standard_normal((Cout, K)) / np.sqrt(K)).astype(np.float32))))
    db = ctx.to_device(rng.standard_normal(Cout).astype(np.float32))
    wpk = ctx.alloc(Cout * K, np.uint16)
    y1, y2 = ctx.alloc(Cout * N, np.float32), ctx.alloc(Cout * N, np.float32)
    p1, p2 = ctx.alloc(Cout * N // 4, np.float32), ctx.alloc(Cout * N // 4, np.float32)
    assert L.msfem_nn_conv3x3_pack(Cin, Cout, dw.ptr, wpk.ptr, st) == 0
    assert L.msfem_nn_conv3x3_bf16(Cin, B, H, W, wpk.ptr, Cout, x.ptr, 0, y1.ptr, db.ptr, 1, None, st) == 0
    assert L.msfem_nn_pool2(Cout * B, H, W, y1.ptr, p1.ptr, 0, st) == 0
    assert L.msfem_nn_conv3x3_bf16_ex(Cin, B, H, W, wpk.ptr, Cout, x.ptr, 0, y2.ptr, db.ptr, 1, None, p2.ptr, 0,
                                      st) == 0
    assert np.array_equal(y1.numpy(), y2.numpy())
    assert np.array_equal(p1.numpy(), p2.numpy()) and np.abs(p1.numpy()).max() > 0
    for a in (x, dw, db, wpk, y1, y2, p1, p2):
        a.free()


@pytest.mark.parametrize("B,H,pooled", [(1, 64, 1), (2, 64, 0), (64, 64, 1), (3, 16, 1)])
def test_conv3x3_first_fused(ctx, B, H, pooled):
    """msfem_nn_conv3x3_first_bf16 (a 1 -> 32 3x3 layer formed in the staging of
    the 32 -> 32 one) is bitwise msfem_nn_conv_bf16 followed by
    msfem_nn_conv3x3_bf16_ex, with one tile and many tiles per workgroup and
    short images (the first layer's zero rows at both image edges)."""
    say(f"fused first layer B={B} H={H} pool={pooled}")
    L, st = S._nn(), ctx.stream
    rng = np.random.default_rng(11 * B + H)
    W, N = 64, B * H * 64
    x = ctx.to_device(rng.standard_normal((B, H, W)).astype(np.float32))
    w0 = ctx.to_device(bf16_bits(R.bf16((rng.standard_normal((32, 9)) / 3).astype(np.float32))))
    b0 = ctx.to_device((0.3 * rng.standard_normal(32)).astype(np.float32))
    dw = ctx.to_device(bf16_bits(R.bf16((rng.standard_normal((32, 288)) / np.sqrt(288)).astype(np.float32))))
    db = ctx.to_device(rng.standard_normal(32).astype(np.float32))
    wpk = ctx.alloc(32 * 288, np.uint16)
    h = ctx.alloc(32 * N, np.float32)
    y1, y2 = ctx.alloc(32 * N, np.float32), ctx.alloc(32 * N, np.float32)
    p1, p2 = ctx.alloc(32 * N // 4, np.float32), ctx.alloc(32 * N // 4, np.float32)
    assert L.msfem_nn_conv3x3_pack(32, 32, dw.ptr, wpk.ptr, st) == 0
    assert L.msfem_nn_conv_bf16(1, B, H, W, 3, 3, 1, w0.ptr, 32, x.ptr, h.ptr, 1, None, b0.ptr, 1, st) == 0
    assert L.msfem_nn_conv3x3_bf16_ex(32, B, H, W, wpk.ptr, 32, h.ptr, 0, y1.ptr, db.ptr, 1, None,
                                      p1.ptr if pooled else None, 0, st) == 0
    assert L.msfem_nn_conv3x3_first_bf16(B, H, W, w0.ptr, b0.ptr, 1, x.ptr, wpk.ptr, 32, y2.ptr, db.ptr, 1,
                                         p2.ptr if pooled else None, st) == 0
    ref = y1.numpy()
    assert np.array_equal(ref, y2.numpy()) and (ref > 0).any() and (ref == 0).any()
    if pooled:
        assert np.array_equal(p1.numpy(), p2.numpy())
    assert L.msfem_nn_conv3x3_first_bf16(B, H, 32, w0.ptr, b0.ptr, 1, x.ptr, wpk.ptr, 32, y2.ptr, db.ptr, 1, None,
                                         st) == msfem.ERR_ARG
    for a in (x, w0, b0, dw, db, wpk, h, y1, y2, p1, p2):
        a.free()


@pytest.mark.parametrize("Cin,Cout,B,H,W,up", [
    (64, 32, 2, 64, 64, 1), (32, 32, 3, 64, 64, 0), (128, 64, 2, 32, 32, 1), (64, 64, 2, 32, 32, 0),
    (256, 128, 4, 16, 16, 1), (256, 256, 8, 8, 8, 1), (256, 256, 16, 4, 4, 1), (32, 64, 16, 4, 4, 0)])
def test_conv3x3_bf16_storage(ctx, Cin, Cout, B, H, W, up):
    """msfem_nn_conv3x3_bf16_ex's BF16 activation storage: with io bit 1 the
    output is bitwise the BF16 rounding of the FP32 output, and with io bit 0
    the BF16 bits of the input give bitwise the output of the FP32 input, with
    and without upsampling, skip and pooling, BM 32 and 64 and every tile shape."""
    say(f"bf16 storage conv3x3 Cin={Cin} Cout={Cout} B={B} H={H} W={W} up={up}")
    L, st = S._nn(), ctx.stream
    rng = np.random.default_rng(3 * Cin + Cout + B + H + up)
    K, N = Cin * 9, B * H * W
    Ni = N // 4 if up else N
    xf = rng.standard_normal(Cin * Ni).astype(np.float32)
    x, xb = ctx.to_device(xf), ctx.to_device(bf16_bits(R.bf16(xf)))
    dw = ctx.to_device(bf16_bits(R.bf16((rng.standard_normal((Cout, K)) / np.sqrt(K)).astype(np.float32))))
    db = ctx.to_device(rng.standard_normal(Cout).astype(np.float32))
    sk = ctx.to_device(rng.standard_normal(Cout * N).astype(np.float32))
    wpk = ctx.alloc(Cout * K, np.uint16)
    assert L.msfem_nn_conv3x3_pack(Cin, Cout, dw.ptr, wpk.ptr, st) == 0
    y0, y1 = ctx.alloc(Cout * N, np.float32), ctx.alloc(Cout * N, np.float32)
    h2, h3 = ctx.alloc(Cout * N, np.uint16), ctx.alloc(Cout * N, np.uint16)
    p0, p1 = ctx.alloc(Cout * N // 4, np.float32), ctx.alloc(Cout * N // 4, np.float32)
    for add in (None, sk.ptr):
        pool = (p0.ptr, p1.ptr) if add is None and not up else (None, None)
        assert L.msfem_nn_conv3x3_bf16_ex(Cin, B, H, W, wpk.ptr, Cout, x.ptr, up, y0.ptr, db.ptr, 1, add, pool[0], 0,
                                          st) == 0
        assert L.msfem_nn_conv3x3_bf16_ex(Cin, B, H, W, wpk.ptr, Cout, xb.ptr, up, y1.ptr, db.ptr, 1, add, pool[1], 1,
                                          st) == 0
        assert L.msfem_nn_conv3x3_bf16_ex(Cin, B, H, W, wpk.ptr, Cout, x.ptr, up, h2.ptr, db.ptr, 1, add, None, 2,
                                          st) == 0
        assert L.msfem_nn_conv3x3_bf16_ex(Cin, B, H, W, wpk.ptr, Cout, xb.ptr, up, h3.ptr, db.ptr, 1, add, None, 3,
                                          st) == 0
        ref = y0.numpy()
        assert np.abs(ref).max() > 0
        assert np.array_equal(ref, y1.numpy())
        assert np.array_equal(bf16_bits(R.bf16(ref)), h2.numpy()) and np.array_equal(h2.numpy(), h3.numpy())
        if pool[0] is not None:
            assert np.array_equal(p0.numpy(), p1.numpy())
    assert L.msfem_nn_conv3x3_bf16_ex(Cin, B, H, W, wpk.ptr, Cout, xb.ptr, up, y1.ptr, db.ptr, 1, None, None, 4,
                                      st) == msfem.ERR_ARG
    for a in (x, xb, dw, db, sk, wpk, y0, y1, h2, h3, p0, p1):
        a.free()


@pytest.mark.parametrize("Cin,B,nf", [(32, 3, 64), (8, 2, 16), (13, 1, 5)])
def test_head_bases(ctx, Cin, B, nf):
    """msfem_nn_head_bases is bitwise msfem_nn_conv_small (2x2 valid, 4
    channels) + msfem_nn_axpby(scale, h, 0) + msfem_nn_delta_to_bases, from FP32
    and from BF16 activations, with and without ReLU."""
    say(f"head as bases Cin={Cin} B={B} nf={nf}")
    L, st = S._nn(), ctx.stream
    rng = np.random.default_rng(Cin + B + nf)
    K, ni, np1 = Cin * 4, nf - 1, nf + 1
    xf = rng.standard_normal(Cin * B * nf * nf).astype(np.float32)
    x, xb = ctx.to_device(xf), ctx.to_device(bf16_bits(R.bf16(xf)))
    dw = ctx.to_device(bf16_bits(R.bf16((rng.standard_normal((4, K)) / np.sqrt(K)).astype(np.float32))))
    db = ctx.to_device(rng.standard_normal(4).astype(np.float32))
    h = ctx.alloc(4 * B * ni * ni, np.float32)
    ref, got = ctx.alloc(B * 4 * np1 * np1, np.float64), ctx.alloc(B * 4 * np1 * np1, np.float64)
    scale = np.float32(0.731)
    for relu in (0, 1):
        assert L.msfem_nn_conv_small(Cin, B, nf, nf, 2, 2, 0, dw.ptr, 4, x.ptr, h.ptr, db.ptr, relu, st) == 0
        assert L.msfem_nn_axpby(4 * B * ni * ni, scale, h.ptr, 0.0, None, h.ptr, st) == 0
        assert L.msfem_nn_delta_to_bases(B, nf, h.ptr, ref.ptr, st) == 0
        want = ref.numpy()
        for src, ib in ((x, 0), (xb, 1)):
            assert L.msfem_nn_head_bases(Cin, B, nf, dw.ptr, src.ptr, ib, db.ptr, relu, scale, got.ptr, st) == 0
            assert np.array_equal(got.numpy(), want)
    assert L.msfem_nn_head_bases(Cin, B, 1, dw.ptr, x.ptr, 0, db.ptr, 0, scale, got.ptr, st) == msfem.ERR_ARG
    for a in (x, xb, dw, db, h, ref, got):
        a.free()


def test_serving_forward_matches_training_form(ctx):
    """The serving forward of the full-size network (first two layers fused,
    pooled encoder outputs, decoder layers with upsampling and skip inside the
    convolution and BF16 activations, head written as bases) gives bitwise the
    bases of the training form's forward + msfem_nn_axpby +
    msfem_nn_delta_to_bases, and without bases bitwise the training form's
    head output.  Random weights and biases; 8 tiles (every layer takes its
    serving kernel)."""
    say("serving forward vs training form, 8 tiles of 64^2")
    net = S.Surrogate(ctx, seed=5)
    rng = np.random.default_rng(5)
    for layer in net.layers:
        w, b = layer.get()
        layer.set(w, 0.1 * rng.standard_normal(b.shape))
    net.dscale = 0.37
    L, st = S._nn(), ctx.stream
    B, t = 8, net.tile
    m, nb = 4 * B * (t - 1) ** 2, B * 4 * (t + 1) ** 2
    x = ctx.to_device((0.5 * rng.standard_normal(B * t * t)).astype(np.float32))
    out = net.forward(x, B, keep=True)
    d, ref, got = ctx.alloc(m, np.float32), ctx.alloc(nb, np.float64), ctx.alloc(nb, np.float64)
    assert L.msfem_nn_axpby(m, net.dscale, out.ptr, 0.0, None, d.ptr, st) == 0
    assert L.msfem_nn_delta_to_bases(B, t, d.ptr, ref.ptr, st) == 0
    raw = out.read(0, m)
    assert np.abs(raw).max() > 0
    assert net.forward(x, B, keep=False, bases=got.ptr) is None
    assert np.array_equal(got.numpy(), ref.numpy())
    assert np.array_equal(net.forward(x, B, keep=False).read(0, m), raw)
    for a in (x, d, ref, got):
        a.free()


@pytest.mark.parametrize("Cin,B,H,W,up", [
    (32, 2, 64, 64, 0), (32, 1, 64, 64, 1), (64, 3, 32, 32, 0), (32, 2, 32, 16, 1), (32, 6, 8, 8, 1),
    (32, 8, 4, 4, 0), (64, 1, 64, 64, 1)])
def test_conv3x3_persistent_matches_tiles(ctx, Cin, B, H, W, up):
    """32 output and 32 input channels take the persistent kernel (weights
    staged once per workgroup, the next tile's loads in flight); its sums run
    in the tile kernel's order, so a 64-channel convolution with weights
    [w; w] (the tile kernel) gives bitwise the same first 32 channels, bias,
    ReLU, upsampling and skip sum included.  (Cin = 64: both sides are tile
    kernels, BM = 32 and 64.)"""
    say(f"persistent conv3x3 Cin={Cin} B={B} H={H} W={W} up={up}")
    L, st = S._nn(), ctx.stream
    rng = np.random.default_rng(Cin + B + H + W + up)
    hi, wi = (H // 2, W // 2) if up else (H, W)
    x = ctx.to_device(rng.standard_normal((Cin, B, hi, wi)).astype(np.float32))
    K, N = Cin * 9, B * H * W
    w = bf16_bits(R.bf16((rng.standard_normal((32, K)) / np.sqrt(K)).astype(np.float32)))
    bias = rng.standard_normal(32).astype(np.float32)
    skip = rng.standard_normal((32, N)).astype(np.float32)
    d32 = [ctx.to_device(a) for a in (w, bias, skip)]
    d64 = [ctx.to_device(np.concatenate([a, a])) for a in (w, bias, skip)]
    outs = []
    for (dw, db, dsk), cout in ((d32, 32), (d64, 64)):
        wpk, y = ctx.alloc(cout * K, np.uint16), ctx.alloc(cout * N, np.float32)
        assert L.msfem_nn_conv3x3_pack(Cin, cout, dw.ptr, wpk.ptr, st) == 0
        assert L.msfem_nn_conv3x3_bf16(Cin, B, H, W, wpk.ptr, cout, x.ptr, up, y.ptr, db.ptr, 1, dsk.ptr, st) == 0
        outs.append(y.numpy().reshape(cout, N)[:32].copy())
        # without bias and skip: the plain sums agree as well
        assert L.msfem_nn_conv3x3_bf16(Cin, B, H, W, wpk.ptr, cout, x.ptr, up, y.ptr, None, 0, None, st) == 0
        outs.append(y.numpy().reshape(cout, N)[:32].copy())
        wpk.free()
        y.free()
    assert np.array_equal(outs[0], outs[2]) and np.array_equal(outs[1], outs[3])
    assert np.abs(outs[1]).max() > 0
    for a in [x] + d32 + d64:
        a.free()


@pytest.mark.parametrize("Cin,Cout,B,H,W,k,pad", [
    (32, 4, 3, 64, 64, 2, 0), (5, 3, 2, 9, 7, 3, 1), (1, 8, 2, 16, 16, 3, 1), (64, 1, 1, 12, 10, 2, 0),
    (11, 6, 2, 8, 9, 1, 0), (3, 7, 2, 10, 12, 3, 1)])
def test_conv_small(ctx, Cin, Cout, B, H, W, k, pad):
    """msfem_nn_conv_small (VALU, few output channels) against the oracle's
    convolution in FP64 on the same BF16 operands within the recursive FP32
    summation bound K 2^-24 sum|w x|, and against im2col + GEMM + bias_act
    within twice that bound."""
    say(f"small conv Cin={Cin} Cout={Cout} B={B} H={H} W={W} k={k}")
    L, st = S._nn(), ctx.stream
    rng = np.random.default_rng(Cin + Cout + H + W + k)
    x = rng.standard_normal((Cin, B, H, W)).astype(np.float32)
    K = Cin * k * k
    w = (rng.standard_normal((Cout, K)) / np.sqrt(K)).astype(np.float32)
    bias = rng.standard_normal(Cout).astype(np.float32)
    dx, dw, db = ctx.to_device(x), ctx.to_device(bf16_bits(R.bf16(w))), ctx.to_device(bias)
    cref, (Ho, Wo) = R.im2col(R.bf16(x.astype(np.float64)), k, pad)
    N = B * Ho * Wo
    wb = R.bf16(w).astype(np.float64)
    ref, bound = wb @ cref, K * 2.0 ** -24 * (np.abs(wb) @ np.abs(cref)) * 1.0001 + 1e-30
    y1, y2, col = ctx.alloc(Cout * N, np.float32), ctx.alloc(Cout * N, np.float32), ctx.alloc(K * N, np.uint16)
    for relu in (0, 1):
        assert L.msfem_nn_conv_small(Cin, B, H, W, k, k, pad, dw.ptr, Cout, dx.ptr, y1.ptr, db.ptr, relu, st) == 0
        assert L.msfem_nn_im2col(Cin, B, H, W, k, k, pad, dx.ptr, col.ptr, st) == 0
        assert L.msfem_nn_gemm_bf16(0, 0, Cout, N, K, dw.ptr, K, col.ptr, N, y2.ptr, N, 0, st) == 0
        assert L.msfem_nn_bias_act(Cout, N, db.ptr, y2.ptr, relu, st) == 0
        a, b = y1.numpy().reshape(Cout, N).astype(np.float64), y2.numpy().reshape(Cout, N).astype(np.float64)
        r = ref + bias[:, None]
        if relu:
            r = np.maximum(r, 0.0)
        assert (np.abs(a - r) <= bound).all(), np.abs(a - r).max()
        assert (np.abs(a - b) <= 2 * bound).all()
    assert L.msfem_nn_conv_small(Cin, B, H, W, k, k, pad, dw.ptr, Cout, dx.ptr, y1.ptr, None, 0, st) == 0
    assert (np.abs(y1.numpy().reshape(Cout, N) - ref) <= bound).all()
    # more than 8 output channels, other kernel sizes, weights beyond the LDS budget: refused
    assert L.msfem_nn_conv_small(Cin, B, H, W, k, k, pad, dw.ptr, 9, dx.ptr, y1.ptr, None, 0, st) == msfem.ERR_ARG
    assert L.msfem_nn_conv_small(Cin, B, H, W, 4, 4, pad, dw.ptr, Cout, dx.ptr, y1.ptr, None, 0, st) == msfem.ERR_ARG
    assert L.msfem_nn_conv_small(4097, 1, 1, 1, 1, 1, 0, dw.ptr, 4, dx.ptr, y1.ptr, None, 0, st) == msfem.ERR_ARG
    for a in (dx, dw, db, y1, y2, col):
        a.free()


def test_mse_and_adam(ctx):
    say("nn mse + adam")
    L, st = S._nn(), ctx.stream
    rng = np.random.default_rng(6)
    n = 10007
    p, t = rng.standard_normal(n).astype(np.float32), rng.standard_normal(n).astype(np.float32)
    dp, dt, g, part = ctx.to_device(p), ctx.to_device(t), ctx.alloc(n, np.float32), ctx.alloc(64, np.float32)
    assert L.msfem_nn_mse(n, dp.ptr, dt.ptr, g.ptr, part.ptr, 64, st) == 0
    d = p.astype(np.float64) - t
    assert abs(part.numpy().reshape(-1).astype(np.float64).sum() / n - (d * d).mean()) < 1e-5 * (d * d).mean()
    assert np.allclose(g.numpy().reshape(-1), 2 * d / n, rtol=1e-5, atol=1e-12)
    # two Adam steps against the textbook update (bias-corrected, L2 decay added to the gradient)
    w = rng.standard_normal(n).astype(np.float32)
    dw, m, v, wbf = ctx.to_device(w), ctx.to_device(np.zeros(n, np.float32)), ctx.to_device(np.zeros(n, np.float32)), \
        ctx.alloc(n, np.uint16)
    lr, b1, b2, eps, wd = 1e-2, 0.9, 0.999, 1e-8, 1e-3
    wr, mr, vr = w.astype(np.float64), np.zeros(n), np.zeros(n)
    for step in (1, 2):
        gr = g.numpy().reshape(-1).astype(np.float64) + wd * dw.numpy().reshape(-1).astype(np.float64)
        assert L.msfem_nn_adam(n, dw.ptr, g.ptr, m.ptr, v.ptr, wbf.ptr, lr, b1, b2, eps, step, wd, st) == 0
        mr = b1 * mr + (1 - b1) * gr
        vr = b2 * vr + (1 - b2) * gr * gr
        mh, vh = mr / (1 - b1 ** step), vr / (1 - b2 ** step)
        wr = wr - lr * mh / (np.sqrt(vh) + eps)
    got = dw.numpy().reshape(-1)
    assert np.abs(got - wr).max() < 1e-5
    assert np.array_equal(wbf.numpy().reshape(-1), bf16_bits(R.bf16(got)))


def tiny_net(ctx, chans, tile, seed):
    net = S.Surrogate(ctx, tile=tile, chans=chans, seed=seed)
    rng = np.random.default_rng(seed + 100)
    for layer in net.layers:  # nonzero biases so every bias path is exercised
        w, _ = layer.get()
        layer.set(w, rng.standard_normal(layer.cout).astype(np.float32) * 0.1)
    return net


def test_tiny_net_forward_and_gradients(ctx):
    """Whole forward/backward of a 3-level net (16x16 tiles, 4/8/8 channels) vs the oracle."""
    say("tiny surrogate forward/backward vs oracle")
    chans, tile, B = (4, 8, 8), 16, 3
    net = tiny_net(ctx, chans, tile, seed=7)
    rng = np.random.default_rng(8)
    x = rng.standard_normal((1, B, tile, tile)).astype(np.float32)
    target = (rng.standard_normal((4, B, tile - 1, tile - 1)) * 0.1).astype(np.float32)
    dx, dtg = ctx.to_device(x), ctx.to_device(target)
    out = net.forward(dx, B)
    n = 4 * B * (tile - 1) ** 2
    got_out = out.numpy().reshape(-1)[:n].astype(np.float64)
    loss, grad = net.loss_and_grad(out, dtg, n)
    net.backward(grad)
    ctx.synchronize()
    params = {l.name: (l.get()[0].astype(np.float64), l.get()[1].astype(np.float64)) for l in net.layers}
    ref = R.Net(params, chans)
    ref_out = ref.forward(x.astype(np.float64))
    assert frob(got_out, ref_out.reshape(-1)) < 1e-3
    ref_loss, dref = R.loss_and_grad(ref_out, target.astype(np.float64))
    assert abs(loss - ref_loss) < 1e-3 * ref_loss
    ref.backward(dref)
    rg = ref.grads()
    for layer in net.layers:
        gw, gb = layer.gw.numpy().reshape(-1), layer.gb.numpy().reshape(-1)
        ew, eb = frob(gw, rg[layer.name][0].reshape(-1)), frob(gb, rg[layer.name][1])
        say(f"  {layer.name}: dW rel {ew:.2e}  db rel {eb:.2e}")
        assert ew < 1e-3 and eb < 1e-3, layer.name


def test_tiny_net_training_reduces_loss(ctx):
    """Adam on a fixed batch: 60 steps cut the loss at least 5x (all kernels in the loop)."""
    say("tiny surrogate training")
    chans, tile, B = (8, 16, 16), 16, 8
    net = tiny_net(ctx, chans, tile, seed=9)
    rng = np.random.default_rng(10)
    x = rng.standard_normal((1, B, tile, tile)).astype(np.float32)
    # a smooth target that depends on the input: local means of x, per output channel
    xs = x[0].astype(np.float64)
    t = np.stack([0.25 * (xs[:, :-1, :-1] + xs[:, 1:, :-1] + xs[:, :-1, 1:] + xs[:, 1:, 1:]) * (i + 1) * 0.1
                  for i in range(4)]).astype(np.float32)
    dx, dt = ctx.to_device(x), ctx.to_device(t)
    losses = [net.train_step(dx, dt, B, lr=3e-3) for _ in range(60)]
    say(f"  loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    assert np.isfinite(losses).all()
    assert losses[-1] < losses[0] / 5


def test_delta_bases_round_trip(ctx):
    """Solver bases -> delta (FP32) -> bases: boundary exact, interior within FP32 rounding."""
    say("delta <-> bases")
    nf, n = 16, 5
    g = msfem.grf_log(n, nf, 40, 4.0, S.CLIP)
    boxes = msfem.unit_boxes(2, 4)[:n]
    out = ctx.batch_build(2, [nf] * n, boxes, msfem.COEF_SCALAR_CELL, np.exp(g.astype(np.float64)).reshape(-1),
                          msfem.RHS_CONST, 0.0)
    bases = out["bases"]
    L, st = S._nn(), ctx.stream
    db = ctx.to_device(bases)
    dl = ctx.alloc(4 * n * (nf - 1) ** 2, np.float32)
    assert L.msfem_nn_bases_to_delta(n, nf, db.ptr, dl.ptr, st) == 0
    back = ctx.alloc(bases.size, np.float64)
    assert L.msfem_nn_delta_to_bases(n, nf, dl.ptr, back.ptr, st) == 0
    b2 = back.numpy().reshape(-1).reshape(n, 4, nf + 1, nf + 1)
    b1 = bases.reshape(n, 4, nf + 1, nf + 1)
    edge = np.ones((nf + 1, nf + 1), bool)
    edge[1:-1, 1:-1] = False
    assert np.array_equal(b2[:, :, edge], b1[:, :, edge])
    assert np.abs(b2 - b1).max() < 1e-7
