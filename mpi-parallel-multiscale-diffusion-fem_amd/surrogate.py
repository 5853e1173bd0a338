"""Learned basis surrogate: a convolutional encoder-decoder that maps a cell's
64x64 log-coefficient tile to the interior corrections delta_i = phi_i - N_i of
its four bases, trained on the GPU with the library's own BF16-MFMA kernels
(msfem_nn_* in include/msfem.h) and served as the initial guess x0 of the
batched basis build.  The build's FP64 residual test is the accept/fallback
gate: a prediction that already meets the tolerance costs no CG iteration, any
other is finished by the (warm-started) CG.  There is no counterpart in the
reference.

Network (6 levels, 64 -> 2 cells per side, 32 -> 256 channels):
  encoder  3x3 conv+ReLU: 1->32, 32->32 @64 | pool, 32->64 @32 | pool, 64->128 @16 |
           pool, 128->256 @8 | pool, 256->256 @4 | pool, 256->256 @2
  decoder  nearest x2, 3x3 conv+ReLU, + encoder skip: 256->256 @4, 256->256 @8,
           256->128 @16, 128->64 @32, 64->32 @64
  head     2x2 valid conv (cells -> the 63x63 interior nodes), 32->4, linear
Weights: FP32 masters + BF16 copies for the matrix cores; activations FP32.
"""
import ctypes
import os

import numpy as np

import msfem

CLIP = 0.5 * np.log(1e3)


def _nn():
    L = msfem.lib()
    if not hasattr(L, "_nn_declared"):
        vp, i32, u64, f32 = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_float
        L.msfem_nn_gemm_bf16.argtypes = [i32, i32, i32, i32, i32, vp, i32, vp, i32, vp, i32, i32, vp]
        L.msfem_nn_gemm_bf16_splitk.argtypes = [i32, i32, i32, i32, i32, vp, i32, vp, i32, vp, i32, i32, i32, vp, vp]
        L.msfem_nn_im2col.argtypes = [i32, i32, i32, i32, i32, i32, i32, vp, vp, vp]
        L.msfem_nn_conv_bf16.argtypes = [i32, i32, i32, i32, i32, i32, i32, vp, i32, vp, vp, i32, vp, vp, i32, vp]
        L.msfem_nn_conv3x3_pack.argtypes = [i32, i32, vp, vp, vp]
        L.msfem_nn_conv_small.argtypes = [i32, i32, i32, i32, i32, i32, i32, vp, i32, vp, vp, vp, i32, vp]
        L.msfem_nn_conv3x3_bf16.argtypes = [i32, i32, i32, i32, vp, i32, vp, i32, vp, vp, i32, vp, vp]
        L.msfem_nn_conv3x3_bf16_ex.argtypes = [i32, i32, i32, i32, vp, i32, vp, i32, vp, vp, i32, vp, vp, i32, vp]
        L.msfem_nn_head_bases.argtypes = [i32, i32, i32, vp, vp, i32, vp, i32, f32, vp, vp]
        L.msfem_nn_conv3x3_first_bf16.argtypes = [i32, i32, i32, vp, vp, i32, vp, vp, i32, vp, vp, i32, vp, vp]
        L.msfem_nn_col2im.argtypes = [i32, i32, i32, i32, i32, i32, i32, vp, vp, vp]
        L.msfem_nn_bias_act.argtypes = [i32, u64, vp, vp, i32, vp]
        L.msfem_nn_act_bwd.argtypes = [u64, vp, vp, vp, i32, vp]
        L.msfem_nn_bias_grad.argtypes = [i32, u64, vp, vp, vp]
        L.msfem_nn_pool2.argtypes = [i32, i32, i32, vp, vp, i32, vp]
        L.msfem_nn_up2.argtypes = [i32, i32, i32, vp, vp, i32, vp]
        L.msfem_nn_axpy.argtypes = [u64, f32, vp, vp, vp]
        L.msfem_nn_axpby.argtypes = [u64, f32, vp, f32, vp, vp, vp]
        L.msfem_nn_to_bf16.argtypes = [u64, vp, vp, vp]
        L.msfem_nn_mse.argtypes = [u64, vp, vp, vp, vp, i32, vp]
        L.msfem_nn_adam.argtypes = [u64, vp, vp, vp, vp, vp, f32, f32, f32, f32, i32, f32, vp]
        L.msfem_nn_delta_to_bases.argtypes = [i32, i32, vp, vp, vp]
        L.msfem_nn_bases_to_delta.argtypes = [i32, i32, vp, vp, vp]
        L._nn_declared = True
    return L


def bf16_bits(a):
    """FP32 values -> BF16 bit patterns (round to nearest even), as uint16."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def _ok(rc):
    if rc != 0:
        raise msfem.MsfemError(rc, "nn primitive failed")


class Buffers:
    """Grow-only device scratch by name (one set per network)."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.bufs = {}

    def get(self, name, n, dtype):
        b = self.bufs.get(name)
        need = int(n) * np.dtype(dtype).itemsize
        if b is None or b.nbytes < need:
            if b is not None:
                b.free()
            b = self.ctx.alloc(max(int(n), 1), dtype)
            self.bufs[name] = b
        return b


class Conv:
    """kxk convolution (stride 1, padding pad) + bias + optional ReLU."""

    def __init__(self, ctx, name, cin, cout, k, pad, relu, rng):
        self.ctx, self.name, self.cin, self.cout, self.k, self.pad, self.relu = ctx, name, cin, cout, k, pad, relu
        self.K = cin * k * k
        w = rng.standard_normal((cout, self.K)).astype(np.float32) * np.float32(np.sqrt(2.0 / self.K))
        if not relu:
            w *= np.float32(0.1)
        self.w = ctx.to_device(w)
        self.b = ctx.to_device(np.zeros(cout, np.float32))
        self.wbf = ctx.alloc(w.size, np.uint16)
        self.gw = ctx.alloc(w.size, np.float32)
        self.gb = ctx.alloc(cout, np.float32)
        self.mw, self.vw = ctx.to_device(np.zeros_like(w)), ctx.to_device(np.zeros_like(w))
        self.mb, self.vb = ctx.to_device(np.zeros(cout, np.float32)), ctx.to_device(np.zeros(cout, np.float32))
        # BF16 weights regrouped for the direct 3x3 kernel, rebuilt when wbf changes
        self.wver, self.wpk, self.wpk_ver = 0, None, -1
        self.stream = None  # None: the context stream
        self.sync_bf16()

    @property
    def st(self):
        """The stream this layer's kernels are enqueued on."""
        return self.ctx.stream if self.stream is None else self.stream

    def n_params(self):
        return self.cout * self.K + self.cout

    def sync_bf16(self):
        _ok(_nn().msfem_nn_to_bf16(self.cout * self.K, self.w.ptr, self.wbf.ptr, self.st))
        self.wver += 1

    def packed(self):
        """Device address of the weights regrouped for msfem_nn_conv3x3_bf16 (repacked after a change)."""
        if self.wpk_ver != self.wver:
            if self.wpk is None:
                self.wpk = self.ctx.alloc(self.cout * self.K, np.uint16)
            _ok(_nn().msfem_nn_conv3x3_pack(self.cin, self.cout, self.wbf.ptr, self.wpk.ptr, self.st))
            self.wpk_ver = self.wver
        return self.wpk.ptr

    def forward_up_skip(self, x, B, H, W, skip, out, io=0):
        """Serving form of a decoder layer: out = act(conv(up2(x)) + b) + skip with x
        at H/2 x W/2, upsampled inside the convolution's staging (no saved state);
        io as msfem_nn_conv3x3_bf16_ex (bit 0: x holds BF16, bit 1: out gets BF16)."""
        _ok(_nn().msfem_nn_conv3x3_bf16_ex(self.cin, B, H, W, self.packed(), self.cout, x.ptr, 1, out.ptr,
                                           self.b.ptr, 1 if self.relu else 0, skip.ptr, None, io, self.st))

    def forward_pool(self, x, B, H, W, buf, pool):
        """Serving form of an encoder layer followed by 2x2 average pooling:
        y = act(conv(x) + b) (kept, the skip) and pool = pool2(y) in one pass,
        bitwise the values of forward + msfem_nn_pool2."""
        y = buf.get(self.name + ".y", self.cout * B * H * W, np.float32)
        _ok(_nn().msfem_nn_conv3x3_bf16_ex(self.cin, B, H, W, self.packed(), self.cout, x.ptr, 0, y.ptr, self.b.ptr,
                                           1 if self.relu else 0, None, pool.ptr, 0, self.st))
        return y

    def small_direct(self):
        """Whether msfem_nn_conv_small takes this layer (the head: at most 8 output channels)."""
        return self.cout <= 8 and self.k <= 3 and self.K * (4 if self.cout <= 4 else 8) <= 16384

    def direct3x3(self, B, H, W):
        """Whether msfem_nn_conv3x3_bf16 takes this layer on B images of H x W."""
        if not (self.k == 3 and self.pad == 1 and self.cin % 32 == 0 and self.cout % 32 == 0):
            return False
        if W in (4, 8):  # tiles of 128 / W^2 whole W x W images
            return H == W and (B * H * W) % 128 == 0
        return W in (16, 32, 64) and H % (128 // W) == 0

    def out_hw(self, H, W):
        return H + 2 * self.pad - self.k + 1, W + 2 * self.pad - self.k + 1

    def forward(self, x, B, H, W, buf):
        L, st = _nn(), self.st
        Ho, Wo = self.out_hw(H, W)
        N = B * Ho * Wo
        y = buf.get(self.name + ".y", self.cout * N, np.float32)
        # deep layers have few output tiles: split K (fixed-order reduction) to fill the GPU
        tiles = -(-self.cout // (32 if self.cout <= 32 else 64 if self.cout <= 64 else 128)) * -(-N // 128)
        splits = int(max(1, min(256 // tiles, self.K // 256))) if tiles < 256 else 1
        work = buf.get("splitk", splits * self.cout * N, np.float32) if splits > 1 else None
        wptr = work.ptr if work is not None else None
        if self.direct3x3(B, H, W):  # input rows staged once per channel block, 9 taps from LDS
            _ok(L.msfem_nn_conv3x3_bf16(self.cin, B, H, W, self.packed(), self.cout, x.ptr, 0, y.ptr, self.b.ptr,
                                        1 if self.relu else 0, None, st))
        elif self.small_direct():  # the head: too few output channels for the matrix cores
            _ok(L.msfem_nn_conv_small(self.cin, B, H, W, self.k, self.k, self.pad, self.wbf.ptr, self.cout, x.ptr,
                                      y.ptr, self.b.ptr, 1 if self.relu else 0, st))
        elif Wo % 8 == 0:  # implicit GEMM, bias and activation in its epilogue
            _ok(L.msfem_nn_conv_bf16(self.cin, B, H, W, self.k, self.k, self.pad, self.wbf.ptr, self.cout, x.ptr,
                                     y.ptr, splits, wptr, self.b.ptr, 1 if self.relu else 0, st))
        else:
            col = buf.get("col", self.K * N, np.uint16)
            _ok(L.msfem_nn_im2col(self.cin, B, H, W, self.k, self.k, self.pad, x.ptr, col.ptr, st))
            _ok(L.msfem_nn_gemm_bf16_splitk(0, 0, self.cout, N, self.K, self.wbf.ptr, self.K, col.ptr, N, y.ptr, N, 0,
                                            splits, wptr, st))
            _ok(L.msfem_nn_bias_act(self.cout, N, self.b.ptr, y.ptr, 1 if self.relu else 0, st))
        self.saved = (x, B, H, W, y)
        return y, Ho, Wo

    def backward(self, dy, buf, want_dx=True):
        """dy: gradient w.r.t. the layer output (modified in place).  Returns dx or None."""
        L, st = _nn(), self.st
        x, B, H, W, y = self.saved
        Ho, Wo = self.out_hw(H, W)
        N = B * Ho * Wo
        dybf = buf.get("dybf", self.cout * N, np.uint16)
        _ok(L.msfem_nn_act_bwd(self.cout * N, y.ptr, dy.ptr, dybf.ptr, 1 if self.relu else 0, st))
        _ok(L.msfem_nn_bias_grad(self.cout, N, dy.ptr, self.gb.ptr, st))
        col = buf.get("col", self.K * N, np.uint16)
        _ok(L.msfem_nn_im2col(self.cin, B, H, W, self.k, self.k, self.pad, x.ptr, col.ptr, st))
        # weight gradient: M x N = cout x K small, K = pixels long -> split K
        tiles = -(-self.cout // (32 if self.cout <= 32 else 64 if self.cout <= 64 else 128)) * -(-self.K // 128)
        splits = int(max(1, min(512 // tiles, N // 1024)))
        work = buf.get("splitk", splits * self.cout * self.K, np.float32) if splits > 1 else None
        _ok(L.msfem_nn_gemm_bf16_splitk(0, 1, self.cout, self.K, N, dybf.ptr, N, col.ptr, N, self.gw.ptr, self.K, 0,
                                        splits, work.ptr if work is not None else None, st))
        if not want_dx:
            return None
        dcol = buf.get("dcol", self.K * N, np.float32)
        _ok(L.msfem_nn_gemm_bf16(1, 0, self.K, N, self.cout, self.wbf.ptr, self.K, dybf.ptr, N, dcol.ptr, N, 0, st))
        dx = buf.get(self.name + ".dx", self.cin * B * H * W, np.float32)
        _ok(L.msfem_nn_col2im(self.cin, B, H, W, self.k, self.k, self.pad, dcol.ptr, dx.ptr, st))
        return dx

    def adam(self, lr, step, wd=0.0, b1=0.9, b2=0.999, eps=1e-8):
        L, st = _nn(), self.st
        _ok(L.msfem_nn_adam(self.cout * self.K, self.w.ptr, self.gw.ptr, self.mw.ptr, self.vw.ptr, self.wbf.ptr, lr, b1,
                            b2, eps, step, wd, st))
        self.wver += 1
        _ok(L.msfem_nn_adam(self.cout, self.b.ptr, self.gb.ptr, self.mb.ptr, self.vb.ptr, None, lr, b1, b2, eps, step,
                            0.0, st))

    def get(self):
        return self.w.numpy().reshape(self.cout, self.K), self.b.numpy()

    def set(self, w, b):
        self.w.upload(np.asarray(w, np.float32).reshape(-1))
        self.b.upload(np.asarray(b, np.float32))
        self.sync_bf16()


class Surrogate:
    """Encoder-decoder of `levels` resolutions; `chans[l]` channels at level l."""

    def __init__(self, ctx, tile=64, chans=(32, 64, 128, 256, 256, 256), out_ch=4, seed=0):
        self.ctx, self.tile, self.chans, self.out_ch = ctx, tile, tuple(chans), out_ch
        self.levels = len(chans)
        assert tile % (1 << (self.levels - 1)) == 0
        rng = np.random.default_rng(seed)
        c = self.chans
        self.enc = [Conv(ctx, "e0a", 1, c[0], 3, 1, True, rng), Conv(ctx, "e0", c[0], c[0], 3, 1, True, rng)]
        self.enc += [Conv(ctx, f"e{l}", c[l - 1], c[l], 3, 1, True, rng) for l in range(1, self.levels)]
        self.dec = [Conv(ctx, f"d{l}", c[l + 1], c[l], 3, 1, True, rng) for l in range(self.levels - 1)]
        self.head = Conv(ctx, "head", c[0], out_ch, 2, 0, False, rng)
        self.stream = None  # None: the context stream
        self.layers = self.enc + self.dec + [self.head]
        self.buf = Buffers(ctx)
        self.step_count = 0
        self.dscale = 1.0  # network output = delta / dscale

    @property
    def st(self):
        """The stream the network's kernels are enqueued on."""
        return self.ctx.stream if self.stream is None else self.stream

    def set_stream(self, stream):
        """Enqueue the network's kernels on `stream` (None: the context stream)."""
        self.stream = stream
        for layer in self.layers:
            layer.stream = stream

    def n_params(self):
        return sum(l.n_params() for l in self.layers)

    def flops_per_sample(self):
        """Forward multiply-adds x 2 of all convolutions for one tile."""
        f, s = 0, self.tile
        for l, conv in enumerate(self.enc):
            res = s >> max(0, l - 1)
            f += 2 * conv.K * conv.cout * res * res
        for l, conv in enumerate(self.dec):
            res = s >> l
            f += 2 * conv.K * conv.cout * res * res
        f += 2 * self.head.K * self.out_ch * (s - 1) ** 2
        return f

    # ---- forward / backward ---------------------------------------------------------
    def forward(self, x, B, keep=True, bases=None):
        """x: DeviceArray [1][B][tile][tile] FP32 (normalized log-coefficient).
        Returns the head output [4][B][tile-1][tile-1] (DeviceArray view); with
        `bases` (keep=False only: a device address of FP64 [B][4][(tile+1)^2])
        the head writes N_i + dscale * prediction there instead
        (msfem_nn_head_bases) and None is returned.
        keep=False (inference): the first two layers run as one pass, encoder
        layers write their pooled output beside their own, and decoder layers
        that the direct 3x3 kernel takes read the upsampled input and add the
        skip inside the convolution, storing their output as BF16 where its only
        consumer rounds it to BF16 anyway (the next such layer, the head written
        as bases); the values are bitwise those of the training form, which
        keeps the intermediates the backward pass needs."""
        assert bases is None or not keep
        L, st, buf = _nn(), self.st, self.buf
        H = self.tile
        skips = []
        fused = not keep  # serving: each encoder layer writes its pooled output too
        pend = None  # pooled output of the last encoder layer, when it was fused
        if fused and self.first_fused(H):  # the first two layers in one pass, pooled
            pend = buf.get("pool1", self.chans[0] * B * (H // 2) ** 2, np.float32)
            e0a, e0 = self.enc[0], self.enc[1]
            h = buf.get(e0.name + ".y", e0.cout * B * H * H, np.float32)
            _ok(L.msfem_nn_conv3x3_first_bf16(B, H, H, e0a.wbf.ptr, e0a.b.ptr, 1 if e0a.relu else 0, x.ptr,
                                              e0.packed(), e0.cout, h.ptr, e0.b.ptr, 1 if e0.relu else 0, pend.ptr,
                                              st))
        elif fused and self.enc[1].direct3x3(B, H, H):
            h, _, _ = self.enc[0].forward(x, B, H, H, buf)
            pend = buf.get("pool1", self.chans[0] * B * (H // 2) ** 2, np.float32)
            h = self.enc[1].forward_pool(h, B, H, H, buf, pend)
        else:
            h, _, _ = self.enc[0].forward(x, B, H, H, buf)
            h, _, _ = self.enc[1].forward(h, B, H, H, buf)
        skips.append((h, H))
        for l in range(1, self.levels):
            C = self.chans[l - 1]
            if pend is not None:
                p, pend = pend, None
            else:
                p = buf.get(f"pool{l}", C * B * (H // 2) ** 2, np.float32)
                _ok(L.msfem_nn_pool2(C * B, H, H, h.ptr, p.ptr, 0, st))
            H //= 2
            if fused and l + 1 < self.levels and self.enc[l + 1].direct3x3(B, H, H):
                pend = buf.get(f"pool{l + 1}", self.chans[l] * B * (H // 2) ** 2, np.float32)
                h = self.enc[l + 1].forward_pool(p, B, H, H, buf, pend)
            else:
                h, _, _ = self.enc[l + 1].forward(p, B, H, H, buf)
            skips.append((h, H))
        self.skips = skips
        hb = False  # h holds BF16 values
        for l in range(self.levels - 2, -1, -1):
            C = self.chans[l + 1]
            if not keep and self.dec[l].direct3x3(B, 2 * H, 2 * H):
                H *= 2
                ob = bases is not None if l == 0 else self.dec[l - 1].direct3x3(B, 2 * H, 2 * H)
                out = buf.get(f"sum{l}" + ("h" if ob else ""), self.chans[l] * B * H * H,
                              np.uint16 if ob else np.float32)
                self.dec[l].forward_up_skip(h, B, H, H, skips[l][0], out, (1 if hb else 0) | (2 if ob else 0))
                h, hb = out, ob
                continue
            u = buf.get(f"up{l}", C * B * (2 * H) ** 2, np.float32)
            _ok(L.msfem_nn_up2(C * B, H, H, h.ptr, u.ptr, 0, st))
            H *= 2
            y, _, _ = self.dec[l].forward(u, B, H, H, buf)
            sk, _ = skips[l]
            # the sum goes to its own buffer: y (post-ReLU) is kept for the backward mask
            h = buf.get(f"sum{l}", self.chans[l] * B * H * H, np.float32)
            _ok(L.msfem_nn_axpby(self.chans[l] * B * H * H, 1.0, y.ptr, 1.0, sk.ptr, h.ptr, st))
        self.B = B
        if bases is not None:
            hd = self.head
            assert hd.k == 2 and hd.pad == 0 and hd.cout == 4 and H == self.tile
            _ok(L.msfem_nn_head_bases(hd.cin, B, H, hd.wbf.ptr, h.ptr, 1 if hb else 0, hd.b.ptr, 1 if hd.relu else 0,
                                      self.dscale, bases, st))
            return None
        out, _, _ = self.head.forward(h, B, H, H, buf)
        return out

    def first_fused(self, H):
        """Whether msfem_nn_conv3x3_first_bf16 takes the first two layers (1 -> 32 -> 32
        channels, 3x3, on 64-wide tiles)."""
        a, b = self.enc[0], self.enc[1]
        return (H == 64 and a.cin == 1 and a.k == 3 and a.pad == 1 and a.cout == 32 and b.cin == 32 and b.k == 3
                and b.pad == 1 and b.cout == 32)

    def backward(self, dout):
        """dout: gradient of the loss w.r.t. the head output (consumed)."""
        L, st, buf = _nn(), self.st, self.buf
        B = self.B
        dh = self.head.backward(dout, buf)
        H = self.tile
        # gradient w.r.t. each encoder output: its skip share, then its consumer's share
        dskip = [None] * self.levels
        for l in range(self.levels - 1):
            # d(h_l) where h_l = dec_l(up(h_{l+1})) + skip_l: goes to skip_l and through dec_l
            n = self.chans[l] * B * H * H
            ds = buf.get(f"dskip{l}", n, np.float32)
            _ok(L.msfem_nn_axpby(n, 1.0, dh.ptr, 0.0, None, ds.ptr, st))  # ds = dh (the decoder masks dh in place)
            dskip[l] = ds
            du = self.dec[l].backward(dh, buf)
            C = self.chans[l + 1]
            dlow = buf.get(f"dlow{l}", C * B * (H // 2) ** 2, np.float32)
            _ok(L.msfem_nn_up2(C * B, H // 2, H // 2, du.ptr, dlow.ptr, 1, st))
            dh = dlow
            H //= 2
        # encoder, coarsest first: dh holds d(h_{levels-1})
        for l in range(self.levels - 1, 0, -1):
            if dskip[l] is not None:
                _ok(L.msfem_nn_axpy(self.chans[l] * B * H * H, 1.0, dskip[l].ptr, dh.ptr, st))
            dp = self.enc[l + 1].backward(dh, buf)
            C = self.chans[l - 1]
            dprev = buf.get(f"dprev{l}", C * B * (2 * H) ** 2, np.float32)
            _ok(L.msfem_nn_pool2(C * B, 2 * H, 2 * H, dp.ptr, dprev.ptr, 1, st))
            dh = dprev
            H *= 2
        _ok(L.msfem_nn_axpy(self.chans[0] * B * H * H, 1.0, dskip[0].ptr, dh.ptr, st))
        dh = self.enc[1].backward(dh, buf)
        self.enc[0].backward(dh, buf, want_dx=False)

    def loss_and_grad(self, out, target, n):
        """MSE over n values; returns (loss, grad DeviceArray)."""
        L, st = _nn(), self.st
        grad = self.buf.get("dout", n, np.float32)
        nb = 256
        part = self.buf.get("loss_part", nb, np.float32)
        _ok(L.msfem_nn_mse(n, out.ptr, target.ptr, grad.ptr, part.ptr, nb, st))
        loss = float(part.numpy()[:nb].astype(np.float64).sum() / n)
        return loss, grad

    def train_step(self, x, target, B, lr, wd=0.0):
        out = self.forward(x, B)
        n = self.out_ch * B * (self.tile - 1) ** 2
        loss, grad = self.loss_and_grad(out, target, n)
        self.backward(grad)
        self.step_count += 1
        for layer in self.layers:
            layer.adam(lr, self.step_count, wd)
        return loss

    # ---- parameters -------------------------------------------------------------------
    def state(self):
        return {l.name + suf: v for l in self.layers for suf, v in zip((".w", ".b"), l.get())}

    def load_state(self, st):
        for l in self.layers:
            l.set(st[l.name + ".w"], st[l.name + ".b"])

    def save(self, path, compact=False):
        """compact: weights stored as their BF16 bits (what the matrix cores use;
        serving from a compact file is bit-identical), biases FP32."""
        st = self.state()
        if compact:
            st = {(k + "bf" if k.endswith(".w") else k): (bf16_bits(v) if k.endswith(".w") else v) for k, v in st.items()}
        save = np.savez_compressed if compact else np.savez
        save(path, **st, chans=np.array(self.chans), tile=self.tile, dscale=self.dscale)

    @staticmethod
    def load(ctx, path):
        d = np.load(path)
        s = Surrogate(ctx, int(d["tile"]), tuple(int(v) for v in d["chans"]))
        st = {}
        for k in d.files:
            if k.endswith(".w") or k.endswith(".b"):
                st[k] = d[k]
            elif k.endswith(".wbf"):
                st[k[:-2]] = (d[k].astype(np.uint32) << 16).view(np.float32)
        s.load_state(st)
        s.dscale = float(d["dscale"])
        return s

    # ---- serving ------------------------------------------------------------------------
    def predict_bases(self, x, n, bases_ptr, batch=2048):
        """x: device FP32 [n][tile*tile] normalized log-coefficients (g / CLIP);
        bases_ptr: device FP64 bases [n][4][(tile+1)^2], filled with N_i + dscale *
        prediction (the x0 of msfem_batch_build_device), `batch` tiles per forward
        pass.  Enqueued on the network's stream (the context stream unless
        set_stream chose another), not synchronized."""
        t = self.tile
        nbase = 4 * (t + 1) ** 2
        for s in range(0, n, batch):
            b = min(batch, n - s)
            self.forward(Ptr(x.ptr + s * t * t * 4), b, keep=False, bases=bases_ptr + s * nbase * 8)


class Ptr:
    """A bare device address (a batch slice of a DeviceArray)."""

    def __init__(self, ptr):
        self.ptr = ptr


CORR_LEN = 4.0  # correlation length of the training/serving fields, in fine cells


class TileSet:
    """n cells of nf x nf fine cells with alpha = exp(g), g a Gaussian random
    field (msfem_gen_grf_log, clipped to contrast 1e3), resident on the GPU:
    x = g / CLIP (FP32, the network input), coef = exp(g) (FP64, the solver
    input), unit boxes (the bases of a scalar coefficient on a square cell do
    not depend on its size)."""

    def __init__(self, ctx, n, seed, nf=64):
        self.ctx, self.n, self.nf, self.seed = ctx, n, nf, seed
        g = msfem.grf_log(n, nf, seed, CORR_LEN, CLIP)
        self.x = ctx.to_device((g / np.float32(CLIP)).astype(np.float32).reshape(-1))
        self.coef_host = np.exp(g.astype(np.float64)).reshape(-1)
        self.coef = ctx.to_device(self.coef_host)
        side = int(np.ceil(np.sqrt(n)))
        self.box_host = msfem.unit_boxes(2, side)[:n]
        self.box = ctx.to_device(np.ascontiguousarray(self.box_host).reshape(-1))
        self.n_fine = np.full(n, nf, np.int32)
        nb = 4 * (nf + 1) ** 2
        self.bases = ctx.alloc(n * nb, np.float64)
        self.Ae = ctx.alloc(n * 16, np.float64)
        self.Fe = ctx.alloc(n * 4, np.float64)
        self.iters = ctx.alloc(n * 4, np.int32)
        self.status = ctx.alloc(n * 4, np.int32)

    def solve(self, opts, x0_ptr=None, c0=0, c1=None, bound_ptr=None):
        """Batched basis build of cells c0..c1-1 (default: all) into self.bases, on
        the context stream (enqueued; returns nothing).  x0_ptr: initial bases of
        all n cells (the slice of c0..c1-1 is used).  bound_ptr: device FP64
        [n][4] for the a posteriori energy-norm bounds of all n cells (the slice
        of c0..c1-1 is written)."""
        c1 = self.n if c1 is None else c1
        nb = 4 * (self.nf + 1) ** 2
        self.ctx.batch_build_device(2, self.n_fine[c0:c1], self.box.ptr + c0 * 4 * 8, msfem.COEF_SCALAR_CELL,
                                    self.coef.ptr + c0 * self.nf ** 2 * 8, self.Ae.ptr + c0 * 16 * 8,
                                    self.Fe.ptr + c0 * 4 * 8, self.bases.ptr + c0 * nb * 8,
                                    self.iters.ptr + c0 * 4 * 4, self.status.ptr + c0 * 4 * 4,
                                    rhs_kind=msfem.RHS_CONST, rhs_value=1.0,
                                    x0_ptr=None if x0_ptr is None else x0_ptr + c0 * nb * 8, opts=opts,
                                    stream=self.ctx.stream,
                                    energy_bound_ptr=None if bound_ptr is None else bound_ptr + c0 * 4 * 8)

    def free(self):
        for a in (self.x, self.coef, self.box, self.bases, self.Ae, self.Fe, self.iters, self.status):
            a.free()


def targets(net, data, s, b, out):
    """Scaled training targets of cells s..s+b: delta / dscale, [4][b][(nf-1)^2] FP32."""
    L, st = _nn(), net.ctx.stream
    nf = data.nf
    _ok(L.msfem_nn_bases_to_delta(b, nf, data.bases.ptr + s * 4 * (nf + 1) ** 2 * 8, out.ptr, st))
    _ok(L.msfem_nn_axpby(4 * b * (nf - 1) ** 2, 1.0 / net.dscale, out.ptr, 0.0, None, out.ptr, st))


def evaluate(net, data, start, count, batch=64):
    """Mean squared error (in scaled units) of the predictions for cells start..start+count."""
    n_per = 4 * (net.tile - 1) ** 2
    tgt = net.buf.get("eval_t", batch * n_per, np.float32)
    tot, m = 0.0, 0
    for s in range(start, start + count, batch):
        b = min(batch, start + count - s)
        targets(net, data, s, b, tgt)
        out = net.forward(Ptr(data.x.ptr + s * data.nf ** 2 * 4), b, keep=False)
        loss, _ = net.loss_and_grad(out, tgt, b * n_per)
        tot += loss * b
        m += b
    return tot / m


def train(net, data, n_train, epochs, batch=32, lr=1e-3, lr_min=2e-5, wd=0.0, log=print, ckpt=None, seed=0,
          n_val=0):
    """Adam with a cosine learning-rate decay over `epochs` passes of the first
    n_train cells (minibatches of consecutive cells, visited in a fresh random
    order every epoch); the next n_val cells are the validation set."""
    import time
    rng = np.random.default_rng(seed)
    n_per = 4 * (net.tile - 1) ** 2
    tgt = net.buf.get("train_t", batch * n_per, np.float32)
    nb = n_train // batch
    total, step, hist = epochs * nb, 0, []
    t0 = tlast = time.perf_counter()
    for ep in range(epochs):
        run = 0.0
        for k, mb in enumerate(rng.permutation(nb)):
            s = int(mb) * batch
            targets(net, data, s, batch, tgt)
            a = lr_min + 0.5 * (lr - lr_min) * (1 + np.cos(np.pi * step / total))
            run += net.train_step(Ptr(data.x.ptr + s * data.nf ** 2 * 4), tgt, batch, float(a), wd)
            step += 1
            if time.perf_counter() - tlast > 30:
                tlast = time.perf_counter()
                log(f"epoch {ep} step {k + 1}/{nb} train mse {run / (k + 1):.4e} lr {a:.2e} "
                    f"{tlast - t0:.0f} s")
        val = evaluate(net, data, n_train, n_val) if n_val else float("nan")
        hist.append({"epoch": ep, "train_mse": run / nb, "val_mse": val, "seconds": time.perf_counter() - t0})
        log(f"epoch {ep}: train mse {run / nb:.4e} val mse {val:.4e} ({time.perf_counter() - t0:.1f} s)")
        if ckpt:
            net.save(ckpt)
    return hist


def delta_std(ctx, data, count=256):
    """RMS of the interior corrections of the first `count` cells (the target scale)."""
    nf = data.nf
    d = ctx.alloc(4 * count * (nf - 1) ** 2, np.float32)
    _ok(_nn().msfem_nn_bases_to_delta(count, nf, data.bases.ptr, d.ptr, ctx.stream))
    v = d.numpy().astype(np.float64)
    d.free()
    return float(np.sqrt((v * v).mean()))


def main(argv=None):
    """Train the surrogate on GPU solves: python surrogate.py --out weights.npz [--cells 5000 --epochs 30]."""
    import argparse
    import json
    import sys
    import time
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None, help="checkpoint (.npz), rewritten after every epoch")
    ap.add_argument("--cells", type=int, default=5000, help="training cells (4 local solves each)")
    ap.add_argument("--val", type=int, default=256, help="validation cells (drawn after the training cells)")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=40, help="Philox seed of the training fields")
    ap.add_argument("--init-seed", type=int, default=0)
    ap.add_argument("--history", default=None, help="JSON file for the per-epoch losses")
    ap.add_argument("--load", default=None, help="skip training: serve this checkpoint")
    ap.add_argument("--serve", type=int, default=0, help="after training, report serving on an N x N coarse mesh")
    ap.add_argument("--serve-seed", type=int, default=41)
    ap.add_argument("--report", default=None, help="JSON file for the serving report")
    ap.add_argument("--time-forward", type=int, default=0,
                    help="time N forward passes over random tiles (after one untimed pass)")
    ap.add_argument("--forward-batch", type=int, default=128, help="--time-forward: tiles per pass")
    ap.add_argument("--forward-form", choices=("serving", "training", "both", "predict"), default="both",
                    help="--time-forward: the serving form (fused layers) and/or the training form, or "
                         "predict: the serving form with the head written as bases (predict_bases' pass)")
    a = ap.parse_args(argv)
    if not a.load and not a.out:
        ap.error("--out (train) or --load (serve) is required")

    def log(msg):
        print(msg, flush=True)
        sys.stdout.flush()

    ctx = msfem.Context(0)
    if a.load:
        net = Surrogate.load(ctx, a.load)
    else:
        net = train_new(ctx, a, log)
    if a.time_forward:
        fb = a.forward_batch
        x = ctx.to_device(np.random.default_rng(0).standard_normal(fb * net.tile ** 2).astype(np.float32))
        forms = {"serving": ("serving",), "training": ("training",), "both": ("serving", "training"),
                 "predict": ("predict",)}[a.forward_form]
        bases = ctx.alloc(fb * 4 * (net.tile + 1) ** 2, np.float64) if "predict" in forms else None
        for form in forms:
            run = lambda: net.forward(x, fb, keep=form == "training",
                                      bases=bases.ptr if form == "predict" else None)
            run()
            ctx.synchronize()
            t = time.perf_counter()
            for _ in range(a.time_forward):
                run()
            ctx.synchronize()
            dt = (time.perf_counter() - t) / a.time_forward
            log(f"{form} forward, {fb} tiles: {dt * 1e3:.3f} ms "
                f"({net.flops_per_sample() * fb / dt / 1e12:.1f} TFLOP/s)")
        x.free()
        if bases is not None:
            bases.free()
    if a.serve:
        rep = serve_report(ctx, net, n_side=a.serve, seed=a.serve_seed, log=log)
        log(json.dumps(rep))
        if a.report:
            with open(a.report, "w") as f:
                json.dump(rep, f, indent=1)
    ctx.close()


def train_new(ctx, a, log):
    import json
    import time
    t0 = time.perf_counter()
    data = TileSet(ctx, a.cells + a.val, a.seed)
    log(f"fields: {a.cells + a.val} cells of 64x64, corr. length {CORR_LEN}, seed {a.seed} "
        f"({time.perf_counter() - t0:.1f} s)")
    t1 = time.perf_counter()
    data.solve(msfem.default_opts(precision=msfem.PREC_FP64))
    ctx.synchronize()
    st = data.status.numpy()
    if st.any():
        raise RuntimeError(f"{int((st != 0).sum())} training solves did not converge")
    its = data.iters.numpy()
    log(f"training solves: {4 * (a.cells + a.val)} FP64 local solves in {time.perf_counter() - t1:.2f} s, "
        f"mean {its.mean():.1f} it")
    net = Surrogate(ctx, seed=a.init_seed)
    net.dscale = delta_std(ctx, data)
    log(f"network: {net.n_params()} parameters, {net.flops_per_sample() / 1e6:.1f} MFLOP/tile forward, "
        f"target rms {net.dscale:.4e}")
    hist = train(net, data, a.cells, a.epochs, batch=a.batch, lr=a.lr, log=log, ckpt=a.out, n_val=a.val)
    if a.history:
        with open(a.history, "w") as f:
            json.dump({"cells": a.cells, "val": a.val, "epochs": a.epochs, "batch": a.batch, "lr": a.lr,
                       "params": net.n_params(), "dscale": net.dscale, "history": hist}, f, indent=1)
    data.free()
    return net


# Q1 stiffness of a square fine cell (any size in 2D), nodes (0,0),(1,0),(0,1),(1,1)
_KQ1 = np.array([[4, -1, -1, -2], [-1, 4, -2, -1], [-1, -2, 4, -1], [-2, -1, -1, 4]], np.float64) / 6.0


def energy_norm2(alpha, v):
    """v^T K v for the fine Q1 stiffness of one cell: alpha [nf][nf] per fine
    cell, v [..., nf+1, nf+1] nodal values (any leading axes)."""
    c = [v[..., :-1, :-1], v[..., :-1, 1:], v[..., 1:, :-1], v[..., 1:, 1:]]
    e = sum(_KQ1[i, j] * c[i] * c[j] for i in range(4) for j in range(4))
    return (alpha * e).sum(axis=(-2, -1))


def energy_error(alpha, phi, ref):
    """max over the cell's bases of ||phi_i - ref_i||_K / ||ref_i||_K."""
    return float(np.max(np.sqrt(energy_norm2(alpha, phi - ref) / energy_norm2(alpha, ref))))


def serve_pipelined(net, data, x0, opts, chunks=4, mark0=2, batch=2048, bound_ptr=None):
    """One serving step -- prediction and warm-started basis build of all cells
    of `data` -- in `chunks` chunks: the inference of every chunk runs on the
    context's second stream (a mark after each), the build of chunk k on the
    context stream behind chunk k's mark, enqueued right after the inference of
    chunk k + 1, so that inference (matrix cores, memory) and builds (VALU)
    overlap.  The values are bitwise those of predict_bases + one build of all
    cells.  Uses marks mark0 .. mark0 + chunks; the inference first waits for
    the context stream, so x0 is not rewritten under a previous step's builds."""
    ctx, t = net.ctx, net.tile
    n, nb = data.n, 4 * (data.nf + 1) ** 2
    if mark0 < 0 or mark0 + chunks >= msfem.N_MARKS:
        raise ValueError("not enough marks for the chunks")
    # chunk bounds on multiples of the inference batch
    per = -(-n // chunks)
    step = -(-per // batch) * batch if n > batch else n
    bounds = [(c, min(n, c + step)) for c in range(0, n, step)]
    aux = ctx.aux_stream
    ctx.mark_on(mark0, None)
    ctx.wait_mark(mark0, aux)

    def infer(k):
        c0, c1 = bounds[k]
        net.set_stream(aux)
        try:
            net.predict_bases(Ptr(data.x.ptr + c0 * t * t * 4), c1 - c0, x0.ptr + c0 * nb * 8, batch=batch)
        finally:
            net.set_stream(None)
        ctx.mark_on(mark0 + 1 + k, aux)

    def build(k):
        ctx.wait_mark(mark0 + 1 + k, None)
        data.solve(opts, x0.ptr, *bounds[k], bound_ptr=bound_ptr)

    infer(0)
    for k in range(len(bounds)):
        if k + 1 < len(bounds):
            infer(k + 1)
        build(k)


def serve_report(ctx, net, n_side=128, seed=41, rel_tol=1e-8, reps=3, warmup=1, sample=64, log=None, data=None,
                 pipe_batch=2048):
    """Times one step -- the basis build of n_side^2 fresh GRF cells plus the
    coarse assembly -- five ways: plain FP64, FP32+FP64 refinement, and each of
    them warm-started from the network's prediction (inference included in the
    step), FP32 also with the inference of one chunk of cells overlapping the
    builds of the previous one (serve_pipelined).  Reports cells/s (wall clock over `reps` steps), device times of the
    inference and of the local solves (HIP events on the context stream),
    iterations, the fallback rate (share of local problems whose prediction
    failed the residual gate and needed CG iterations) and the energy-norm
    error of a cell sample against plain FP64."""
    import time
    n = n_side * n_side
    data = data or TileSet(ctx, n, seed)
    nf = data.nf
    per = 4 * (nf + 1) ** 2
    x0 = ctx.alloc(n * per, np.float64)
    ncn = (n_side + 1) ** 2
    Kc, Fc = ctx.alloc(ncn * 9, np.float64), ctx.alloc(ncn, np.float64)
    cm, cv = ctx.alloc(ncn, np.uint8), ctx.alloc(ncn, np.float64)
    pick = np.unique(np.linspace(0, n - 1, min(sample, n)).astype(np.int64))
    fp32 = msfem.default_opts(precision=msfem.PREC_FP32_REFINE, rel_tol=rel_tol)
    legs = [("fp64", msfem.default_opts(precision=msfem.PREC_FP64), None),
            ("fp32_refine", fp32, None),
            ("fp32_refine_surrogate", fp32, "serial"),
            ("fp32_refine_surrogate_pipelined", fp32, "pipelined"),
            ("fp64_surrogate", msfem.default_opts(precision=msfem.PREC_FP64), "serial")]
    ctx.set_profiling(True)
    # inference alone (device time between two marks, per pass over all cells)
    for _ in range(warmup):
        net.predict_bases(data.x, n, x0.ptr)
    ctx.mark(0)
    for _ in range(reps):
        net.predict_bases(data.x, n, x0.ptr)
    ctx.mark(1)
    infer_ms = ctx.elapsed(0, 1) / reps
    pred = {int(c): x0.read(int(c) * per, per).reshape(4, nf + 1, nf + 1) for c in pick}
    alpha = {int(c): data.coef_host[int(c) * nf * nf:(int(c) + 1) * nf * nf].reshape(nf, nf) for c in pick}
    out, ref = {}, None
    for name, opts, use in legs:
        solve_ms = []

        def step():
            if use == "pipelined":  # inference of chunk k+1 beside the builds of chunk k
                serve_pipelined(net, data, x0, opts, batch=pipe_batch)
            else:
                if use:
                    net.predict_bases(data.x, n, x0.ptr)
                data.solve(opts, x0.ptr if use else None)
            ctx.coarse_assemble_device(2, n_side, data.Ae.ptr, data.Fe.ptr, msfem.BC_ZERO_DIRICHLET, Kc.ptr, Fc.ptr,
                                       cm.ptr, cv.ptr, stream=ctx.stream)
        for _ in range(warmup):
            step()
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            step()
            if use != "pipelined":  # (reading the build timers waits for the build)
                solve_ms.append(ctx.profile()[0])
        ctx.synchronize()
        ms = (time.perf_counter() - t) * 1e3 / reps
        st = data.status.numpy()
        if st.any():
            raise RuntimeError(f"{name}: {int((st != 0).sum())} local solves did not converge")
        its = data.iters.numpy().astype(np.int64)
        bases = {int(c): data.bases.read(int(c) * per, per).reshape(4, nf + 1, nf + 1) for c in pick}
        if ref is None:
            ref = bases
        r = {"ms": round(ms, 3), "cells_per_s": round(n / (ms * 1e-3), 1),
             "solve_kernel_ms": round(float(np.mean(solve_ms)), 3) if solve_ms else None,
             "iters_mean": round(float(its.mean()), 2),
             "iters_p99": int(np.percentile(its, 99)),
             "energy_err_max": float(max(energy_error(alpha[c], bases[c], ref[c]) for c in pick))}
        if use:
            r["fallback_rate"] = round(float((its > 0).mean()), 4)
        out[name] = r
        if log:
            log(f"{name}: {r}")
    errs = [energy_error(alpha[c], pred[c], ref[c]) for c in pick]
    out["prediction"] = {"infer_ms": round(infer_ms, 3), "tiles_per_s": round(n / (infer_ms * 1e-3), 1),
                         "energy_err_median": float(np.median(errs)), "energy_err_max": float(np.max(errs)),
                         "tflops": round(net.flops_per_sample() * n / (infer_ms * 1e-3) / 1e12, 2)}
    out["cells"], out["sample"], out["rel_tol"] = n, len(pick), rel_tol
    for arr in (x0, Kc, Fc, cm, cv):
        arr.free()
    return out

if __name__ == "__main__":
    main()
