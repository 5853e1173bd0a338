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
        L.msfem_nn_im2col.argtypes = [i32, i32, i32, i32, i32, i32, i32, vp, vp, vp]
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
        self.sync_bf16()

    def n_params(self):
        return self.cout * self.K + self.cout

    def sync_bf16(self):
        _ok(_nn().msfem_nn_to_bf16(self.cout * self.K, self.w.ptr, self.wbf.ptr, self.ctx.stream))

    def out_hw(self, H, W):
        return H + 2 * self.pad - self.k + 1, W + 2 * self.pad - self.k + 1

    def forward(self, x, B, H, W, buf):
        L, st = _nn(), self.ctx.stream
        Ho, Wo = self.out_hw(H, W)
        N = B * Ho * Wo
        col = buf.get("col", self.K * N, np.uint16)
        y = buf.get(self.name + ".y", self.cout * N, np.float32)
        _ok(L.msfem_nn_im2col(self.cin, B, H, W, self.k, self.k, self.pad, x.ptr, col.ptr, st))
        _ok(L.msfem_nn_gemm_bf16(0, 0, self.cout, N, self.K, self.wbf.ptr, self.K, col.ptr, N, y.ptr, N, 0, st))
        _ok(L.msfem_nn_bias_act(self.cout, N, self.b.ptr, y.ptr, 1 if self.relu else 0, st))
        self.saved = (x, B, H, W, y)
        return y, Ho, Wo

    def backward(self, dy, buf, want_dx=True):
        """dy: gradient w.r.t. the layer output (modified in place).  Returns dx or None."""
        L, st = _nn(), self.ctx.stream
        x, B, H, W, y = self.saved
        Ho, Wo = self.out_hw(H, W)
        N = B * Ho * Wo
        dybf = buf.get("dybf", self.cout * N, np.uint16)
        _ok(L.msfem_nn_act_bwd(self.cout * N, y.ptr, dy.ptr, dybf.ptr, 1 if self.relu else 0, st))
        _ok(L.msfem_nn_bias_grad(self.cout, N, dy.ptr, self.gb.ptr, st))
        col = buf.get("col", self.K * N, np.uint16)
        _ok(L.msfem_nn_im2col(self.cin, B, H, W, self.k, self.k, self.pad, x.ptr, col.ptr, st))
        _ok(L.msfem_nn_gemm_bf16(0, 1, self.cout, self.K, N, dybf.ptr, N, col.ptr, N, self.gw.ptr, self.K, 0, st))
        if not want_dx:
            return None
        dcol = buf.get("dcol", self.K * N, np.float32)
        _ok(L.msfem_nn_gemm_bf16(1, 0, self.K, N, self.cout, self.wbf.ptr, self.K, dybf.ptr, N, dcol.ptr, N, 0, st))
        dx = buf.get(self.name + ".dx", self.cin * B * H * W, np.float32)
        _ok(L.msfem_nn_col2im(self.cin, B, H, W, self.k, self.k, self.pad, dcol.ptr, dx.ptr, st))
        return dx

    def adam(self, lr, step, wd=0.0, b1=0.9, b2=0.999, eps=1e-8):
        L, st = _nn(), self.ctx.stream
        _ok(L.msfem_nn_adam(self.cout * self.K, self.w.ptr, self.gw.ptr, self.mw.ptr, self.vw.ptr, self.wbf.ptr, lr, b1,
                            b2, eps, step, wd, st))
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
        self.layers = self.enc + self.dec + [self.head]
        self.buf = Buffers(ctx)
        self.step_count = 0

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
    def forward(self, x, B):
        """x: DeviceArray [1][B][tile][tile] FP32 (normalized log-coefficient).
        Returns the head output [4][B][tile-1][tile-1] (DeviceArray view)."""
        L, st, buf = _nn(), self.ctx.stream, self.buf
        H = self.tile
        h, _, _ = self.enc[0].forward(x, B, H, H, buf)
        skips = []
        h, _, _ = self.enc[1].forward(h, B, H, H, buf)
        skips.append((h, H))
        for l in range(1, self.levels):
            C = self.chans[l - 1]
            p = buf.get(f"pool{l}", C * B * (H // 2) ** 2, np.float32)
            _ok(L.msfem_nn_pool2(C * B, H, H, h.ptr, p.ptr, 0, st))
            H //= 2
            h, _, _ = self.enc[l + 1].forward(p, B, H, H, buf)
            skips.append((h, H))
        self.skips = skips
        for l in range(self.levels - 2, -1, -1):
            C = self.chans[l + 1]
            u = buf.get(f"up{l}", C * B * (2 * H) ** 2, np.float32)
            _ok(L.msfem_nn_up2(C * B, H, H, h.ptr, u.ptr, 0, st))
            H *= 2
            y, _, _ = self.dec[l].forward(u, B, H, H, buf)
            sk, _ = skips[l]
            # the sum goes to its own buffer: y (post-ReLU) is kept for the backward mask
            h = buf.get(f"sum{l}", self.chans[l] * B * H * H, np.float32)
            _ok(L.msfem_nn_axpby(self.chans[l] * B * H * H, 1.0, y.ptr, 1.0, sk.ptr, h.ptr, st))
        out, _, _ = self.head.forward(h, B, H, H, buf)
        self.B = B
        return out

    def backward(self, dout):
        """dout: gradient of the loss w.r.t. the head output (consumed)."""
        L, st, buf = _nn(), self.ctx.stream, self.buf
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
        L, st = _nn(), self.ctx.stream
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

    def save(self, path):
        np.savez(path, **self.state(), chans=np.array(self.chans), tile=self.tile)

    @staticmethod
    def load(ctx, path):
        d = np.load(path)
        s = Surrogate(ctx, int(d["tile"]), tuple(int(v) for v in d["chans"]))
        s.load_state({k: d[k] for k in d.files if k.endswith(".w") or k.endswith(".b")})
        return s

    # ---- serving ------------------------------------------------------------------------
    def predict_bases(self, g_tiles, bases_out, batch=128):
        """g_tiles: host float32 [n][tile*tile] log-coefficients; bases_out: device
        FP64 bases buffer [n][4][(tile+1)^2] (the x0 of msfem_batch_build_device)."""
        L, st = _nn(), self.ctx.stream
        n = g_tiles.shape[0]
        t = self.tile
        xin = self.buf.get("serve_x", batch * t * t, np.float32)
        nbase = 4 * (t + 1) ** 2
        for s in range(0, n, batch):
            b = min(batch, n - s)
            xin.upload(np.concatenate([(g_tiles[s:s + b] / CLIP).astype(np.float32).reshape(-1),
                                       np.zeros((batch - b) * t * t, np.float32)]))
            out = self.forward(xin, batch)
            # head output is [4][batch][(t-1)^2]; the first b samples of each channel are used
            if b < batch:
                packed = self.buf.get("serve_pack", 4 * b * (t - 1) ** 2, np.float32)
                host = out.numpy()[:4 * batch * (t - 1) ** 2].reshape(4, batch, -1)[:, :b]
                packed.upload(np.ascontiguousarray(host).reshape(-1))
                src = packed
            else:
                src = out
            _ok(L.msfem_nn_delta_to_bases(b, t, src.ptr, bases_out.ptr + s * nbase * 8, st))
        self.ctx.synchronize()
