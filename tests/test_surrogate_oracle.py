"""CPU checks of the surrogate network's FP64 oracle (oracle/surrogate_ref.py).

The oracle's backward pass is pinned by central finite differences of its own
forward pass with BF16 rounding switched off (the rounding makes the forward
piecewise constant), on a tiny net where every layer type appears: 3x3
padded conv + ReLU, 2x2 average pool, nearest upsampling, skip sums, and the
2x2 valid head.  A transposed weight, a dropped skip or a wrong pooling scale
each move at least one gradient entry by O(1).
"""
import numpy as np
import pytest

import surrogate_ref as R


def tiny_params(rng, chans, cin=1, out_ch=4):
    p = {}

    def conv(name, ci, co, k):
        p[name] = (rng.standard_normal((co, ci * k * k)) * np.sqrt(2.0 / (ci * k * k)), rng.standard_normal(co) * 0.1)

    conv("e0a", cin, chans[0], 3)
    conv("e0", chans[0], chans[0], 3)
    for l in range(1, len(chans)):
        conv(f"e{l}", chans[l - 1], chans[l], 3)
    for l in range(len(chans) - 1):
        conv(f"d{l}", chans[l + 1], chans[l], 3)
    conv("head", chans[0], out_ch, 2)
    return p


@pytest.fixture
def no_rounding(monkeypatch):
    monkeypatch.setattr(R, "bf16", lambda a: np.asarray(a, np.float64))


def test_bf16_rounding_known_values():
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties go to the even mantissa (1.0);
    # 1 + 3*2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: goes up to 1 + 2^-6.
    x = np.array([1.0, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, 1 + 2.0 ** -9, 1 + 2.0 ** -8 + 2.0 ** -12, -2.5, 0.0])
    want = np.array([1.0, 1.0, 1 + 2.0 ** -6, 1.0, 1 + 2.0 ** -7, -2.5, 0.0])
    assert np.array_equal(R.bf16(x), want)


def test_im2col_col2im_adjoint():
    """<im2col(x), c> == <x, col2im(c)> for random x, c (col2im is the adjoint)."""
    rng = np.random.default_rng(1)
    for (C, B, H, W, k, pad) in [(2, 3, 5, 6, 3, 1), (3, 2, 4, 4, 2, 0)]:
        x = rng.standard_normal((C, B, H, W))
        col, (Ho, Wo) = R.im2col(x, k, pad)
        c = rng.standard_normal(col.shape)
        assert abs((col * c).sum() - (x * R.col2im(c, x.shape, k, pad)).sum()) < 1e-10


def test_conv_matches_direct_convolution():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 1, 5, 4))
    w = rng.standard_normal((3, 2 * 9))
    b = rng.standard_normal(3)
    conv = R.Conv(w, b, 3, 1, relu=False)
    with pytest.MonkeyPatch().context() as m:
        m.setattr(R, "bf16", lambda a: np.asarray(a, np.float64))
        y = conv.forward(x)
    xp = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1)))
    w4 = w.reshape(3, 2, 3, 3)
    for co in range(3):
        for yy in range(5):
            for xx in range(4):
                ref = b[co] + (w4[co] * xp[:, 0, yy:yy + 3, xx:xx + 3]).sum()
                assert abs(y[co, 0, yy, xx] - ref) < 1e-12


def test_pool_up_adjoints():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 3, 4, 6))
    d = rng.standard_normal((2, 3, 2, 3))
    assert abs((R.pool2(x) * d).sum() - (x * R.pool2_bwd(d)).sum()) < 1e-12
    assert abs((R.up2(d) * x).sum() - (d * R.up2_bwd(x)).sum()) < 1e-12
    assert np.allclose(R.pool2(np.ones((1, 1, 4, 4))), 1.0)


def test_net_gradient_finite_differences(no_rounding):
    rng = np.random.default_rng(4)
    chans = (3, 4, 5)
    p = tiny_params(rng, chans)
    x = rng.standard_normal((1, 2, 8, 8))
    target = rng.standard_normal((4, 2, 7, 7))

    def loss_of(params):
        net = R.Net(params, chans)
        return R.loss_and_grad(net.forward(x), target)[0]

    net = R.Net(p, chans)
    out = net.forward(x)
    _, dout = R.loss_and_grad(out, target)
    net.backward(dout)
    g = net.grads()
    h = 1e-6
    for name in ("e0a", "e0", "e1", "e2", "d0", "d1", "head"):
        w, b = p[name]
        for (arr, gi, idxs) in ((w, 0, [(0, 0), (w.shape[0] - 1, w.shape[1] - 1), (1, w.shape[1] // 2)]),
                                (b, 1, [(0,), (b.shape[0] - 1,)])):
            for idx in idxs:
                q = {k: (v[0].copy(), v[1].copy()) for k, v in p.items()}
                q[name][gi][idx] += h
                lp = loss_of(q)
                q[name][gi][idx] -= 2 * h
                lm = loss_of(q)
                fd = (lp - lm) / (2 * h)
                an = g[name][gi][idx]
                assert abs(fd - an) <= 1e-6 + 1e-5 * abs(fd), (name, gi, idx, fd, an)


def test_energy_norm_matches_oracle_stiffness():
    """surrogate.energy_norm2 (the serving error metric) equals v^T K v with the
    oracle's assembled fine stiffness, for a random field and random nodal vectors."""
    import oracle as O
    import surrogate as S
    rng = np.random.default_rng(11)
    nf = 8
    alpha = np.exp(rng.standard_normal((nf, nf)))
    box = np.array([[0.0, 0.25], [0.0, 0.25]])
    K, _ = O.local_assemble(2, nf, box, O.COEF_SCALAR_CELL, alpha.reshape(-1))
    v = rng.standard_normal((3, nf + 1, nf + 1))
    got = S.energy_norm2(alpha, v)
    for i in range(3):
        want = v[i].reshape(-1) @ O.stencil_apply(2, nf, K, v[i].reshape(-1))
        assert abs(got[i] - want) < 1e-12 * abs(want)
    # the Q1 nodal interpolant of a linear function has energy alpha |grad|^2 integrated
    x = np.linspace(0, 1, nf + 1)
    lin = np.broadcast_to(2.0 * x[None, :], (nf + 1, nf + 1))
    assert abs(S.energy_norm2(np.ones((nf, nf)), lin) - 4.0) < 1e-12
