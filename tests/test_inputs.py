"""Synthetic-input generators of the product library (host code, no GPU needed).

Philox-4x32-10 is pinned by the known-answer vectors published with Random123
(Salmon et al., SC'11); the reference coefficient by the oracle's long-double
restatement of MatrixCoeff (include/coefficients/matrix_coeff.tpp:17-91).
"""
import numpy as np
import pytest

import msfem
import oracle as O


@pytest.mark.parametrize("ctr,key,out", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(ctr, key, out):
    assert list(msfem.philox4x32_10(ctr, key)) == out


def test_unit_boxes():
    b = msfem.unit_boxes(2, 4)
    assert b.shape == (16, 2, 2)
    assert np.allclose(b[5], [[0.25, 0.5], [0.25, 0.5]])  # cell (1, 1)
    b3 = msfem.unit_boxes(3, 2)
    assert np.allclose(b3[6], [[0, 0.5], [0.5, 1], [0.5, 1]])  # cell (0, 1, 1)


@pytest.mark.parametrize("dim", [2, 3])
def test_reference_coefficient_matches_long_double(dim):
    nf = [8, 16]
    box = msfem.unit_boxes(dim, 2)[:2]
    c = msfem.coef_reference(dim, nf, box)
    ns = 3 if dim == 2 else 6
    nq = 1 << dim
    gs = [0.5 - 0.5 / np.sqrt(3), 0.5 + 0.5 / np.sqrt(3)]
    off = 0
    for cell in range(2):
        n = nf[cell]
        h = (box[cell][:, 1] - box[cell][:, 0]) / n
        for f in range(n ** dim):
            fi = [f % n, (f // n) % n, f // (n * n)][:dim]
            for q in range(nq):
                p = [box[cell][k][0] + (box[cell][k][1] - box[cell][k][0]) * fi[k] / n + h[k] * gs[(q >> k) & 1]
                     for k in range(dim)]
                want = O.matrix_coeff_ref(dim, p)
                assert np.abs(c[off:off + ns] - want).max() < 1e-12
                off += ns
    assert off == c.size


PI_D = 3.14592653509793218403  # include/coefficients/coefficients.h:21 (not pi)
GAUSS0 = 0.5 - 0.5 / np.sqrt(3)  # first QGauss(2) abscissa on [0, 1]


def special_point(kind, m):
    """x with 2 PI_D 57 x = theta exactly (up to one rounding of x), where
    sin(theta) is known: kind 'one' -> theta = pi/2 + 2 pi m (sin = 1),
    'minus' -> 3 pi/2 + 2 pi m (sin = -1), 'zero' -> pi m (sin = 0)."""
    theta = {"one": np.pi / 2 + 2 * np.pi * m, "minus": 1.5 * np.pi + 2 * np.pi * m, "zero": np.pi * m}[kind]
    return theta / (2 * PI_D * 57), {"one": 1.0, "minus": -1.0, "zero": 0.0}[kind]


@pytest.mark.parametrize("dim", [2, 3])
def test_reference_coefficient_value_pin(dim):
    """MatrixCoeff (include/coefficients/matrix_coeff.tpp:44-91, k = 57,
    scale 0.9999, PI_D = 3.14592..., matrix_coeff.hpp:45-49) at points where
    the oscillation's sines have exact values: the tensor is d I with
    d = 1 - 0.9999 (sx + sy) / 2 (the rotation R d I R^T only rounds), z
    ignored in 3D.  The expected values need no restatement of the formula:
    pi in place of PI_D moves these points' sines by 7e-4 to 5e-2, k = 56 by
    more than 0.1, scale 0.999 moves d by at least 1e-4; the bar is 1e-12.  Both the
    product generator and the oracle's long-double restatement are checked."""
    ns = 3 if dim == 2 else 6
    h = 1e-3
    diag = [0, 2] if dim == 2 else [0, 3, 5]
    off = [1] if dim == 2 else [1, 2, 4]
    cases = [("one", 10, "one", 3), ("minus", 7, "minus", 12), ("zero", 9, "one", 10), ("one", 4, "minus", 5),
             ("zero", 12, "zero", 3), ("minus", 11, "zero", 1)]
    for kx, mx, ky, my in cases:
        x, sx = special_point(kx, mx)
        y, sy = special_point(ky, my)
        want = 1.0 - 0.9999 * (0.5 * sx + 0.5 * sy)
        lo = [x - h * GAUSS0, y - h * GAUSS0, 0.3][:dim]
        box = np.array([[[l, l + h] for l in lo]])
        c = msfem.coef_reference(dim, [1], box)[:ns]  # first quadrature point of the one fine cell
        p = [lo[k] + h * GAUSS0 for k in range(dim)]
        for got in (c, O.matrix_coeff_ref(dim, p)):
            assert np.abs(np.asarray(got)[diag] - want).max() < 1e-12, (kx, mx, ky, my, got, want)
            assert np.abs(np.asarray(got)[off]).max() < 1e-15
    # the sines' extremes bound the spectrum to [1e-4, 1.9999]: contrast ~2e4
    x1, _ = special_point("one", 3)
    xm, _ = special_point("minus", 3)
    for (px, py), want in (((x1, x1), 1e-4), ((xm, xm), 1.9999)):
        box = np.array([[[px - h * GAUSS0, px - h * GAUSS0 + h], [py - h * GAUSS0, py - h * GAUSS0 + h]]])
        assert abs(msfem.coef_reference(2, [1], box)[0] - want) < 1e-12


def test_reference_coefficient_spectrum_bounds():
    """Every quadrature value of MatrixCoeff has its eigenvalues in
    [1e-4, 1.9999] and is isotropic (the rotation by PI_D/3 of the equal
    diagonal entries)."""
    c = msfem.coef_reference(2, [16], msfem.unit_boxes(2, 1))
    A = c.reshape(-1, 3)
    tr = A[:, 0] + A[:, 2]
    assert tr.min() >= 2 * 1e-4 - 1e-12 and tr.max() <= 2 * 1.9999 + 1e-12
    # the rotation by PI_D/3 mixes the (equal) diagonal entries into an isotropic tensor
    assert np.abs(A[:, 1]).max() < 1e-12
    assert np.allclose(A[:, 0], A[:, 2])


def test_periodic_coefficient_formula():
    box = msfem.unit_boxes(2, 8)[:1]
    c = msfem.coef_periodic([32], box, 1 / 64).reshape(32 * 32, 4)
    gs = [0.5 - 0.5 / np.sqrt(3), 0.5 + 0.5 / np.sqrt(3)]
    h = 0.125 / 32
    for f in (0, 31, 500):
        for q in range(4):
            x = (f % 32) * h + h * gs[q & 1]
            y = (f // 32) * h + h * gs[q >> 1]
            assert abs(c[f, q] - (2 + np.sin(2 * np.pi * x * 64) * np.cos(2 * np.pi * y * 64))) < 1e-12


def test_lognormal_statistics_and_clip():
    clip = 0.5 * np.log(1e3)
    a = msfem.coef_lognormal(2, [64] * 64, 2, clip)
    g = np.log(a)
    assert g.size == 64 * 64 * 64
    assert g.min() >= -clip - 1e-12 and g.max() <= clip + 1e-12
    assert a.max() / a.min() <= 1e3 * (1 + 1e-12)
    assert abs(g.mean()) < 0.01 and abs(g.std() - 0.99) < 0.02  # N(0,1) clipped at 3.45 sigma
    assert np.array_equal(a, msfem.coef_lognormal(2, [64] * 64, 2, clip))  # deterministic
    assert not np.array_equal(a, msfem.coef_lognormal(2, [64] * 64, 3, clip))


def test_tensor_field_is_spd_with_lognormal_eigenvalues():
    t = msfem.coef_tensor([16] * 8, 3, 1.5).reshape(-1, 3)
    det = t[:, 0] * t[:, 2] - t[:, 1] ** 2
    assert (t[:, 0] > 0).all() and (det > 0).all()
    tr = t[:, 0] + t[:, 2]
    lmax = 0.5 * (tr + np.sqrt(tr ** 2 - 4 * det))
    lmin = det / lmax
    logs = np.log(np.concatenate([lmax, lmin]))
    assert abs(logs.std() - 1.5) < 0.1


def test_resolution_classes_quartiles():
    nf, con = msfem.resolution_classes(32, 3, 1.5)
    vals, counts = np.unique(nf, return_counts=True)
    assert list(vals) == [16, 32, 64, 128]
    assert np.abs(counts - 256).max() <= 2
    for a, b in zip(vals[:-1], vals[1:]):  # higher contrast, finer grid
        assert con[nf == a].max() <= con[nf == b].min()


def test_grf_unit_variance_and_correlation():
    g = msfem.grf_log(64, 64, 40, 4.0, 10.0).astype(np.float64).reshape(64, 64, 64)
    assert abs(g.var() - 1) < 0.1
    # Gaussian covariance exp(-r^2 / (2 l^2)), l = 4: lag 4 -> e^-0.5, lag 8 -> e^-2
    c4 = np.mean(g[:, :, :-4] * g[:, :, 4:]) / g.var()
    c8 = np.mean(g[:, :, :-8] * g[:, :, 8:]) / g.var()
    assert abs(c4 - np.exp(-0.5)) < 0.08 and abs(c8 - np.exp(-2)) < 0.08
