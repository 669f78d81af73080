"""Pins the oracle's spectral layer (filter_fftw.f90 restated on numpy FFTs)
with closed forms: derivatives of trigonometric modes, 3/2-grid interpolation
of band-limited fields, the band truncation of compress, the Nyquist removal
of deal and the forward normalisation of ffthrz."""
import math

import numpy as np
import pytest

import cf_oracle as O

NX, NY = 32, 16
XL, YL = 2 * math.pi, math.pi


def grid(nx, ny):
    x = np.arange(nx) * XL / nx
    y = np.arange(ny) * YL / ny
    return np.meshgrid(y, x, indexing="ij")  # Y, X with shape (ny, nx)


@pytest.mark.parametrize("m", [0, 1, 3, NX // 2 - 1])
def test_xderiv_cosine(m):
    Y, X = grid(NX, NY)
    k = m * 2 * math.pi / XL
    f = np.cos(k * X + 0.3) * (1 + np.cos(2 * Y * 2 * math.pi / YL))
    expect = -k * np.sin(k * X + 0.3) * (1 + np.cos(2 * Y * 2 * math.pi / YL))
    assert np.allclose(O.xderiv(f, XL), expect, atol=1e-12)


def test_xderiv_drops_nyquist():
    Y, X = grid(NX, NY)
    f = np.cos(NX // 2 * 2 * math.pi / XL * X)
    assert np.allclose(O.xderiv(f, XL), 0.0, atol=1e-13)


@pytest.mark.parametrize("m", [1, 2, NY // 2 - 1])
def test_yderiv_sine(m):
    Y, X = grid(NX, NY)
    k = m * 2 * math.pi / YL
    f = np.sin(k * Y) * np.cos(X)
    assert np.allclose(O.yderiv(f, YL), k * np.cos(k * Y) * np.cos(X), atol=1e-12)


def trig(X, Y, modes):
    out = np.zeros_like(X)
    for (mx, my, a, ph) in modes:
        out += a * np.cos(mx * 2 * math.pi / XL * X + my * 2 * math.pi / YL * Y + ph)
    return out


MODES = [(0, 0, 1.5, 0.0), (1, 0, 0.7, 0.2), (3, -2, 0.4, 1.0), (NX // 2 - 1, NY // 2 - 1, 0.3, 0.5),
         (2, -(NY // 2 - 1), 0.2, -0.4), (0, 5, 0.1, 0.9)]


def test_expand_interpolates_band_limited_fields():
    Ys, Xs = grid(NX, NY)
    Yb, Xb = grid(3 * NX // 2, 3 * NY // 2)
    small = trig(Xs, Ys, MODES)
    big = O.expand(small, 3 * NX // 2, 3 * NY // 2)
    assert np.allclose(big, trig(Xb, Yb, MODES), atol=1e-12)


def test_compress_truncates_to_band():
    Yb, Xb = grid(3 * NX // 2, 3 * NY // 2)
    Ys, Xs = grid(NX, NY)
    out_of_band = [(NX // 2 + 2, 1, 0.5, 0.0), (3, NY // 2 + 1, 0.25, 0.3), (NX // 2, 0, 0.6, 0.0)]
    big = trig(Xb, Yb, MODES + out_of_band)
    assert np.allclose(O.compress(big, NX, NY), trig(Xs, Ys, MODES), atol=1e-12)


def test_compress_expand_identity_on_dealiased_fields():
    rng = np.random.default_rng(3)
    a = O.deal(rng.standard_normal((4, NY, NX)))
    b = O.compress(O.expand(a, 3 * NX // 2, 3 * NY // 2), NX, NY)
    assert np.allclose(a, b, atol=1e-13)


def test_deal_removes_only_nyquist():
    Ys, Xs = grid(NX, NY)
    keep = trig(Xs, Ys, MODES)
    nyq = np.cos(NX // 2 * 2 * math.pi / XL * Xs) + np.cos(NY // 2 * 2 * math.pi / YL * Ys)
    assert np.allclose(O.deal(keep + nyq), keep, atol=1e-12)


def test_ffthrz_forward_normalisation():
    c = O.ffthrz_fwd(np.full((NY, NX), 2.5))
    assert abs(c[0, 0] - 2.5) < 1e-14
    assert np.abs(c).sum() - 2.5 < 1e-12
    Ys, Xs = grid(NX, NY)
    c = O.ffthrz_fwd(np.cos(3 * 2 * math.pi / XL * Xs))
    assert abs(c[0, 3] - 0.5) < 1e-14   # exp(-i k x) convention, scaled by 1/(n1 n2)
