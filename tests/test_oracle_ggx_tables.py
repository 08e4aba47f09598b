"""The GGX energy tables of the oracle against an independent FP64 quadrature.

The BSDF's energy compensation reads two baked tables
(crates/vork-render/shaders/environment.slang:78-114 and
openpbr_core.slang:154-166): per (cosine, roughness) texel the directional
albedo of the height-correlated Smith GGX lobe split by the Schlick weight,
(1 - (1 - vh)^5), (1 - vh)^5 and vh (1 - vh)^6, plus its cosine-weighted
average over the hemisphere, and per (cosine, roughness, eta) the albedo
weighted by the exact dielectric Fresnel.  The reference bakes them by
VNDF-sampled quadrature with 512 points in FP32, as the oracle does.

Here each integral is computed a second, independent way: over half vectors
with the GGX distribution's own CDF as the variable (tan^2 theta =
alpha^2 xi / (1 - xi)), the azimuth restricted in closed form to where
v.h > 0 and n.l > 0, and Gauss-Legendre in both variables in FP64:

    E[w] = 1/(2 pi) int int G2(v, l) (v.h) / (n.v cos theta_h) w(v.h) dxi dphi

The bound is the 512-point bake's own quadrature error, about 2e-3 for the
first two Schlick moments, 1e-4 for the third (5e-3 for Fresnel below total
internal reflection); the last
test shows that plausible slips (an exponent off by one, G2 in place of
G2/G1, a missing half in the Fresnel average) miss by far more.
"""
import ctypes as C

import numpy as np
import pytest

import oracle

N = 256


def _integrate(nv, alpha, weight, g1_division=True):
    x, w = np.polynomial.legendre.leggauss(N)
    xi, wxi = 0.5 * (x + 1), 0.5 * w
    a2 = alpha * alpha
    t2 = a2 * xi / (1 - xi)
    ch = 1 / np.sqrt(1 + t2)
    sh = np.sqrt(t2) * ch
    st = np.sqrt(max(1 - nv * nv, 0.0))
    den = ch * sh * st
    with np.errstate(divide="ignore", invalid="ignore"):
        c_light = np.where(den > 0, nv * (1 - 2 * ch * ch) / (2 * den), np.where(1 - 2 * ch * ch < 0, -np.inf, np.inf))
        c_view = np.where(den > 0, -nv * ch / (sh * st), -np.inf)
    phi0 = np.arccos(np.clip(np.maximum(c_light, c_view), -1, 1))
    phi = 0.5 * (x[None, :] + 1) * phi0[:, None]
    wphi = 0.5 * w[None, :] * phi0[:, None]
    vh = st * sh[:, None] * np.cos(phi) + nv * ch[:, None]
    nl = np.maximum(2 * vh * ch[:, None] - nv, 0)
    g2 = 2 * nv * nl / (nl * np.sqrt(a2 + (1 - a2) * nv * nv) + nv * np.sqrt(a2 + (1 - a2) * nl * nl))
    if not g1_division:   # a slip: the VNDF weight without dividing by G1(v)
        g2 = g2 * 2 * nv / (nv + np.sqrt(a2 + (1 - a2) * nv * nv))
    f = g2 * vh / (nv * ch[:, None]) * weight(np.clip(vh, 0, 1))
    return float(((f * wphi).sum(1) * 2 * wxi).sum() / (2 * np.pi))


def _fresnel(eta, half=0.5):
    def f(c):
        if eta <= 0:
            return np.ones_like(c)
        si2 = (1 - c * c) / (eta * eta)
        ct = np.sqrt(np.maximum(0, 1 - si2))
        rs, rp = (c - eta * ct) / (c + eta * ct), (eta * c - ct) / (eta * c + ct)
        return np.where(si2 >= 1, 1.0, (rs * rs + rp * rp) * half)
    return f


def _cosine(col):
    t = np.float32(col) / np.float32(127)
    return float(np.float32(t * t) * np.float32(t * t))


def _alpha(row):
    if row == 0:
        return float(np.float32(1e-4))
    if row == 127:
        return 1.0
    return float(np.float32(2 ** (np.log2(np.float32(1e-4)) * (1 - row / 127.0))))


def _bake(col, row):
    out = np.zeros(4, np.float32)
    oracle.lib().vo_bake_moments_texel(col, row, oracle.ptr(out))
    return out


SCHLICK = (lambda m: 1 - (1 - m) ** 5, lambda m: (1 - m) ** 5, lambda m: m * (1 - m) ** 6)
BOUND = np.array([3e-3, 3e-3, 3e-4])   # the third moment is an order of magnitude smaller


@pytest.mark.parametrize("row", [0, 8, 20, 40, 64, 90, 110, 127])
def test_schlick_moments(row):
    for col in (1, 4, 16, 40, 64, 100, 127):
        nv, alpha = _cosine(col), _alpha(row)
        b = _bake(col, row)
        want = [_integrate(nv, alpha, w) for w in SCHLICK]
        assert (np.abs(b[:3] - want) < BOUND).all(), (row, col, b[:3], want)


def test_hemispherical_average():
    x, w = np.polynomial.legendre.leggauss(64)
    t, wt = 0.5 * (x + 1), 0.5 * w
    nv = t ** 4   # the table's own cosine spacing, dense near grazing
    for row in (0, 40, 64, 90, 127):
        alpha = _alpha(row)
        e = np.array([_integrate(max(v, 1e-12), alpha, np.ones_like) for v in nv])
        avg = float((2 * nv * e * 4 * t ** 3 * wt).sum())
        assert abs(_bake(0, row)[3] - avg) < 1e-3, (row, _bake(0, row)[3], avg)


@pytest.mark.parametrize("eta", [0.0, 0.5, 1.5, 3.0])
def test_dielectric_albedo(eta):
    L = oracle.lib()
    for r in (0.1, 0.3, 0.7, 1.0):
        for no in (0.05, 0.3, 0.8, 1.0):
            b = L.vo_dielectric_directional_albedo(C.c_float(no), C.c_float(r), C.c_float(eta), 512)
            want = _integrate(float(np.float32(no)), float(np.float32(max(r * r, 1e-4))), _fresnel(eta))
            assert abs(b - want) < 8e-3, (eta, r, no, b, want)


def test_the_bound_rejects_plausible_slips():
    # each slip moves at least one checked texel far beyond the bounds above
    nv, alpha = _cosine(64), _alpha(110)
    b = _bake(64, 110)
    assert abs(b[2] - _integrate(nv, alpha, lambda m: m * (1 - m) ** 5)) > 10 * BOUND[2]
    assert abs(b[1] - _integrate(nv, alpha, lambda m: (1 - m) ** 4)) > 10 * BOUND[1]
    nv, alpha = _cosine(100), _alpha(127)
    assert abs(_bake(100, 127)[0] - _integrate(nv, alpha, SCHLICK[0], g1_division=False)) > 3e-2
    d = oracle.lib().vo_dielectric_directional_albedo(C.c_float(0.8), C.c_float(0.3), C.c_float(1.5), 512)
    assert abs(d - _integrate(0.8, 0.09, _fresnel(1.5, half=1.0))) > 3e-2
