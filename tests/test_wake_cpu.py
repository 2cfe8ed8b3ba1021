"""Wakefield set-up on the host (g4mi_wake_loss: Wake.cpp's single-particle
wakes, Collective.cpp's convolution), pinned against what the physics and
the convolution's structure fix: the resistive wall's short- and
long-range limits (Bane & Stupakov, SLAC-PUB-10707), the geometric wake's
sqrt(s) law, causality, the self-loading half bin and linearity in the
current.  No device is needed."""
import math

import numpy as np
import pytest

import g4mi

E = 1.602176634e-19
C0 = 299792458.0
Z0 = 376.73
CE = 4.8032045e-11


def _dc_round_wake(sigma):
    """Bane & Stupakov's closed form of the DC resistive-wall wake of a round
    pipe, W(s) / W(0+) with sigma = s / s0 (SLAC-PUB-10707, Eq. 6), its
    integral by scipy's adaptive quadrature."""
    from scipy.integrate import quad
    integral = quad(lambda x: x * x * math.exp(-x * x * sigma) / (x ** 6 + 8), 0, np.inf)[0]
    return 4 * (math.exp(-sigma) / 3 * math.cos(math.sqrt(3) * sigma) - math.sqrt(2) / math.pi * integral)


def test_resistive_round_dc_wake():
    """Round copper pipe, a = 2.5 mm, DC conductivity: w(0+) = -e Z0 c /
    (pi a^2) and the whole short-range wake against the closed form, within
    2 % of w(0+) for s up to 20 s0 (the reference's quadrature, k up to
    100 / s0 on 1000 intervals, repeats with period 62.8 s0)."""
    a, sig = 2.5e-3, 5.813e7
    s0 = (2 * a * a / (Z0 * sig)) ** (1 / 3)
    ds = s0 / 20
    _, w = g4mi.wake_loss(np.zeros(401), ds, conductivity=sig, radius=a)
    wr = w["resistive"]
    w0 = -E * Z0 * C0 / (math.pi * a * a)
    assert abs(wr[0] / w0 - 1) < 0.02
    for i in (10, 20, 40, 100, 200, 400):
        assert abs(wr[i] / w0 - _dc_round_wake(i * ds / s0)) < 0.02, i


def test_geometric_sqrt_law_and_flat_factor():
    ds, a, g, lg = 1e-7, 3e-3, 1e-2, 0.5
    _, wround = g4mi.wake_loss(np.zeros(64), ds, radius=a, gap=g, lgap=lg)
    _, wflat = g4mi.wake_loss(np.zeros(64), ds, radius=a, gap=g, lgap=lg, roundpipe=False)
    s = ds * np.arange(64)
    ref = -Z0 * CE / (math.pi ** 2 * a * lg) * 2 * math.sqrt(0.5 * g) * np.sqrt(s)
    np.testing.assert_allclose(wround["geometric"], ref, rtol=1e-14, atol=0)
    np.testing.assert_allclose(wflat["geometric"], 0.956 * ref, rtol=1e-14, atol=0)


@pytest.mark.parametrize("kind", ["resistive", "geometric", "roughness"])
def test_convolution_causality_linearity_self_loading(kind):
    """A current in one slice k only: slices ahead of it (index > k) feel
    nothing, slice k feels half of its own zero-distance wake, the loss is
    linear in the current and the external loss adds on."""
    ns, ds = 32, 2e-7
    kw = dict(resistive=dict(conductivity=5.813e7), geometric=dict(gap=1e-2, lgap=0.5),
              roughness=dict(hrough=1e-7, lrough=1e-4))[kind]
    cur = np.zeros(ns)
    cur[20] = 1000.0
    e1, w = g4mi.wake_loss(cur, ds, **kw)
    e2, _ = g4mi.wake_loss(2 * cur, ds, loss=-5e3, **kw)
    assert np.all(e1[21:] == 0)
    np.testing.assert_allclose(e2 + 5e3, 2 * e1, rtol=1e-13, atol=1e-9)
    if kind != "geometric":  # the geometric wake is driven by dI/ds
        n = 1000.0 * ds / CE  # electrons in the bin
        wk = w[kind]
        assert e1[20] == pytest.approx(n * 0.5 * wk[0], rel=1e-12)
        assert e1[19] == pytest.approx(n * wk[1], rel=1e-12)
