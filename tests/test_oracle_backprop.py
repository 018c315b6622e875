"""Pins of the oracle's back-propagation (reference microscope.cu:423-489,
applyBackpropagation :592-605) by the physics it encodes: the DC sample
takes -exp(i k d); propagating samples keep their modulus and the others
vanish, and the propagating set is the disk |f| < 1/lambda; two
propagations compose to one over the summed distance (up to the factor -1
each carries); a single spatial frequency takes the closed-form phase
-exp(i sqrt(k^2 - kx^2) d); with the reference's frequency off-by-one the
last index maps to frequency 0."""
import math

import numpy as np
import pytest

PI_F = float(np.float32(3.14159))


def spectrum(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))


@pytest.mark.parametrize("fp32", [0, 1])
def test_dc_sample(orc, fp32):
    n, L, lam, d = 32, 40.0, 3.0, 1.7
    F = spectrum(n, 1)
    out = orc.backpropagate(F, 1.0 / L, 1.0 / L, lam, d, fp32=fp32)
    k = 2 * PI_F / lam
    want = -F[0, 0] * np.exp(1j * k * d)
    assert abs(out[0, 0] - want) <= 1e-5 * abs(F[0, 0])


@pytest.mark.parametrize("fp32", [0, 1])
@pytest.mark.parametrize("quirks", [0, 1])
def test_modulus_and_propagating_disk(orc, fp32, quirks):
    n, L, lam = 48, 30.0, 2.45
    F = spectrum(n, 2)
    out = orc.backpropagate(F, 1.0 / L, 1.0 / L, lam, 2.5, quirks=quirks, fp32=fp32)
    freq = np.array([i if i <= n // 2 else -((n - 1 - i) if quirks else (n - i))
                     for i in range(n)], dtype=np.float64) / L
    fu, fv = np.meshgrid(freq, freq)
    r2 = (fu ** 2 + fv ** 2) * lam ** 2
    inside = r2 < 1.0
    assert np.min(np.abs(r2 - 1.0)) > 1e-3   # no sample on the circle here
    assert np.allclose(np.abs(out[inside]), np.abs(F[inside]), rtol=1e-6)
    assert np.all(out[~inside] == 0)
    assert inside.sum() > 50 and (~inside).sum() > 50


@pytest.mark.parametrize("fp32", [0, 1])
def test_composition(orc, fp32):
    n, L, lam, d1, d2 = 32, 25.0, 2.0, 0.75, 1.5
    F = spectrum(n, 3)
    two = orc.backpropagate(orc.backpropagate(F, 1 / L, 1 / L, lam, d1, fp32=fp32), 1 / L, 1 / L,
                            lam, d2, fp32=fp32)
    one = orc.backpropagate(F, 1 / L, 1 / L, lam, d1 + d2, fp32=fp32)
    assert np.max(np.abs(two + one)) <= 2e-5 * np.max(np.abs(F))


def test_single_frequency_phase(orc):
    """A spectrum with one sample at (nu, 0): that sample times
    -exp(i sqrt(k^2 - (2 pi nu / L)^2) d)."""
    n, L, lam, d, nu = 64, 50.0, 4.0, 3.0, 7
    F = np.zeros((n, n), dtype=complex)
    F[0, nu] = 2.0 - 0.5j
    out = orc.backpropagate(F, 1.0 / L, 1.0 / L, lam, d)
    k = 2 * PI_F / lam
    kx = 2 * PI_F * nu / L
    want = -F[0, nu] * np.exp(1j * math.sqrt(k * k - kx * kx) * d)
    assert abs(out[0, nu] - want) <= 1e-5
    assert np.count_nonzero(out) == 1


def test_negative_frequency_off_by_one(orc):
    """With quirks the last index is frequency 0 (the DC factor); without,
    frequency -1, whose factor equals that of +1."""
    n, L, lam, d = 16, 20.0, 2.5, 1.1
    F = np.zeros((n, n), dtype=complex)
    F[0, 0] = F[0, 1] = F[0, n - 1] = 1.0
    q = orc.backpropagate(F, 1 / L, 1 / L, lam, d, quirks=1)
    assert q[0, n - 1] == q[0, 0]
    c = orc.backpropagate(F, 1 / L, 1 / L, lam, d, quirks=0)
    assert c[0, n - 1] == c[0, 1] and c[0, n - 1] != c[0, 0]
