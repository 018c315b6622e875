"""Pins of the oracle's Mie coefficients, incident field, scattered field,
field sum and detector against what the mathematics fixes, independently of
the reference's formulas:
  - Mie coefficients satisfy the scalar boundary conditions at r = a
    (continuity of the field and its radial derivative), evaluated with
    scipy's Bessel functions; a sphere of index 1 does not scatter;
  - the incident-field series equals the condenser-annulus integral of plane
    waves, 2 pi A int J0(k rho sin t) e^{i k z cos t} sin t dt (quadrature);
  - inside a sphere of index 1 the internal series is the plane-wave sum;
  - the detector passes or removes single spatial-frequency tones as their
    frequency lies inside or outside the objective annulus, including the
    reference's off-by-one negative-frequency index;
  - constant fields give closed-form intensities and absorbance.
The reference's own quirks (recurrence thresholds, 3.14159 scale, typo) are
turned off here (quirks=0) where they would break the identities; with
quirks on, the same routines are pinned against the reference's compiled
code in test_oracle_pins.py."""
import math

import numpy as np
import pytest
from scipy import integrate, special

PI_F = 3.1415901184082031


def sph_h(l, x):
    return special.spherical_jn(l, x) + 1j * special.spherical_yn(l, x)


def sph_hp(l, x):
    return (special.spherical_jn(l, x, derivative=True) +
            1j * special.spherical_yn(l, x, derivative=True))


def sph_j_complex(l, z):
    return special.spherical_jn(l, z)


def sph_jp_complex(l, z):
    return special.spherical_jn(l, z, derivative=True)


@pytest.mark.parametrize("a,n,lam", [(2.0, 1.33 + 0.0j, 3.0), (5.0, 1.4 + 0.05j, 2.5),
                                     (1.0, 1.6 + 0.01j, 4.0), (3.0, 1.2 + 0.2j, 2.0)])
def test_mie_coefficients_satisfy_boundary_conditions(orc, a, n, lam):
    kmag = 2.0 * PI_F / lam
    nl = orc.sphere_order(a, lam)
    A, B = orc.mie_coeffs(nl, kmag, a, n, quirks=0)
    l = np.arange(nl + 1)
    ka = kmag * a
    inc = (2 * l + 1) * (1j ** l)
    jin, jpin = sph_j_complex(l, n * ka), sph_jp_complex(l, n * ka)
    # field: A j_l(n k a) = (2l+1) i^l j_l(k a) + B h_l(k a)
    lhs0 = A * jin
    rhs0 = inc * special.spherical_jn(l, ka) + B * sph_h(l, ka)
    # radial derivative: A n j_l'(n k a) = (2l+1) i^l j_l'(k a) + B h_l'(k a)
    lhs1 = A * n * jpin
    rhs1 = inc * special.spherical_jn(l, ka, derivative=True) + B * sph_hp(l, ka)
    scale = np.abs(inc) + np.abs(A * jin) + np.abs(B * sph_h(l, ka))
    assert np.max(np.abs(lhs0 - rhs0) / scale) < 1e-9
    assert np.max(np.abs(lhs1 - rhs1) / scale) < 1e-9


@pytest.mark.parametrize("quirks", [0, 1])
def test_index_one_sphere_does_not_scatter(orc, quirks):
    nl, kmag, a = 20, 2.0 * PI_F / 2.5, 4.0
    A, B = orc.mie_coeffs(nl, kmag, a, 1.0 + 0.0j, quirks=quirks)
    l = np.arange(nl + 1)
    assert np.max(np.abs(B)) == 0.0
    assert np.max(np.abs(A - (2 * l + 1) * (1j ** l)) / (2 * l + 1)) < 1e-12


def annulus_field(rho, z, kmag, ca, cb, amp):
    """A * 2 pi * int_{cb}^{ca} J0(k rho sqrt(1-t^2)) exp(i k z t) dt."""
    def f(t, part):
        v = special.j0(kmag * rho * math.sqrt(max(0.0, 1 - t * t))) * np.exp(1j * kmag * z * t)
        return v.real if part == 0 else v.imag
    re = integrate.quad(f, cb, ca, args=(0,), epsabs=1e-13, epsrel=1e-12, limit=400)[0]
    im = integrate.quad(f, cb, ca, args=(1,), epsabs=1e-13, epsrel=1e-12, limit=400)[0]
    return amp * 2 * math.pi * (re + 1j * im)


@pytest.mark.parametrize("z,NAin,NAout", [(0.0, 0.0, 0.5), (1.5, 0.2, 0.62), (-2.0, 0.0, 0.9)])
def test_incident_series_is_the_annulus_integral(orc, z, NAin, NAout):
    import oracle
    lam = 2.5
    kmag = 2.0 * PI_F / lam
    N, L = 8, 16.0
    q = oracle.Quad.xy(L, z)
    ca, cb = math.cos(math.asin(NAin)), math.cos(math.asin(NAout))
    dmax = math.sqrt(2 * (L / 2) ** 2 + z * z)
    kd = kmag * dmax
    m = int(math.ceil(kd + 4 * kd ** (1 / 3) + 2)) + 20
    Uf = orc.nf_uf(q, N, N, (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), kmag, 1.3, ca, cb, m, quirks=0)
    want = np.zeros((N, N), dtype=complex)
    for iv in range(N):
        for iu in range(N):
            x, y = -L / 2 + L * iu / N, -L / 2 + L * iv / N
            want[iv, iu] = annulus_field(math.hypot(x, y), z, kmag, ca, cb, 1.3)
    # the series carries the reference's float PI in its 2*PI prefactor
    want *= PI_F / math.pi
    assert np.max(np.abs(Uf - want)) / np.max(np.abs(want)) < 1e-9


def test_internal_series_of_index_one_sphere_is_the_plane_wave_sum(orc):
    import oracle
    lam, a = 3.0, 6.0
    kmag = 2.0 * PI_F / lam
    N, L = 12, 20.0
    q = oracle.Quad.xy(L, 0.5)
    p = (1.0, -0.5, 0.0)
    nl = orc.sphere_order(a, lam) + 25
    A, B = orc.mie_coeffs(nl, kmag, a, 1.0 + 0.0j, quirks=0)
    s = dict(p=p, a=a, n=1.0 + 0.0j, Nl=nl, coeff_off=0)
    rng = np.random.default_rng(3)
    waves = rng.normal(size=(7, 3))
    waves[:, 2] = np.abs(waves[:, 2]) + 1.0
    waves /= np.linalg.norm(waves, axis=1)[:, None]
    subA = 0.8
    focus = (0.3, 0.2, -0.1)
    U = orc.nf_us(q, N, N, focus, kmag, [s], A, B, waves, subA, quirks=0)
    w1 = float(np.float32(1.0) / np.float32(len(waves)))
    iv, iu = np.mgrid[0:N, 0:N]
    P = np.stack([-L / 2 + L * iu / N, -L / 2 + L * iv / N, np.full((N, N), 0.5)], axis=-1)
    d = np.linalg.norm(P - np.array(p), axis=-1)
    phase = np.exp(1j * kmag * np.einsum("wc,ijc->wij", waves, P - np.array(focus)))
    want = w1 * subA * phase.sum(axis=0)
    inside = d <= a
    assert inside.sum() > 20 and (~inside).sum() > 20
    assert np.max(np.abs(U[inside] - want[inside])) < 1e-10
    assert np.max(np.abs(U[~inside])) == 0.0


def tone(N, a, b):
    iv, iu = np.mgrid[0:N, 0:N]
    return np.exp(2j * np.pi * (a * iu + b * iv) / N)


@pytest.mark.parametrize("quirks", [1, 0])
def test_detector_passes_and_removes_tones(orc, quirks):
    N, L, lam = 32, 32.0, 1.0
    du = dv = 1.0 / L               # bin (a, b) has |f| = sqrt(a^2 + b^2) / L
    NAin, NAout = 2.5 / L, 6.5 / L  # pass 2.5 <= |bin| <= 6.5 (lam = 1)
    Uf = np.full((N, N), 2.0 + 0j)  # DC: removed by the obscuration
    cases = [((3, 0), True), ((0, 5), True), ((4, 4), True), ((1, 1), False), ((7, 0), False),
             ((5, 5), False)]
    for (a, b), passes in cases:
        out = orc.detector(tone(N, a, b), Uf, du, dv, NAin, NAout, lam, 0, N, 1, quirks=quirks)
        D, Di = out[2], out[3]
        assert np.allclose(D, 1.0 if passes else 0.0, atol=1e-12), (a, b)
        assert np.allclose(Di, 0.0, atol=1e-12)
    # negative frequency -3 sits at index N-3: the reference reads it as
    # -(N-1-(N-3)) = -2 (outside the pass band), the corrected index as -3
    out = orc.detector(tone(N, N - 3, 0), Uf, du, dv, NAin, NAout, lam, 0, N, 1, quirks=quirks)
    assert np.allclose(out[2], 0.0 if quirks else 1.0, atol=1e-12)


def test_detector_constant_fields_crop_and_supersampling(orc):
    Nd, ss, pad = 8, 2, 1
    N = Nd * ss * (2 * pad + 1)
    U = np.full((N, N), 0.5 - 0.5j)
    Uf = np.full((N, N), 1.0 + 0.0j)
    out = orc.detector(U, Uf, 1.0 / 30, 1.0 / 30, 0.0, 0.5, 2.0, pad, Nd, ss)
    assert np.allclose(out[2], 0.5, atol=1e-13) and np.allclose(out[3], 1.0, atol=1e-13)
    assert np.allclose(out[4], math.log10(2.0), atol=1e-13)
    # D and Di accumulate (+=)
    out2 = orc.detector(U, Uf, 1.0 / 30, 1.0 / 30, 0.0, 0.5, 2.0, pad, Nd, ss, D=out[2], Di=out[3])
    assert np.allclose(out2[2], 1.0, atol=1e-13)


def test_sum_without_spheres_is_the_incident_field(orc):
    import oracle
    q = oracle.Quad.xy(10.0)
    rng = np.random.default_rng(0)
    Uf = rng.normal(size=(6, 6)) + 1j * rng.normal(size=(6, 6))
    U = rng.normal(size=(6, 6)) + 1j * rng.normal(size=(6, 6))
    assert np.array_equal(orc.nf_sum_uf(q, 6, 6, [], Uf, U), Uf)
    # a sphere covering the whole slice keeps U; one far away adds Uf everywhere
    big = dict(p=(0.0, 0.0, 0.0), a=100.0, n=1.5 + 0j, Nl=1, coeff_off=0)
    far = dict(p=(0.0, 0.0, 50.0), a=1.0, n=1.5 + 0j, Nl=1, coeff_off=0)
    assert np.array_equal(orc.nf_sum_uf(q, 6, 6, [big], Uf, U), U)
    assert np.allclose(orc.nf_sum_uf(q, 6, 6, [far], Uf, U), U + Uf, atol=0)


@pytest.mark.parametrize("n,z", [(1.45 + 0.03j, 0.0), (1.3 + 0.0j, 0.7)])
def test_scattered_series_is_the_brute_force_sum(orc, n, z):
    """With the reference recurrences off, the oracle's scattered/internal
    series equals sum_w (1/nk) subA e^{i k K_w.c} sum_l c_l P_l(K_w.r-hat)
    with c_l = B_l h_l(kd) outside (A_l j_l(n kd) inside), every term from
    scipy's Bessel functions and Legendre polynomials."""
    import oracle
    lam, a = 2.5, 3.0
    kmag = 2.0 * PI_F / lam
    N, L = 10, 14.0
    q = oracle.Quad.xy(L, z)
    p = (0.6, -1.1, 0.0)
    nl = orc.sphere_order(a, lam)
    A, B = orc.mie_coeffs(nl, kmag, a, n, quirks=0)
    assert np.max(np.abs(B)) > 0.1
    s = dict(p=p, a=a, n=n, Nl=nl, coeff_off=0)
    rng = np.random.default_rng(11)
    waves = rng.normal(size=(6, 3))
    waves[:, 2] = np.abs(waves[:, 2]) + 1.5
    waves /= np.linalg.norm(waves, axis=1)[:, None]
    subA, focus = 0.7, (0.2, 0.1, 0.0)
    U = orc.nf_us(q, N, N, focus, kmag, [s], A, B, waves, subA, quirks=0)
    w1 = float(np.float32(1.0) / np.float32(len(waves)))
    c = np.array(p) - np.array(focus)
    l = np.arange(nl + 1)
    want = np.zeros((N, N), dtype=complex)
    for iv in range(N):
        for iu in range(N):
            r = np.array([-L / 2 + L * iu / N, -L / 2 + L * iv / N, z]) - np.array(p)
            d = np.linalg.norm(r)
            rh = r / d
            if d <= a:
                cl = A * special.spherical_jn(l, n * kmag * d)
            else:
                cl = B * sph_h(l, kmag * d)
            for k in waves:
                P = special.eval_legendre(l, float(k @ rh))
                want[iv, iu] += w1 * subA * np.exp(1j * kmag * float(k @ c)) * np.sum(cl * P)
    inside = np.array([[np.linalg.norm(np.array([-L / 2 + L * iu / N, -L / 2 + L * iv / N, z]) -
                                       np.array(p)) <= a for iu in range(N)] for iv in range(N)])
    assert inside.sum() > 5 and (~inside).sum() > 30
    scale = np.max(np.abs(want))
    assert np.max(np.abs(U[~inside] - want[~inside])) / scale < 1e-10
    assert np.max(np.abs(U[inside] - want[inside])) / scale < 1e-10


def test_mc_incident_field_is_the_internal_field_of_an_index_one_sphere(orc):
    """The Monte-Carlo incident field sums the same weighted plane waves that
    an index-1 sphere's internal series reproduces (by the Mie and Legendre
    path of the scattered field), so the two agree inside such a sphere."""
    import oracle
    lam, a = 3.0, 6.0
    kmag = 2.0 * PI_F / lam
    N, L = 12, 20.0
    q = oracle.Quad.xy(L, 0.5)
    p = (1.0, -0.5, 0.0)
    nl = orc.sphere_order(a, lam) + 25
    A, B = orc.mie_coeffs(nl, kmag, a, 1.0 + 0.0j, quirks=0)
    s = dict(p=p, a=a, n=1.0 + 0.0j, Nl=nl, coeff_off=0)
    rng = np.random.default_rng(4)
    waves = rng.normal(size=(9, 3))
    waves[:, 2] = np.abs(waves[:, 2]) + 1.0
    waves /= np.linalg.norm(waves, axis=1)[:, None]
    subA, focus = 0.8, (0.3, 0.2, -0.1)
    Ui = orc.nf_us(q, N, N, focus, kmag, [s], A, B, waves, subA, quirks=0)
    Uf = orc.nf_uf_mc(q, N, N, focus, kmag, waves, subA)
    iv, iu = np.mgrid[0:N, 0:N]
    P = np.stack([-L / 2 + L * iu / N, -L / 2 + L * iv / N, np.full((N, N), 0.5)], axis=-1)
    inside = np.linalg.norm(P - np.array(p), axis=-1) <= a
    assert inside.sum() > 20
    assert np.max(np.abs(Uf[inside] - Ui[inside])) < 1e-10
    # and everywhere it is the weighted plane-wave sum itself
    w1 = float(np.float32(1.0) / np.float32(len(waves)))
    want = w1 * subA * np.exp(1j * kmag * np.einsum("wc,ijc->wij", waves,
                                                    P - np.array(focus))).sum(axis=0)
    assert np.max(np.abs(Uf - want)) < 1e-12


def test_mc_incident_field_of_one_wave_is_the_plane_wave(orc):
    import oracle
    q = oracle.Quad.xy(9.0, 0.3)
    k = (0.6, 0.0, 0.8)
    kmag = 2.0 * PI_F / 2.0
    Up = orc.nf_uf(q, 7, 7, (0.1, 0.0, 0.0), k, kmag, 1.7, 1.0, 0.5, 10, mode=1)
    Um = orc.nf_uf_mc(q, 7, 7, (0.1, 0.0, 0.0), kmag, np.array([k]), 1.7)
    assert np.max(np.abs(Up - Um)) < 1e-14


def test_mc_incident_field_converges_to_the_series(orc, pkg):
    """With many waves drawn uniformly over the condenser annulus the sum
    tends to the focused field (the Legendre series, itself pinned to the
    annulus integral), within its Monte-Carlo error."""
    import oracle
    lam, z = 2.5, 1.2
    kmag = 2.0 * PI_F / lam
    NAin, NAout = 0.2, 0.6
    ca, cb = math.cos(math.asin(NAin)), math.cos(math.asin(NAout))
    N, L = 6, 8.0
    q = oracle.Quad.xy(L, z)
    nk = 40000
    waves = orc.mc_sample_na(nk, (0.0, 0.0, 1.0), NAin, NAout, 1, 77)
    subA = 2.0 * PI_F * 1.0 * ((1.0 - cb) - (1.0 - ca))
    Umc = orc.nf_uf_mc(q, N, N, (0.0, 0.0, 0.0), kmag, waves, subA)
    m = 60
    Us = orc.nf_uf(q, N, N, (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), kmag, 1.0, ca, cb, m, quirks=0)
    err = np.max(np.abs(Umc - Us)) / np.max(np.abs(Us))
    assert np.max(np.abs(Us.imag)) > 0.1 * np.max(np.abs(Us))   # a phase-sign slip would show
    assert err < 3.0 / math.sqrt(nk), err
