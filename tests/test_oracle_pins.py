"""Pins of the oracle's special functions, recurrences and sampler against the
reference's own compiled code (golden fixtures from tests/golden/, and
oracle/_ref live when present), scipy and closed forms."""
import ctypes
import math
import os

import numpy as np
import pytest
from scipy import special

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PI_REF = 3.14159   # bessjy.cpp:18


def normrel(a, b):
    m = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (m if m > 0 else 1.0)


# ---------------------------------------------------------------- golden
def test_real_bessel_matches_reference_fixtures(orc):
    g = np.load(os.path.join(GOLD, "ref_bessel.npz"))
    for nl, x, out in zip(g["real_nl"], g["real_x"], g["real_out"]):
        mine = orc.sph_jy(int(nl), float(x), quirks=1)
        for k in range(4):
            ref = out[k, : nl + 1]
            assert normrel(mine[k], ref) <= 1e-14, (nl, x, k)


def test_complex_bessel_matches_reference_fixtures(orc):
    g = np.load(os.path.join(GOLD, "ref_bessel.npz"))
    for nl, z, out in zip(g["cplx_nl"], g["cplx_z"], g["cplx_out"]):
        mine = orc.sph_jy_c(int(nl), complex(z), quirks=1)
        for k in range(4):
            assert normrel(mine[k], out[k, : nl + 1]) <= 1e-14, (nl, z, k)


def test_gamma_matches_reference_fixtures(orc):
    g = np.load(os.path.join(GOLD, "ref_bessel.npz"))
    for x, v in zip(g["gamma_x"], g["gamma_out"]):
        assert orc.gamma(float(x)) == v


def test_recurrences_match_reference_fixtures(orc):
    g = np.load(os.path.join(GOLD, "ref_recur.npz"))
    for i, (nl, x) in enumerate(zip(g["nl"], g["x"])):
        for fp32 in (0, 1):
            j, y = orc.sbessel_up(int(nl), float(x), fp32)
            assert np.array_equal(j, g["jy"][i, fp32, 0, : nl + 1]), (nl, x, fp32)
            assert np.array_equal(y, g["jy"][i, fp32, 1, : nl + 1]), (nl, x, fp32)
    for fp32 in (0, 1):
        for x, P in zip(g["leg_x"], g["leg"][fp32]):
            assert np.array_equal(orc.legendre_up(80, float(x), fp32), P)
    for nl, z, jc in zip(g["cz_nl"], g["cz"], g["jc32"]):
        assert np.array_equal(orc.sbessel_up_c(int(nl), complex(z), 1), jc[: nl + 1])


def test_mc_sampler_matches_reference_fixtures(orc):
    g = np.load(os.path.join(GOLD, "ref_mc.npz"))
    for i in range(3):
        out = g[f"out{i}"]
        mine = orc.mc_sample_na(out.shape[0], g["k"][i], g["na"][i][0], g["na"][i][1], 0,
                                int(g["seed"][i]))
        assert np.array_equal(mine, out)


# ------------------------------------------------------------ live _ref
def test_bessel_random_against_live_reference(orc, ref):
    rng = np.random.default_rng(5)
    for _ in range(150):
        nl = int(rng.integers(2, 120))
        x = float(rng.uniform(1e-3, 150.0))
        for a, b in zip(orc.sph_jy(nl, x), ref.sph_jy(nl, x)):
            assert normrel(a, b) <= 1e-14
        z = complex(rng.uniform(1e-2, 120.0), rng.uniform(0.0, 10.0))
        for a, b in zip(orc.sph_jy_c(nl, z), ref.sph_jy_c(nl, z)):
            assert normrel(a, b) <= 1e-14
    for _ in range(60):
        x = float(rng.uniform(0.01, 300.0))
        mp = int(rng.integers(5, 250))
        assert orc.msta1(x, mp) == ref.msta1(x, mp)
        n = int(rng.integers(1, 300))
        assert orc.msta2(x, n, 15) == ref.msta2(x, n, 15)


def test_mc_and_recurrences_against_live_reference(orc, ref):
    rng = np.random.default_rng(9)
    for _ in range(5):
        k = rng.normal(size=3)
        k /= np.linalg.norm(k)
        k = k.astype(np.float32).astype(np.float64)
        na_in, na_out = sorted(rng.uniform(0, 0.95, size=2))
        seed = int(rng.integers(0, 2 ** 31))
        assert np.array_equal(orc.mc_sample_na(257, k, na_in, na_out, 0, seed),
                              ref.mc_sample_na(257, k, na_in, na_out, seed))
    for _ in range(100):
        nl = int(rng.integers(2, 90))
        x = float(rng.uniform(0.05, 80.0))
        for fp32 in (0, 1):
            a = orc.sbessel_up(nl, x, fp32)
            b = ref.sbessel_up(nl, x, fp32)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_glibc_stream():
    from oracle import Oracle
    libc = ctypes.CDLL("libc.so.6")
    o = Oracle()
    for seed in (0, 1, 12345):
        libc.srand(seed)
        seq = [libc.rand() for _ in range(700)]
        assert list(o.glibc_rand(seed, 700)) == seq


# ------------------------------------------------ scipy and closed forms
def test_real_bessel_against_scipy(orc):
    s = math.sqrt(PI_REF / math.pi)   # the reference's sqrt(3.14159/2x) scale
    for x in [0.3, 2.0, 7.5, 12.0, 12.5, 33.0, 71.0]:
        nl = 60
        j0, y0, jp0, yp0 = orc.sph_jy(nl, x, quirks=0)
        j1, y1, _, _ = orc.sph_jy(nl, x, quirks=1)
        l = np.arange(nl + 1)
        J = special.spherical_jn(l, x)
        Y = special.spherical_yn(l, x)
        keep = np.abs(J) > 1e-250
        assert np.max(np.abs(j0[keep] - J[keep]) / np.abs(J[keep])) < 1e-11
        ky = np.abs(Y) < 1e250
        assert np.max(np.abs(y0[ky] - Y[ky]) / np.abs(Y[ky])) < 1e-11
        assert np.max(np.abs(j1[keep] - s * J[keep]) / np.abs(s * J[keep])) < 1e-11
        # derivatives
        assert normrel(jp0, special.spherical_jn(l, x, derivative=True)) < 1e-11
        # Wronskian j_l y_l' - j_l' y_l = 1/x^2
        w = j0 * yp0 - jp0 * y0
        good = (np.abs(y0) < 1e100)
        assert np.max(np.abs(w[good] * x * x - 1.0)) < 1e-9


def test_complex_bessel_against_scipy_and_typo(orc):
    s = math.sqrt(PI_REF / math.pi)
    for z in [3.0 + 0.2j, 8.0 + 0.5j, 11.5 + 0.05j, 15.0 + 0.7j, 30.0 + 1.5j, 12.566 + 0.0j,
              60.0 + 0.0j]:
        nl = 40
        l = np.arange(nl + 1)
        J = special.spherical_jn(l, z)
        Y = special.spherical_yn(l, z)
        j0, y0, _, _ = orc.sph_jy_c(nl, z, quirks=0)
        j1, y1, _, _ = orc.sph_jy_c(nl, z, quirks=1)
        keep = np.abs(J) > 1e-200
        assert np.max(np.abs(j0[keep] - J[keep]) / np.abs(J[keep])) < 1e-10
        assert np.max(np.abs(j1[keep] - s * J[keep]) / np.abs(s * J[keep])) < 1e-10
        assert normrel(y0, Y) < 1e-10
        if abs(z) <= 12.0:
            assert normrel(y1, s * Y) < 1e-10
        elif z.imag == 0.0:
            # cbessjy.cpp:541: for v0 = 1/2 the asymptotic P = 1, Q = 0 and the
            # typo gives Y_{1/2} + J_{1/2}: y_0 = s (sin z - cos z)/z, and the
            # error rides the upward recurrence into every order.  (For complex
            # z the Wronskian pass of cbessjy.cpp:668-713 rebuilds Y and can
            # repair it when it normalises through Y_{3/2}.)
            assert abs(y1[0] - s * (np.sin(z) - np.cos(z)) / z) < 1e-12 * abs(y1[0])
            assert normrel(y1, s * Y) > 1e-5


def test_legendre_and_hacked_recurrences_closed_forms(orc):
    l = np.arange(81)
    for x in [-1.0, -0.37, 0.0, 0.5, 0.999, 1.0]:
        P = orc.legendre_up(80, x)
        assert np.max(np.abs(P - special.eval_legendre(l, x))) < 1e-12
    # below l = x the thresholds never act: the recurrence is the textbook one
    for x in [20.0, 45.5]:
        j, y = orc.sbessel_up(int(x) - 1, x)
        ll = np.arange(int(x))
        assert np.max(np.abs(j - special.spherical_jn(ll, x))) < 1e-10
        assert normrel(y, special.spherical_yn(ll, x)) < 1e-10
    # above l = x: j is zeroed outside [0.0145, 0.4] and stays zero
    j, y = orc.sbessel_up(60, 10.0)
    assert np.all(j[11:][(j[11:] != 0)] >= 0.0145)
    first_zero = np.argmax((np.arange(61) > 10) & (j == 0))
    assert np.all(j[first_zero + 1:] == 0)


def test_gamma_and_msta(orc):
    for x in [0.5, 1.5, 2.5, -0.5, -1.5, 3.7, 10.2]:
        assert abs(orc.gamma(x) - math.gamma(x)) < 1e-14 * abs(math.gamma(x))
    for x in [0.5, 5.0, 50.0]:
        m = orc.msta1(x, 200)
        # order at which |J_m(x)| ~ 10^-200
        assert -205 < math.log10(abs(special.jv(m, x))) < -190
