"""Pins of the oracle's look-up-table path and rotation resample against
independent computations: scipy Bessel/Legendre sums at table nodes, the
direct series with exact Bessel functions, and an exact 90-degree rotation."""
import math
import sys

import numpy as np
import pytest
from scipy import special

S_REF = math.sqrt(3.14159 / math.pi)   # the tables use sqrt(3.14159/2x)-scaled Bessels


def normrel(a, b):
    m = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (m if m > 0 else 1.0))


@pytest.fixture(scope="module")
def W(pkg):
    import bimsim_amd.workloads as W
    return W


def test_sphere_tables_at_nodes_match_scipy_sums(orc):
    nl, kmag, a, n = 18, 2.0 * 3.1415901184082031 / 3.0, 3.0, complex(1.4, 0.02)
    A, B = orc.mie_coeffs(nl, kmag, a, n)
    dmin, dmax, dR, thetaR = 0.0, 20.0, 64, 40
    h = orc.lut_hankel(nl, kmag, a, dmin, dmax, dR)
    Usp = orc.lut_eval(nl, B, h, thetaR)
    lo = max(a, dmin)
    r = lo + np.arange(dR) * (dmax - lo) / (dR - 1)
    th = np.arange(thetaR) * np.float32(np.float32(3.14159) / np.float32(thetaR - 1))
    l = np.arange(nl + 1)
    H = S_REF * (special.spherical_jn(l[None, :], kmag * r[:, None]) +
                 1j * special.spherical_yn(l[None, :], kmag * r[:, None]))
    P = special.eval_legendre(l[None, :], np.cos(th)[:, None])
    want = np.einsum("tl,rl,l->tr", P, H, B)
    assert normrel(Usp, want) < 1e-10
    aR = 9
    j = orc.lut_bessel(nl, kmag, a, n, aR)
    rho = np.arange(aR) * (a / (aR - 1))
    J = S_REF * special.spherical_jn(l[None, :], kmag * n * rho[1:, None])
    assert normrel(j[1:], J) < 1e-10
    assert j[0, 0] == 1 and np.all(j[0, 1:] == 0)


def test_table_lookup_matches_exact_series(pkg, orc, W):
    """gpuScalarUpLut on fine tables reproduces the direct series with exact
    Bessel functions, up to the tables' sqrt(3.14159/pi) scale and bilinear
    interpolation error."""
    import bimsim_amd.pipeline as P
    sph = [dict(p=(1.0, -2.0, 0.0), a=3.0, n=complex(1.4, 0.02))]
    inst = W.make_instance(sph, 3.0, 24, 24.0, (0.0, 0.5), (0.0, 0.5), 6, 5)
    q = P.quad_of(inst)
    N = inst["N"]
    s = inst["spheres"][0]
    waves = W.mc_waves(inst, pkg.mc_sample_na)
    A, B = orc.mie_coeffs(s["Nl"], inst["kmag"], s["a"], s["n"])
    aR, dR, dmin, dmax = orc.sphere_lut_dims(q, s, N)
    dR, thetaR, aR = 4096, 4000, 600      # fine tables: small interpolation error
    h = orc.lut_hankel(s["Nl"], inst["kmag"], s["a"], dmin, dmax, dR)
    Usp = orc.lut_eval(s["Nl"], B, h, thetaR)
    j = orc.lut_bessel(s["Nl"], inst["kmag"], s["a"], s["n"], aR)
    Uip = orc.lut_eval(s["Nl"], A, j, thetaR)
    U_lut = orc.nf_us_lut(q, N, N, inst["focus"], inst["kmag"], s, Usp, Uip, dmin, dmax, waves,
                          inst["subA"])
    U_ser = orc.nf_us(q, N, N, inst["focus"], inst["kmag"], [s], A, B, waves, inst["subA"],
                      quirks=0)
    # Pixels exactly on the sphere surface are inside for the series
    # (nfScalarUs.cu:373, d <= a) and outside for the tables
    # (nfScalarUpLut.cu:80, d < a); compare everywhere else.
    iv, iu = np.mgrid[0:N, 0:N]
    A0, X, Y = (np.array(v) for v in (q.A, q.X, q.Y))
    pos = A0 + X * (iu / N)[..., None] + Y * (iv / N)[..., None]
    d = np.linalg.norm(pos - np.array(s["p"]), axis=-1)
    off = np.abs(d - s["a"]) > 1e-9
    assert (~off).sum() == 4                 # the grid meets the surface at 4 pixels
    assert normrel(U_lut[off], S_REF * U_ser[off]) < 2e-3


def test_uf_table_matches_exact_series(pkg, orc, W):
    import bimsim_amd.pipeline as P
    inst = W.config(1, N=32)
    q = P.quad_of(inst)
    N = inst["N"]
    dmin, dmax = orc.quad_dist(q, inst["focus"])
    dR = 16 * N
    lut = orc.uf_bessel_lut(inst["m"], inst["kmag"], dmin, dmax, dR)
    args = (inst["focus"], inst["kdir"], inst["kmag"], inst["amp"], inst["cos_alpha"],
            inst["cos_beta"], inst["m"])
    U_lut = orc.nf_uf(q, N, N, *args, mode=2, lut=lut, dR=dR, dmin=dmin, dmax=dmax)
    U_ser = orc.nf_uf(q, N, N, *args, mode=0, quirks=0)
    assert normrel(U_lut, S_REF * U_ser) < 2e-3


def test_quad_distance_closed_form(orc):
    import oracle
    q = oracle.Quad.xy(40.0, 0.0)
    dmin, dmax = orc.quad_dist(q, (0.0, 0.0, 3.0))
    assert abs(dmin - 3.0) < 1e-12 and abs(dmax - math.sqrt(800 + 9)) < 1e-12
    dmin, dmax = orc.quad_dist(q, (25.0, 0.0, 0.0))
    assert abs(dmin - 5.0) < 1e-12
    dmin, dmax = orc.quad_dist(q, (25.0, 30.0, 4.0))
    assert abs(dmin - math.sqrt(25 + 100 + 16)) < 1e-12


@pytest.mark.parametrize("N", [16, 24])
def test_rotation_by_quarter_turn_is_a_permutation(orc, N):
    rng = np.random.default_rng(N)
    U = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    assert normrel(orc.nf_interpolate(U, 0.0, centre_float=1), U) < 1e-15
    out = orc.nf_interpolate(U, math.pi / 2, centre_float=1)
    c = N // 2
    want = np.zeros_like(U)
    for iv in range(N):
        for iu in range(N):
            x, y = c - (iv - c), c + (iu - c)
            if 0 <= x < N and 0 <= y < N:
                want[iv, iu] = U[y, x]
    assert normrel(out, want) < 1e-12
