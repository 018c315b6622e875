"""Seeded BIMSim instances (sphere lists, optics, slice, Monte-Carlo waves).

Every parameter is drawn from SplitMix64 with seed 0xB1A50000 + config id and
rounded to float32, because the reference reads every input as float
(rts/tools/arguments.h:100-116, ptype = float).  Quantities the reference
derives on the host in ptype (k = 2*PI/lambda, cos(asin(NA)), the condenser
amplitude subA, du = 1/|X|) are derived here in double, i.e. the reference
with ptype = double.

Named workloads (slice of N x N samples over L um at z = 0, focus at the
origin, k = +z, detector without padding and ss = 1):
  cfg1   one sphere a=5, n=1.4+0.05i, lambda=2.5, N=128, L=40, condenser 0-0.5,
         objective 0-0.5, Legendre Uf, direct Us series with 400 waves
  cfg2   3 spheres a~U(2,8), n~U(1.3,1.6)+iU(0,0.1), centres U(-15,15)^2,
         1000 waves over NA 0.2-0.62, lambda=3, N=512, L=60, objective 0-0.62
  cfg3   50 non-overlapping spheres a~U(1,10), n as cfg2, centres
         U(-100,100)^2, lambda=4, N=1024, L=200, 400 waves, condenser and
         objective 0-0.5; scattered field from the sphere tables (radial
         tables of 4096 samples, 1000 angles)
  cfg4   spectral sweep: 256 wavenumbers 900-4000 cm^-1 (lambda = 1e4/nu), 20
         non-overlapping spheres a~U(2,15), centres U(-50,50)^2, n(nu) =
         1.5 + 0.08 (g/2)/((nu0 - nu) - i g/2), nu0 = 1650, g = 30; N=512,
         L=100, condenser and objective 0.2-0.5, 400 waves
  cfg5   ragged: single-sphere instances, a log-uniform [0.5, 25],
         lambda~U(2.5,11), n~U(1.2,1.7)+iU(0,0.2), N from {256,384,512,768,
         1024}, L = 4a, 200-2000 Monte-Carlo waves each; the incident field
         by the Legendre series or (Bernoulli 1/2) as the Monte-Carlo sum of
         the same waves; condenser and objective 0-0.5
"""
import math
import struct

import numpy as np

PI_F = 3.1415901184082031   # float(3.14159f)
SEED_BASE = 0xB1A50000
MASK64 = (1 << 64) - 1


def f32(x):
    """Round a double to the nearest float32 value."""
    return struct.unpack("f", struct.pack("f", x))[0]


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & MASK64

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / 9007199254740992.0))


def sphere_order(a, lam):
    """sphere::calcNl (sphere.h:71-74), evaluated by the library's host helper."""
    from . import sphere_order as lib_order
    return lib_order(a, lam)


def incident_order(kmag, dmax):
    """m = ceil(kd + 4 kd^(1/3) + 2) at the farthest slice point."""
    kd = kmag * dmax
    return int(math.ceil(kd + 4.0 * kd ** (1.0 / 3.0) + 2.0))


def make_instance(spheres, lam, N, L, cond, obj, nk, wave_seed, amp=1.0, m=None,
                  plane_wave=False, uf_mc=False):
    lam = f32(lam)
    kmag = (2.0 * PI_F) / lam
    spheres = [dict(s) for s in spheres]
    off = 0
    for s in spheres:
        s["p"] = tuple(f32(c) for c in s["p"])
        s["a"] = f32(s["a"])
        s["n"] = complex(f32(s["n"].real), f32(s["n"].imag))
        s["Nl"] = sphere_order(s["a"], lam)
        s["coeff_off"] = off
        off += s["Nl"] + 1
    cond = (f32(cond[0]), f32(cond[1]))
    obj = (f32(obj[0]), f32(obj[1]))
    ca = math.cos(math.asin(cond[0]))
    cb = math.cos(math.asin(cond[1]))
    half = L / 2.0
    if m is None:
        m = incident_order(kmag, math.hypot(half, half))
    inst = dict(
        lam=lam, kmag=kmag, N=N, L=float(L), A=(-half, -half, 0.0), X=(float(L), 0.0, 0.0),
        Y=(0.0, float(L), 0.0), focus=(0.0, 0.0, 0.0), kdir=(0.0, 0.0, 1.0), amp=amp,
        cond=cond, obj=obj, cos_alpha=ca,
        cos_beta=cb, m=m, spheres=spheres, ncoef=off, nk=nk, wave_seed=wave_seed,
        plane_wave=plane_wave, uf_mc=uf_mc, padding=0, Nd=N, ss=1, du=1.0 / L, dv=1.0 / L,
    )
    inst["subA"] = (2.0 * PI_F) * amp * ((1.0 - cb) - (1.0 - ca))
    return inst


def mc_waves(inst, sampler):
    """Monte-Carlo directions for an instance, rounded to float32.
    `sampler(N, k, NAin, NAout, rng, seed)` is mcSampleNA with SplitMix64."""
    w = np.asarray(sampler(inst["nk"], inst["kdir"], inst["cond"][0], inst["cond"][1], 1,
                           inst["wave_seed"]), dtype=np.float64).reshape(-1, 3)
    return w.astype(np.float32).astype(np.float64)


def config(cid, N=None, nk=None):
    """Named workload `cid` (1 or 2), optionally reduced in N / nk for tests."""
    g = SplitMix64(SEED_BASE + cid)
    if cid == 1:
        sph = [dict(p=(0.0, 0.0, 0.0), a=5.0, n=complex(1.4, 0.05))]
        return make_instance(sph, 2.5, N or 128, 40.0, (0.0, 0.5), (0.0, 0.5), nk or 400,
                             SEED_BASE + 0x100 + cid)
    if cid == 2:
        sph = []
        for _ in range(3):
            a = g.uniform(2.0, 8.0)
            nr = g.uniform(1.3, 1.6)
            ni = g.uniform(0.0, 0.1)
            cx = g.uniform(-15.0, 15.0)
            cy = g.uniform(-15.0, 15.0)
            sph.append(dict(p=(cx, cy, 0.0), a=a, n=complex(nr, ni)))
        return make_instance(sph, 3.0, N or 512, 60.0, (0.2, 0.62), (0.0, 0.62), nk or 1000,
                             SEED_BASE + 0x100 + cid)
    raise ValueError(f"unknown configuration {cid}")


def _non_overlapping(g, count, a_lo, a_hi, half):
    """Rejection sampling of `count` spheres (radius, centre) with
    |c_i - c_j| >= a_i + a_j."""
    out = []
    tries = 0
    while len(out) < count:
        tries += 1
        if tries > 200000:
            raise RuntimeError("cannot place non-overlapping spheres")
        a = g.uniform(a_lo, a_hi)
        cx = g.uniform(-half, half)
        cy = g.uniform(-half, half)
        if all(math.hypot(cx - o[1], cy - o[2]) >= a + o[0] for o in out):
            out.append((a, cx, cy))
    return out


def lorentz_index(nu, nu0=1650.0, gamma=30.0, strength=0.08, n0=1.5):
    """n(nu) = n0 + strength (gamma/2) / ((nu0 - nu) - i gamma/2): imaginary
    peak `strength` at nu0, real part Kramers-Kronig consistent."""
    return n0 + strength * (gamma / 2.0) / complex(nu0 - nu, -gamma / 2.0)


def config3(N=1024, ns=50, nk=400, L=200.0):
    g = SplitMix64(SEED_BASE + 3)
    sph = []
    for a, cx, cy in _non_overlapping(g, ns, 1.0, 10.0, L / 2.0):
        nr, ni = g.uniform(1.3, 1.6), g.uniform(0.0, 0.1)
        sph.append(dict(p=(cx, cy, 0.0), a=a, n=complex(nr, ni)))
    inst = make_instance(sph, 4.0, N, L, (0.0, 0.5), (0.0, 0.5), nk, SEED_BASE + 0x103)
    inst["lut"] = dict(dR=4096, thetaR=1000)
    return inst


def config4(n_nu=256, N=512, ns=20, nk=400, L=100.0, nu_lo=900.0, nu_hi=4000.0, stride=1):
    """The spectral sweep as a list of instances (every `stride`-th
    wavenumber of the n_nu-point grid)."""
    g = SplitMix64(SEED_BASE + 4)
    geo = _non_overlapping(g, ns, 2.0, 15.0, L / 2.0)
    out = []
    for i in range(0, n_nu, stride):
        nu = nu_lo + i * (nu_hi - nu_lo) / (n_nu - 1)
        lam = 1e4 / nu
        n = lorentz_index(nu)
        sph = [dict(p=(cx, cy, 0.0), a=a, n=n) for a, cx, cy in geo]
        inst = make_instance(sph, lam, N, L, (0.2, 0.5), (0.2, 0.5), nk, SEED_BASE + 0x104)
        inst["nu"] = nu
        out.append(inst)
    return out


def config5(count=1024, sizes=(256, 384, 512, 768, 1024), nk_range=(200, 2000)):
    g = SplitMix64(SEED_BASE + 5)
    out = []
    for i in range(count):
        a = math.exp(g.uniform(math.log(0.5), math.log(25.0)))
        lam = g.uniform(2.5, 11.0)
        n = complex(g.uniform(1.2, 1.7), g.uniform(0.0, 0.2))
        N = sizes[min(len(sizes) - 1, int(g.uniform(0, len(sizes))))]
        mc = g.uniform() < 0.5
        nk = nk_range[0] + int(g.uniform(0, nk_range[1] - nk_range[0] + 1))
        inst = make_instance([dict(p=(0.0, 0.0, 0.0), a=a, n=n)], lam, N, 4.0 * a, (0.0, 0.5),
                             (0.0, 0.5), nk, SEED_BASE + 0x1000 + i, uf_mc=mc)
        out.append(inst)
    return out


def terms(inst):
    """Mie series terms of one image: N^2 nk sum_s (Nl+1) + N^2 (m+1), the
    incident-field part N^2 nk when it is the Monte-Carlo sum."""
    n2 = inst["N"] * inst["N"]
    uf = inst["nk"] if inst.get("uf_mc") else inst["m"] + 1
    return n2 * inst["nk"] * sum(s["Nl"] + 1 for s in inst["spheres"]) + n2 * uf


def us_flops(inst):
    """Algorithmic FLOP of the scattered-field series: per (pixel, sphere)
    (Nl+1)(8 nk + 12) + 13 nk."""
    n2 = inst["N"] * inst["N"]
    nk = inst["nk"]
    return n2 * sum((s["Nl"] + 1) * (8 * nk + 12) + 13 * nk for s in inst["spheres"])
