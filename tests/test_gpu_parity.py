"""Parity of the HIP library (through the C ABI) with the oracle on the same
seeded inputs.  Tolerances (normwise: max |gpu - oracle| / max |oracle|):
  FP64_PARITY   <= 1e-10  (the reference algorithm in double)
  FP64          bit-identical to FP64_PARITY
  FP32_FAST     field <= 2e-5, absorbance |dA| <= 1e-4 where the incident
                image is above 1e-3 of its peak
"""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def normrel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    m = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (m if m > 0 else 1.0))


def dev(x, dtype):
    return torch.as_tensor(np.ascontiguousarray(x)).to(dtype).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.fixture(scope="module")
def W(pkg):
    import bimsim_amd.workloads as W
    return W


def instance_waves(pkg, W, inst):
    return W.mc_waves(inst, pkg.mc_sample_na)


# ------------------------------------------------------------ Mie coefficients
@pytest.mark.parametrize("quirks", [1, 0])
def test_mie_coeffs_match_oracle(pkg, ctx, orc, W, quirks):
    cases = [(5.0, 1.4 + 0.05j, 2.5), (2.0, 1.33 + 0.0j, 3.0), (7.9, 1.55 + 0.09j, 3.0),
             (10.0, 1.2 + 0.2j, 4.0), (1.0, 1.6 + 0.01j, 4.0), (15.0, 1.5 + 0.08j, 2.5),
             (0.6, 1.7 + 0.2j, 11.0), (25.0, 1.25 + 0.1j, 3.5)]
    spheres, off = [], 0
    for a, n, lam in cases:
        Nl = pkg.sphere_order(a, lam)
        spheres.append(dict(p=(0.0, 0.0, 0.0), a=a, n=n, Nl=Nl, coeff_off=off, lam=lam))
        off += Nl + 1
    A = torch.zeros(off, dtype=torch.complex128, device="cuda")
    B = torch.zeros(off, dtype=torch.complex128, device="cuda")
    # one call per wavelength
    for s in spheres:
        kmag = 2.0 * pkg.PI_F / s["lam"]
        ctx.mie_coeffs([s], kmag, A, B, quirks=quirks)
    Ah, Bh = host(A), host(B)
    for s in spheres:
        kmag = 2.0 * pkg.PI_F / s["lam"]
        Ao, Bo = orc.mie_coeffs(s["Nl"], kmag, s["a"], s["n"], quirks)
        sl = slice(s["coeff_off"], s["coeff_off"] + s["Nl"] + 1)
        assert normrel(Ah[sl], Ao) <= 1e-11, (s, normrel(Ah[sl], Ao))
        assert normrel(Bh[sl], Bo) <= 1e-11, (s, normrel(Bh[sl], Bo))


# ------------------------------------------------------------ incident field
@pytest.mark.parametrize("mode", [0, 1])
def test_incident_field_matches_oracle(pkg, ctx, orc, W, mode):
    inst = W.config(1, N=48)
    import bimsim_amd.pipeline as P
    q = P.quad_of(inst)
    N = inst["N"]
    args = (inst["focus"], inst["kdir"], inst["kmag"], inst["amp"], inst["cos_alpha"],
            inst["cos_beta"], inst["m"])
    ref = orc.nf_uf(q, N, N, *args, mode=mode).ravel()
    U64 = torch.empty(N * N, dtype=torch.complex128, device="cuda")
    ctx.nf_uf(q, N, N, *args, U64, mode=mode, prec=pkg.FP64_PARITY)
    assert normrel(host(U64), ref) <= 1e-10
    U32 = torch.empty(N * N, dtype=torch.complex64, device="cuda")
    ctx.nf_uf(q, N, N, *args, U32, mode=mode, prec=pkg.FP32_FAST)
    assert normrel(host(U32), ref) <= 1e-6


@pytest.mark.parametrize("N,nk,z", [(48, 40, 0.0), (70, 333, 0.8), (130, 65, -0.4)])
def test_mc_incident_field_matches_oracle(pkg, ctx, orc, W, N, nk, z):
    """The Monte-Carlo incident field: FP64 per pixel in wave order, FP32 by
    LDS tiles of row and column factors (ragged edges at N = 70, 130; wave
    chunks with a remainder at nk = 333, 65)."""
    import bimsim_amd.pipeline as P
    inst = W.make_instance([], 2.7, N, 26.0, (0.15, 0.55), (0.0, 0.5), nk, 31)
    inst["A"] = (inst["A"][0], inst["A"][1], z)
    q = P.quad_of(inst)
    waves = W.mc_waves(inst, pkg.mc_sample_na)
    focus = (0.4, -0.3, 0.1)
    ref = orc.nf_uf_mc(q, N, N, focus, inst["kmag"], waves, inst["subA"]).ravel()
    wd = dev(waves.ravel(), torch.float64)
    U64 = torch.empty(N * N, dtype=torch.complex128, device="cuda")
    ctx.nf_uf_mc(q, N, N, focus, inst["kmag"], wd, inst["subA"], U64, prec=pkg.FP64_PARITY)
    assert normrel(host(U64), ref) <= 1e-10
    U32 = torch.empty(N * N, dtype=torch.complex64, device="cuda")
    ctx.nf_uf_mc(q, N, N, focus, inst["kmag"], wd, inst["subA"], U32, prec=pkg.FP32_FAST)
    assert normrel(host(U32), ref) <= 2e-6


@pytest.mark.parametrize("fill,N", [(2, 64), (2, 96), (3, 32)])
def test_zero_filled_detector_matches_oracle(pkg, ctx, orc, fill, N):
    """The detector on the field zero-filled to fill x N: the oracle's
    detector on the same embedded arrays, frequency step 1/(fill L)."""
    rng = np.random.default_rng(fill * 100 + N)
    U = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    Uf = 1.0 + 0.1 * (rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N)))
    L, lam, NAin, NAout = 30.0, 2.0, 0.1, 0.5
    M = fill * N
    Ue = np.zeros((M, M), dtype=complex)
    Ufe = np.zeros((M, M), dtype=complex)
    Ue[:N, :N] = U
    Ufe[:N, :N] = Uf
    out = orc.detector(Ue, Ufe, 1.0 / (fill * L), 1.0 / (fill * L), NAin, NAout, lam, 0, N, 1)
    for prec, tol in ((pkg.FP64_PARITY, 1e-10), (pkg.FP32_FAST, 2e-5)):
        cd = torch.complex64 if prec == pkg.FP32_FAST else torch.complex128
        rd = torch.float32 if prec == pkg.FP32_FAST else torch.float64
        Ud, Ufd = dev(U.ravel(), cd), dev(Uf.ravel(), cd)
        D = torch.zeros(N * N, dtype=rd, device="cuda")
        Di, Ab = torch.zeros_like(D), torch.zeros_like(D)
        ctx.detector_zero_filled(Ud, Ufd, N, N, fill, 1.0 / L, 1.0 / L, NAin, NAout, lam, D, Di,
                                 Ab, prec=prec)
        assert normrel(host(D), out[2].ravel()) <= tol
        assert normrel(host(Di), out[3].ravel()) <= tol
        # the inputs are left unchanged
        assert np.array_equal(host(Ud), U.ravel().astype(np.complex64 if prec == 2 else complex))


# ------------------------------------------------------------ scattered field
def _us_case(pkg, W, cid, N, nk):
    inst = W.config(cid, N=N, nk=nk)
    waves = instance_waves(pkg, W, inst)
    return inst, waves


@pytest.mark.parametrize("cid,N,nk", [(1, 32, 24), (2, 48, 16)])
def test_scattered_field_matches_oracle(pkg, ctx, orc, W, cid, N, nk):
    import bimsim_amd.pipeline as P
    inst, waves = _us_case(pkg, W, cid, N, nk)
    q = P.quad_of(inst)
    sph = inst["spheres"]
    A = torch.empty(inst["ncoef"], dtype=torch.complex128, device="cuda")
    B = torch.empty(inst["ncoef"], dtype=torch.complex128, device="cuda")
    ctx.mie_coeffs(sph, inst["kmag"], A, B)
    Ah, Bh = host(A), host(B)
    ref = orc.nf_us(q, N, N, inst["focus"], inst["kmag"], sph, Ah, Bh, waves,
                    inst["subA"]).ravel()
    wd = dev(waves, torch.float64)
    outs = {}
    for prec, dt in [(pkg.FP64_PARITY, torch.complex128), (pkg.FP64, torch.complex128),
                     (pkg.FP32_FAST, torch.complex64)]:
        U = torch.empty(N * N, dtype=dt, device="cuda")
        ctx.nf_us(q, N, N, inst["focus"], inst["kmag"], sph, A, B, wd, inst["subA"], U,
                  prec=prec)
        outs[prec] = host(U)
    assert normrel(outs[pkg.FP64_PARITY], ref) <= 1e-10
    assert np.array_equal(outs[pkg.FP64], outs[pkg.FP64_PARITY])
    assert normrel(outs[pkg.FP32_FAST], ref) <= 2e-5


def test_scattered_field_mixed_orders(pkg, ctx, orc, W):
    """Spheres whose Nl+1 is below 8, a multiple of 8 and ragged, one pixel at
    a sphere centre (calc_Ui's A_0 branch) and pixels inside every sphere."""
    import bimsim_amd.pipeline as P
    sph = [dict(p=(0.0, 0.0, 0.0), a=0.2, n=complex(1.5, 0.05)),
           dict(p=(-6.0, 4.0, 0.0), a=3.0, n=complex(1.33, 0.0)),
           dict(p=(-9.0, -9.0, 0.0), a=2.6, n=complex(1.45, 0.03)),
           dict(p=(7.0, -5.0, 0.5), a=6.0, n=complex(1.6, 0.1)),
           dict(p=(2.0, 9.0, -1.0), a=1.9, n=complex(1.2, 0.02))]
    inst = W.make_instance(sph, 3.0, 32, 30.0, (0.1, 0.5), (0.0, 0.5), 18, 77)
    orders = [s["Nl"] + 1 for s in inst["spheres"]]
    assert min(orders) < 8 and any(o % 8 == 0 for o in orders)
    waves = instance_waves(pkg, W, inst)
    q = P.quad_of(inst)
    N = inst["N"]
    A = torch.empty(inst["ncoef"], dtype=torch.complex128, device="cuda")
    B = torch.empty(inst["ncoef"], dtype=torch.complex128, device="cuda")
    ctx.mie_coeffs(inst["spheres"], inst["kmag"], A, B)
    Ah, Bh = host(A), host(B)
    ref = orc.nf_us(q, N, N, inst["focus"], inst["kmag"], inst["spheres"], Ah, Bh, waves,
                    inst["subA"]).ravel()
    wd = dev(waves, torch.float64)
    for prec, dt, tol in [(pkg.FP64_PARITY, torch.complex128, 1e-10),
                          (pkg.FP32_FAST, torch.complex64, 2e-5)]:
        U = torch.empty(N * N, dtype=dt, device="cuda")
        ctx.nf_us(q, N, N, inst["focus"], inst["kmag"], inst["spheres"], A, B, wd, inst["subA"],
                  U, prec=prec)
        assert normrel(host(U), ref) <= tol, (prec, orders)


def test_scattered_field_accumulates_and_single_plane_wave(pkg, ctx, orc, W):
    import bimsim_amd.pipeline as P
    inst, _ = _us_case(pkg, W, 1, 32, 1)
    q = P.quad_of(inst)
    N = inst["N"]
    sph = inst["spheres"]
    A = torch.empty(inst["ncoef"], dtype=torch.complex128, device="cuda")
    B = torch.empty(inst["ncoef"], dtype=torch.complex128, device="cuda")
    ctx.mie_coeffs(sph, inst["kmag"], A, B)
    Ah, Bh = host(A), host(B)
    # gpuScalarUsp path: one wave along k with amplitude A (nfScalarUs.cu:547-561)
    waves = np.array([[0.0, 0.0, 1.0]])
    base = np.random.default_rng(3).normal(size=(N * N, 2)).view(np.complex128).ravel()
    ref = orc.nf_us(q, N, N, inst["focus"], inst["kmag"], sph, Ah, Bh, waves, 1.0,
                    U=base.reshape(N, N)).ravel()
    for prec, dt, tol in [(pkg.FP64_PARITY, torch.complex128, 1e-10),
                          (pkg.FP32_FAST, torch.complex64, 2e-5)]:
        U = dev(base, dt)
        ctx.nf_us(q, N, N, inst["focus"], inst["kmag"], sph, A, B, dev(waves, torch.float64),
                  1.0, U, prec=prec, accumulate=True)
        assert normrel(host(U), ref) <= tol


def test_sum_uf_matches_oracle(pkg, ctx, orc, W):
    import bimsim_amd.pipeline as P
    inst = W.config(2, N=40, nk=4)
    q = P.quad_of(inst)
    N = inst["N"]
    rng = np.random.default_rng(11)
    Uf = rng.normal(size=(N * N, 2)).view(np.complex128).ravel()
    U0 = rng.normal(size=(N * N, 2)).view(np.complex128).ravel()
    ref = orc.nf_sum_uf(q, N, N, inst["spheres"], Uf.reshape(N, N), U0.reshape(N, N)).ravel()
    for prec, dt in [(pkg.FP64_PARITY, torch.complex128), (pkg.FP32_FAST, torch.complex64)]:
        U = dev(U0, dt)
        ctx.nf_sum_uf(q, N, N, inst["spheres"], dev(Uf, dt), U, prec=prec)
        tol = 0 if dt == torch.complex128 else 1e-6
        assert normrel(host(U), ref) <= tol


# ------------------------------------------------------------------ detector
@pytest.mark.parametrize("N", [64, 96, 80, 128])
def test_detector_matches_oracle(pkg, ctx, orc, N):
    rng = np.random.default_rng(N)
    # smooth fields: a few plane waves plus a bump, so the band-pass matters
    y, x = np.mgrid[0:N, 0:N] / N
    U = sum(np.exp(2j * np.pi * (a * x + b * y)) * c for a, b, c in
            [(1, 2, 1.0), (7, -3, 0.5), (N // 3, 5, 0.3)])
    U = U + 0.2 * np.exp(-((x - 0.5) ** 2 + (y - 0.4) ** 2) / 0.01)
    Uf = np.exp(2j * np.pi * (2 * x + 1 * y)) + 0.5 + 0.1 * rng.normal(size=(N, N))
    L, lam, NAin, NAout = 40.0, 2.5, 0.05, 0.5
    du = dv = 1.0 / L
    _, _, D, Di, Ab = orc.detector(U, Uf, du, dv, NAin, NAout, lam, 0, N, 1)
    for prec, cd, rd, tol in [(pkg.FP64_PARITY, torch.complex128, torch.float64, 1e-10),
                              (pkg.FP32_FAST, torch.complex64, torch.float32, 2e-5)]:
        Ug, Ufg = dev(U.ravel(), cd), dev(Uf.ravel(), cd)
        Dg = torch.zeros(N * N, dtype=rd, device="cuda")
        Dig = torch.zeros(N * N, dtype=rd, device="cuda")
        Ag = torch.empty(N * N, dtype=rd, device="cuda")
        ctx.detector(Ug, Ufg, N, N, du, dv, NAin, NAout, lam, 0, N, 1, Dg, Dig, Ag, prec=prec)
        assert normrel(host(Dg), D.ravel()) <= tol
        assert normrel(host(Dig), Di.ravel()) <= tol
        bright = Di.ravel() > 1e-3 * Di.max()
        assert np.max(np.abs(host(Ag)[bright] - Ab.ravel()[bright])) <= (1e-9 if tol < 1e-9 else 1e-4)


def test_detector_padding_and_supersampling(pkg, ctx, orc):
    # padding p = 1 and ss = 2: near field (2p+1)*Nd*ss = 96, crop at p*Nd*ss
    Nd, ss, p = 16, 2, 1
    N = (2 * p + 1) * Nd * ss
    rng = np.random.default_rng(7)
    U = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    Uf = 1.0 + 0.1 * (rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N)))
    D0 = rng.uniform(0, 1, size=Nd * Nd)
    Di0 = rng.uniform(1, 2, size=Nd * Nd)
    Uo, Ufo, D, Di, Ab = orc.detector(U, Uf, 0.02, 0.02, 0.0, 0.4, 3.0, p, Nd, ss, D=D0, Di=Di0)
    Ug = dev(U.ravel(), torch.complex128)
    Ufg = dev(Uf.ravel(), torch.complex128)
    Dg = dev(D0, torch.float64)
    Dig = dev(Di0, torch.float64)
    Ag = torch.empty(Nd * Nd, dtype=torch.float64, device="cuda")
    ctx.detector(Ug, Ufg, N, N, 0.02, 0.02, 0.0, 0.4, 3.0, p, Nd, ss, Dg, Dig, Ag,
                 prec=pkg.FP64_PARITY, write_filtered=True)
    assert normrel(host(Dg), D.ravel()) <= 1e-10
    assert normrel(host(Dig), Di.ravel()) <= 1e-10
    assert normrel(host(Ug), Uo.ravel()) <= 1e-10   # band-passed field left in U
    assert np.max(np.abs(host(Ag) - Ab.ravel())) <= 1e-9


# ------------------------------------------------------------ whole pipeline
def oracle_pipeline(orc, inst, waves, A, B, fp32=0):
    import bimsim_amd.pipeline as P
    q = P.quad_of(inst)
    N = inst["N"]
    Uf = orc.nf_uf(q, N, N, inst["focus"], inst["kdir"], inst["kmag"], inst["amp"],
                   inst["cos_alpha"], inst["cos_beta"], inst["m"], fp32=fp32)
    Us = orc.nf_us(q, N, N, inst["focus"], inst["kmag"], inst["spheres"], A, B, waves,
                   inst["subA"], fp32=fp32)
    U = orc.nf_sum_uf(q, N, N, inst["spheres"], Uf, Us, fp32=fp32)
    return orc.detector(U, Uf, inst["du"], inst["dv"], inst["obj"][0], inst["obj"][1],
                        inst["lam"], inst["padding"], inst["Nd"], inst["ss"], fp32=fp32)


@pytest.mark.parametrize("cid,N,nk", [(1, 64, 40), (2, 64, 32)])
def test_pipeline_absorbance(pkg, ctx, orc, W, cid, N, nk):
    import bimsim_amd.pipeline as P
    inst = W.config(cid, N=N, nk=nk)
    waves = instance_waves(pkg, W, inst)
    wd = dev(waves, torch.float64)
    b64 = P.simulate(ctx, inst, wd, pkg.FP64_PARITY)
    A, B = host(b64.A), host(b64.B)
    _, _, D, Di, Ab = oracle_pipeline(orc, inst, waves, A, B)
    assert normrel(host(b64.D), D.ravel()) <= 1e-10
    assert normrel(host(b64.Di), Di.ravel()) <= 1e-10
    bright = Di.ravel() > 1e-3 * Di.max()
    assert np.max(np.abs(host(b64.Abs)[bright] - Ab.ravel()[bright])) <= 1e-9
    b32 = P.simulate(ctx, inst, wd, pkg.FP32_FAST)
    assert np.max(np.abs(host(b32.Abs)[bright] - Ab.ravel()[bright])) <= 1e-4
