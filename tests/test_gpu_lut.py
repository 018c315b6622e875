"""Parity of the look-up-table entry points and the rotation resample with
the oracle (through the C ABI).  Tolerances (normwise, as in the direct-path
tests): FP64_PARITY <= 1e-10; FP32_FAST (FP64 tables stored as float,
float lookups) <= 2e-5 against the FP64 oracle."""
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


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.fixture(scope="module")
def W(pkg):
    import bimsim_amd.workloads as W
    return W


TOL = {0: 1e-10, 2: 2e-5}
CT = {0: torch.complex128, 2: torch.complex64}
RT = {0: torch.float64, 2: torch.float32}


@pytest.mark.parametrize("prec", [0, 2])
def test_incident_bessel_table_and_lookup(pkg, ctx, orc, W, prec):
    import bimsim_amd.pipeline as P
    inst = W.config(1, N=40)
    q = P.quad_of(inst)
    N = inst["N"]
    m = inst["m"]
    dmin, dmax = orc.quad_dist(q, inst["focus"])
    dR = 2 * N
    want = orc.uf_bessel_lut(m, inst["kmag"], dmin, dmax, dR)
    lut = torch.empty(dR * (m + 1), dtype=RT[prec], device="cuda")
    ctx.uf_bessel_lut(m, inst["kmag"], dmin, dmax, dR, lut, prec=prec)
    got = host(lut).astype(np.float64)
    assert normrel(got, want) <= (1e-12 if prec == 0 else 1e-7)
    args = (inst["focus"], inst["kdir"], inst["kmag"], inst["amp"], inst["cos_alpha"],
            inst["cos_beta"], m)
    ref = orc.nf_uf(q, N, N, *args, mode=2, lut=want, dR=dR, dmin=dmin, dmax=dmax).ravel()
    Uf = torch.empty(N * N, dtype=CT[prec], device="cuda")
    ctx.nf_uf(q, N, N, *args, Uf, mode=pkg.UF_LUT, prec=prec)
    assert normrel(host(Uf), ref) <= TOL[prec]
    # the table-taking entry point with the same table
    Uf2 = torch.empty_like(Uf)
    a2 = (inst["focus"], inst["kdir"], inst["amp"], inst["cos_alpha"], inst["cos_beta"], m)
    ctx.nf_uf_lut(q, N, N, *a2, lut, dR, dmin, dmax, Uf2, prec=prec)
    assert torch.equal(Uf, Uf2)


def _sphere_case(pkg, W, N=32, nk=12):
    sph = [dict(p=(1.5, -2.0, 0.5), a=4.0, n=complex(1.4, 0.03))]
    inst = W.make_instance(sph, 3.0, N, 24.0, (0.0, 0.5), (0.0, 0.5), nk, 11)
    return inst, inst["spheres"][0]


@pytest.mark.parametrize("prec", [0, 2])
def test_sphere_tables_match_oracle(pkg, ctx, orc, W, prec):
    import bimsim_amd.pipeline as P
    inst, s = _sphere_case(pkg, W)
    q = P.quad_of(inst)
    aR, dR, dmin, dmax = pkg.sphere_lut_dims(q, s, inst["N"])
    thetaR = 90
    A, B = orc.mie_coeffs(s["Nl"], inst["kmag"], s["a"], s["n"])
    h = orc.lut_hankel(s["Nl"], inst["kmag"], s["a"], dmin, dmax, dR)
    Usp_want = orc.lut_eval(s["Nl"], B, h, thetaR)
    j = orc.lut_bessel(s["Nl"], inst["kmag"], s["a"], s["n"], aR)
    Uip_want = orc.lut_eval(s["Nl"], A, j, thetaR)
    Ad = torch.as_tensor(A).cuda()
    Bd = torch.as_tensor(B).cuda()
    sd = dict(s, coeff_off=0)
    Usp = torch.empty(thetaR * dR, dtype=CT[prec], device="cuda")
    Uip = torch.empty(thetaR * aR, dtype=CT[prec], device="cuda")
    ctx.sphere_lut(sd, Ad, Bd, inst["kmag"], dmin, dmax, dR, aR, thetaR, Usp, Uip, prec=prec)
    tol = 1e-10 if prec == 0 else 1e-6
    assert normrel(host(Usp).reshape(thetaR, dR), Usp_want) <= tol
    assert normrel(host(Uip).reshape(thetaR, aR), Uip_want) <= tol


@pytest.mark.parametrize("prec", [0, 2])
def test_scattered_field_from_tables(pkg, ctx, orc, W, prec):
    """gpuScalarUpLut with the oracle's tables uploaded: isolates the lookup."""
    import bimsim_amd.pipeline as P
    inst, s = _sphere_case(pkg, W)
    q = P.quad_of(inst)
    N = inst["N"]
    waves = W.mc_waves(inst, pkg.mc_sample_na)
    aR, dR, dmin, dmax = pkg.sphere_lut_dims(q, s, N)
    thetaR = 120
    A, B = orc.mie_coeffs(s["Nl"], inst["kmag"], s["a"], s["n"])
    Usp = orc.lut_eval(s["Nl"], B, orc.lut_hankel(s["Nl"], inst["kmag"], s["a"], dmin, dmax, dR),
                       thetaR)
    Uip = orc.lut_eval(s["Nl"], A, orc.lut_bessel(s["Nl"], inst["kmag"], s["a"], s["n"], aR),
                       thetaR)
    ref = orc.nf_us_lut(q, N, N, inst["focus"], inst["kmag"], s, Usp, Uip, dmin, dmax, waves,
                        inst["subA"]).ravel()
    Usp_d = torch.as_tensor(Usp.ravel()).to(CT[prec]).cuda()
    Uip_d = torch.as_tensor(Uip.ravel()).to(CT[prec]).cuda()
    w = torch.as_tensor(np.ascontiguousarray(waves, dtype=np.float64).ravel()).cuda()
    U = torch.zeros(N * N, dtype=CT[prec], device="cuda")
    ctx.nf_us_lut(q, N, N, inst["focus"], inst["kmag"], s, Usp_d, dR, Uip_d, aR, thetaR, dmin,
                  dmax, w, inst["subA"], U, prec=prec)
    assert normrel(host(U), ref) <= TOL[prec]
    # accumulation: a second call adds the same field
    ctx.nf_us_lut(q, N, N, inst["focus"], inst["kmag"], s, Usp_d, dR, Uip_d, aR, thetaR, dmin,
                  dmax, w, inst["subA"], U, prec=prec)
    assert normrel(host(U), 2 * ref) <= TOL[prec]


def _fp32_lookup_cases(W):
    out = {}
    # a pixel at the sphere centre (r-hat = 0/0 there)
    out["centre"] = W.make_instance([dict(p=(0.0, 0.0, 0.0), a=3.0, n=complex(1.4, 0.03))], 3.0,
                                    24, 24.0, (0.0, 0.5), (0.0, 0.5), 16, 5)
    # more waves than one workgroup stages at a time
    out["many_waves"] = W.make_instance([dict(p=(2.0, 1.0, 0.5), a=4.0, n=complex(1.5, 0.0))],
                                        3.0, 16, 24.0, (0.0, 0.5), (0.0, 0.5), 1500, 7)
    # tilted illumination: cosines down to -1 (the last table row)
    t = W.make_instance([dict(p=(-3.0, 2.0, -1.0), a=3.5, n=complex(1.3, 0.05))], 2.5, 24,
                        24.0, (0.1, 0.7), (0.0, 0.7), 32, 9)
    t["kdir"] = (math.sin(1.2), 0.0, math.cos(1.2))
    out["tilted"] = t
    # an inner table one column wide
    out["one_column"] = W.make_instance([dict(p=(1.0, 1.0, 0.0), a=0.3, n=complex(1.4, 0.0))],
                                        3.0, 16, 40.0, (0.0, 0.5), (0.0, 0.5), 12, 3)
    return out


@pytest.mark.parametrize("name", ["centre", "many_waves", "tilted", "one_column"])
def test_fp32_lookup_edges(pkg, ctx, orc, W, name):
    """FP32 lookups (tables stored as float) against the FP64 oracle on the
    cases with their own code: the sphere centre, wave staging, angles at the
    table's last row, and a one-column inner table."""
    import bimsim_amd.pipeline as P
    inst = _fp32_lookup_cases(W)[name]
    s = inst["spheres"][0]
    q = P.quad_of(inst)
    N = inst["N"]
    waves = W.mc_waves(inst, pkg.mc_sample_na)
    aR, dR, dmin, dmax = pkg.sphere_lut_dims(q, s, N)
    if name == "one_column":
        assert aR == 1
    else:
        assert aR >= 2
    thetaR = 240
    A, B = orc.mie_coeffs(s["Nl"], inst["kmag"], s["a"], s["n"])
    Usp = orc.lut_eval(s["Nl"], B, orc.lut_hankel(s["Nl"], inst["kmag"], s["a"], dmin, dmax, dR),
                       thetaR)
    Uip = orc.lut_eval(s["Nl"], A, orc.lut_bessel(s["Nl"], inst["kmag"], s["a"], s["n"], aR),
                       thetaR)
    ref = orc.nf_us_lut(q, N, N, inst["focus"], inst["kmag"], s, Usp, Uip, dmin, dmax, waves,
                        inst["subA"]).ravel()
    if name == "centre":
        iu = iv = N // 2
        assert abs(ref[iv * N + iu]) > 0
    w = torch.as_tensor(np.ascontiguousarray(waves, dtype=np.float64).ravel()).cuda()
    U = torch.zeros(N * N, dtype=torch.complex64, device="cuda")
    ctx.nf_us_lut(q, N, N, inst["focus"], inst["kmag"], s,
                  torch.as_tensor(Usp.ravel()).to(torch.complex64).cuda(), dR,
                  torch.as_tensor(Uip.ravel()).to(torch.complex64).cuda(), aR, thetaR, dmin, dmax,
                  w, inst["subA"], U, prec=pkg.FP32_FAST)
    got = host(U)
    assert np.all(np.isfinite(got))
    assert normrel(got, ref) <= TOL[2]


def test_table_path_end_to_end(pkg, ctx, orc, W):
    """Tables built on the GPU then looked up on the GPU, against the oracle
    doing both steps."""
    import bimsim_amd.pipeline as P
    inst, s = _sphere_case(pkg, W, N=48, nk=20)
    q = P.quad_of(inst)
    N = inst["N"]
    waves = W.mc_waves(inst, pkg.mc_sample_na)
    aR, dR, dmin, dmax = pkg.sphere_lut_dims(q, s, N)
    thetaR = 200
    A, B = orc.mie_coeffs(s["Nl"], inst["kmag"], s["a"], s["n"])
    Usp = orc.lut_eval(s["Nl"], B, orc.lut_hankel(s["Nl"], inst["kmag"], s["a"], dmin, dmax, dR),
                       thetaR)
    Uip = orc.lut_eval(s["Nl"], A, orc.lut_bessel(s["Nl"], inst["kmag"], s["a"], s["n"], aR),
                       thetaR)
    ref = orc.nf_us_lut(q, N, N, inst["focus"], inst["kmag"], s, Usp, Uip, dmin, dmax, waves,
                        inst["subA"]).ravel()
    Ad = torch.zeros(s["Nl"] + 1, dtype=torch.complex128, device="cuda")
    Bd = torch.zeros_like(Ad)
    sd = dict(s, coeff_off=0)
    ctx.mie_coeffs([sd], inst["kmag"], Ad, Bd)
    Usp_d = torch.empty(thetaR * dR, dtype=torch.complex128, device="cuda")
    Uip_d = torch.empty(thetaR * aR, dtype=torch.complex128, device="cuda")
    ctx.sphere_lut(sd, Ad, Bd, inst["kmag"], dmin, dmax, dR, aR, thetaR, Usp_d, Uip_d)
    w = torch.as_tensor(np.ascontiguousarray(waves, dtype=np.float64).ravel()).cuda()
    U = torch.zeros(N * N, dtype=torch.complex128, device="cuda")
    ctx.nf_us_lut(q, N, N, inst["focus"], inst["kmag"], sd, Usp_d, dR, Uip_d, aR, thetaR, dmin,
                  dmax, w, inst["subA"], U)
    assert normrel(host(U), ref) <= 1e-10


@pytest.mark.parametrize("prec", [0, 2])
@pytest.mark.parametrize("N,deg", [(32, 0.3), (33, 1.0), (40, -2.2), (24, math.pi / 2)])
def test_rotation_resample(pkg, ctx, orc, prec, N, deg):
    rng = np.random.default_rng(N)
    U = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    Uf = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    want_U = orc.nf_interpolate(U, deg, centre_float=0)
    want_Uf = orc.nf_interpolate(Uf, deg, centre_float=1)
    Us = torch.as_tensor(U.ravel()).to(CT[prec]).cuda()
    Ufs = torch.as_tensor(Uf.ravel()).to(CT[prec]).cuda()
    Uo = torch.empty_like(Us)
    Ufo = torch.empty_like(Ufs)
    ctx.nf_interpolate(deg, N, N, Us, Ufs, Uo, Ufo, prec=prec)
    tol = 1e-12 if prec == 0 else 1e-5
    assert normrel(host(Uo).reshape(N, N), want_U) <= tol
    assert normrel(host(Ufo).reshape(N, N), want_Uf) <= tol


def test_radial_tables_match_oracle(pkg, ctx, orc, W):
    import bimsim_amd.pipeline as P
    inst, s = _sphere_case(pkg, W)
    q = P.quad_of(inst)
    aR, dR, dmin, dmax = pkg.sphere_lut_dims(q, s, inst["N"])
    nl = s["Nl"]
    h = torch.empty(dR * (nl + 1), dtype=torch.complex128, device="cuda")
    j = torch.empty(aR * (nl + 1), dtype=torch.complex128, device="cuda")
    ctx.sphere_radial_lut(dict(s, coeff_off=0), inst["kmag"], dmin, dmax, dR, aR, h, j)
    want_h = orc.lut_hankel(nl, inst["kmag"], s["a"], dmin, dmax, dR)
    want_j = orc.lut_bessel(nl, inst["kmag"], s["a"], s["n"], aR)
    assert normrel(host(h).reshape(dR, nl + 1), want_h) <= 1e-12
    assert normrel(host(j).reshape(aR, nl + 1), want_j) <= 1e-12
