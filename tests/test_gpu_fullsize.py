"""Parity with the oracle at the workloads' full sizes (through the C ABI):

  cfg1   one sphere, 128 x 128, 400 waves: the whole image (incident series,
         Mie coefficients, scattered series, sum, detector), with the
         reference detector and with the 2x zero-filled one;
  cfg2   three spheres, 512 x 512, 1000 waves: the whole image;
  cfg3   the sphere-table path on the 50-sphere workload's own spheres and
         tables (4096 radii x 1000 angles), every fifth sphere, at N = 128.

Tolerances: FP64_PARITY images <= 1e-10 normwise and |dA| <= 1e-9; FP32
absorbance |dA| <= 1e-4 where the incident image is above 1e-3 of its peak;
tables and lookups as in test_gpu_lut (FP64 <= 1e-10, FP32 <= 2e-5)."""
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


def oracle_image(orc, inst, waves, fill=0):
    """The oracle composed as microscopeStruct::Simulate; with fill >= 2 the
    detector runs on the fields embedded in fill x larger zero arrays."""
    import bimsim_amd.pipeline as P
    q = P.quad_of(inst)
    N = inst["N"]
    Uf = orc.nf_uf(q, N, N, inst["focus"], inst["kdir"], inst["kmag"], inst["amp"],
                   inst["cos_alpha"], inst["cos_beta"], inst["m"])
    sph = inst["spheres"]
    A = np.zeros(inst["ncoef"], dtype=complex)
    B = np.zeros(inst["ncoef"], dtype=complex)
    for s in sph:
        a, b = orc.mie_coeffs(s["Nl"], inst["kmag"], s["a"], s["n"])
        A[s["coeff_off"]:s["coeff_off"] + s["Nl"] + 1] = a
        B[s["coeff_off"]:s["coeff_off"] + s["Nl"] + 1] = b
    Us = orc.nf_us(q, N, N, inst["focus"], inst["kmag"], sph, A, B, waves, inst["subA"])
    U = orc.nf_sum_uf(q, N, N, sph, Uf, Us)
    du, dv = inst["du"], inst["dv"]
    if fill >= 2:
        M = fill * N
        Ue = np.zeros((M, M), dtype=complex)
        Ufe = np.zeros((M, M), dtype=complex)
        Ue[:N, :N], Ufe[:N, :N] = U, Uf
        U, Uf, du, dv = Ue, Ufe, du / fill, dv / fill
    out = orc.detector(U, Uf, du, dv, inst["obj"][0], inst["obj"][1], inst["lam"],
                       inst["padding"], inst["Nd"], inst["ss"])
    return out[2].ravel(), out[3].ravel(), out[4].ravel()


def check_image(pkg, ctx, inst, waves, want):
    import bimsim_amd.pipeline as P
    Dw, Diw, Abw = want
    bright = Diw > 1e-3 * Diw.max()
    wd = torch.as_tensor(waves).cuda()
    b64 = P.simulate(ctx, inst, wd, pkg.FP64_PARITY)
    assert normrel(host(b64.D), Dw) <= 1e-10 and normrel(host(b64.Di), Diw) <= 1e-10
    assert np.max(np.abs(host(b64.Abs)[bright] - Abw[bright])) <= 1e-9
    b32 = P.simulate(ctx, inst, wd, pkg.FP32_FAST)
    err = float(np.max(np.abs(host(b32.Abs).astype(np.float64)[bright] - Abw[bright])))
    assert err <= 1e-4, err


@pytest.mark.parametrize("fill", [0, 2])
def test_cfg1_full_image(pkg, ctx, orc, W, fill):
    inst = W.config(1)
    assert inst["N"] == 128 and inst["nk"] == 400
    inst["zero_fill"] = fill
    waves = W.mc_waves(inst, pkg.mc_sample_na)
    check_image(pkg, ctx, inst, waves, oracle_image(orc, inst, waves, fill))


def test_cfg2_full_image(pkg, ctx, orc, W):
    inst = W.config(2)
    assert inst["N"] == 512 and inst["nk"] == 1000 and len(inst["spheres"]) == 3
    waves = W.mc_waves(inst, pkg.mc_sample_na)
    check_image(pkg, ctx, inst, waves, oracle_image(orc, inst, waves))


@pytest.mark.parametrize("prec", [0, 2])
def test_cfg3_table_path_on_its_workload(pkg, ctx, orc, W, prec):
    import bimsim_amd.pipeline as P
    inst = W.config3(N=128)
    spheres = inst["spheres"][::5]
    dR_fix, thetaR = inst["lut"]["dR"], inst["lut"]["thetaR"]
    assert dR_fix == 4096 and thetaR == 1000 and len(spheres) == 10
    q = P.quad_of(inst)
    N = inst["N"]
    waves = W.mc_waves(inst, pkg.mc_sample_na)
    ct = torch.complex64 if prec == pkg.FP32_FAST else torch.complex128
    A = torch.zeros(inst["ncoef"], dtype=torch.complex128, device="cuda")
    B = torch.zeros_like(A)
    ctx.mie_coeffs(spheres, inst["kmag"], A, B)
    Ah, Bh = host(A), host(B)
    wd = torch.as_tensor(np.ascontiguousarray(waves).ravel()).cuda()
    U = torch.zeros(N * N, dtype=ct, device="cuda")
    ref = np.zeros((N, N), dtype=complex)
    for s in spheres:
        aR, dR, dmin, dmax = pkg.sphere_lut_dims(q, s, N, dR_fix)
        assert dR == dR_fix
        sl = slice(s["coeff_off"], s["coeff_off"] + s["Nl"] + 1)
        Usp_o = orc.lut_eval(s["Nl"], Bh[sl], orc.lut_hankel(s["Nl"], inst["kmag"], s["a"], dmin,
                                                             dmax, dR), thetaR)
        Uip_o = orc.lut_eval(s["Nl"], Ah[sl], orc.lut_bessel(s["Nl"], inst["kmag"], s["a"],
                                                             s["n"], aR), thetaR)
        ref = orc.nf_us_lut(q, N, N, inst["focus"], inst["kmag"], s, Usp_o, Uip_o, dmin, dmax,
                            waves, inst["subA"], U=ref)
        Usp = torch.empty(thetaR * dR, dtype=ct, device="cuda")
        Uip = torch.empty(thetaR * aR, dtype=ct, device="cuda")
        ctx.sphere_lut(s, A, B, inst["kmag"], dmin, dmax, dR, aR, thetaR, Usp, Uip, prec=prec)
        if prec == pkg.FP64_PARITY:
            assert normrel(host(Usp), Usp_o.ravel()) <= 1e-10
        ctx.nf_us_lut(q, N, N, inst["focus"], inst["kmag"], s, Usp, dR, Uip, aR, thetaR, dmin, dmax,
                      wd, inst["subA"], U, prec=prec)
    tol = 1e-10 if prec == pkg.FP64_PARITY else 2e-5
    assert normrel(host(U), ref.ravel()) <= tol
