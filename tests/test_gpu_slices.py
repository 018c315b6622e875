"""The imaging steps one at a time (fieldslice / scalarslice / detector
building blocks) against their definitions and the oracle: 2D FFT by the
DFT (numpy.fft), band-pass + crop + integration against the oracle
detector, and the element-wise views and ratios by their formulas."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def normrel(a, b):
    m = np.max(np.abs(b))
    return float(np.max(np.abs(np.asarray(a) - b)) / (m if m > 0 else 1.0))


CT = {0: torch.complex128, 2: torch.complex64}
RT = {0: torch.float64, 2: torch.float32}
TOL = {0: 1e-12, 2: 2e-6}


def field(N, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))


@pytest.mark.parametrize("prec", [0, 2])
@pytest.mark.parametrize("N", [48, 96, 120])
def test_fft_is_the_dft(pkg, ctx, prec, N):
    U = field(N, N)
    F = torch.as_tensor(U.ravel()).to(CT[prec]).cuda()
    ctx.field_fft(F, N, N, prec=prec)
    assert normrel(host(F).reshape(N, N), np.fft.fft2(U)) <= 10 * TOL[prec]
    ctx.field_fft(F, N, N, inverse=True, prec=prec)
    assert normrel(host(F).reshape(N, N), U) <= 10 * TOL[prec]


@pytest.mark.parametrize("prec", [0, 2])
def test_stepwise_imaging_matches_oracle_detector(pkg, ctx, orc, prec):
    """applyBandpass + getFarField + integrateDetector + getAbsorbance as
    separate calls equal the oracle's detector."""
    Nd, pad = 32, 1
    N = Nd * (2 * pad + 1)
    U, Uf = field(N, 1), field(N, 2) + 3.0
    du = dv = 1.0 / 90.0
    lam, NAin, NAout = 2.5, 0.1, 0.5
    want = orc.detector(U, Uf, du, dv, NAin, NAout, lam, pad, Nd, 1)
    D = torch.zeros(Nd * Nd, dtype=RT[prec], device="cuda")
    Di = torch.zeros_like(D)
    for src, dst in [(U, D), (Uf, Di)]:
        F = torch.as_tensor(src.ravel()).to(CT[prec]).cuda()
        ctx.field_fft(F, N, N, prec=prec)
        ctx.field_bandpass(F, N, N, du, dv, NAin, NAout, lam, prec=prec)
        ctx.field_fft(F, N, N, inverse=True, prec=prec)
        C = torch.empty(Nd * Nd, dtype=CT[prec], device="cuda")
        ctx.field_crop(F, N, N, pad * Nd, pad * Nd, Nd, Nd, C, prec=prec)
        ctx.detector_resample(C, Nd, Nd, 1, 1.0, True, dst, prec=prec)
    Ab = torch.empty_like(D)
    ctx.slice_ratio(pkg.RATIO_ABSORBANCE, D, Di, Ab, prec=prec)
    tol = 1e-10 if prec == 0 else 1e-5
    assert normrel(host(D).reshape(Nd, Nd), want[2]) <= tol
    assert normrel(host(Di).reshape(Nd, Nd), want[3]) <= tol
    assert np.max(np.abs(host(Ab).reshape(Nd, Nd) - want[4])) <= (1e-9 if prec == 0 else 1e-4)


@pytest.mark.parametrize("prec", [0, 2])
def test_views_scale_resample_ratio(pkg, ctx, prec):
    N, ss = 24, 3
    U = field(N, 7)
    F = torch.as_tensor(U.ravel()).to(CT[prec]).cuda()
    S = torch.empty(N * N, dtype=RT[prec], device="cuda")
    tol = TOL[prec]
    for op, want in [(pkg.SLICE_INTENSITY, np.abs(U) ** 2), (pkg.SLICE_MAGNITUDE, np.abs(U)),
                     (pkg.SLICE_REAL, U.real), (pkg.SLICE_IMAG, U.imag)]:
        ctx.field_to_scalar(F, op, S, prec=prec)
        assert normrel(host(S).reshape(N, N), want) <= tol
    ctx.field_scale(F, 0.25, prec=prec)
    assert normrel(host(F).reshape(N, N), 0.25 * U) <= tol
    # supersampled detector: mean of |F|^2 over ss x ss blocks
    Dn = N // ss
    blocks = (np.abs(0.25 * U) ** 2).reshape(Dn, ss, Dn, ss).mean(axis=(1, 3))
    D = torch.ones(Dn * Dn, dtype=RT[prec], device="cuda")
    ctx.detector_resample(F, Dn, Dn, ss, 2.0, True, D, prec=prec)
    assert normrel(host(D).reshape(Dn, Dn), 1.0 + 2.0 * blocks) <= tol
    ctx.detector_resample(F, Dn, Dn, ss, 2.0, False, D, prec=prec)
    assert normrel(host(D).reshape(Dn, Dn), blocks) <= tol
    Di = torch.full_like(D, 0.5)
    T = torch.empty_like(D)
    ctx.slice_ratio(pkg.RATIO_TRANSMITTANCE, D, Di, T, prec=prec)
    assert normrel(host(T).reshape(Dn, Dn), blocks / 0.5) <= tol
    ctx.slice_ratio(pkg.RATIO_ABSORBANCE, D, Di, T, prec=prec)
    assert np.max(np.abs(host(T).reshape(Dn, Dn) + np.log10(blocks / 0.5))) <= 10 * tol


@pytest.mark.parametrize("prec", [0, 2])
@pytest.mark.parametrize("quirks", [0, 1])
@pytest.mark.parametrize("N,lam,d", [(48, 2.45, 2.5), (30, 3.0, -1.25), (64, 0.8, 7.0)])
def test_backpropagate(pkg, ctx, orc, prec, quirks, N, lam, d):
    """microscopeStruct::applyBackpropagation's kernel (microscope.cu:423-489)
    on a spectrum, against the oracle; the FFTs around it are the ones
    tested above."""
    L = 30.0
    F0 = field(N, N + 3)
    want = orc.backpropagate(F0, 1.0 / L, 1.0 / L, lam, d, quirks=quirks,
                             fp32=1 if prec == 2 else 0)
    F = torch.as_tensor(F0.ravel()).to(CT[prec]).cuda()
    ctx.field_backpropagate(F, N, N, 1.0 / L, 1.0 / L, lam, d, prec=prec, quirks=quirks)
    assert normrel(host(F).reshape(N, N), want) <= TOL[prec]
