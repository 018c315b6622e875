"""The allocation-free detector (bimsim_detector_noalloc) against
bimsim_detector on the same inputs, through the C ABI: two calls in a row,
the second on a smaller field, FP64 and FP32; the work memory limit is
reported as an error.  Tolerance: the twiddles are evaluated on the device
instead of the host (FP64 <= 1e-12 normwise, FP32 <= 1e-6)."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def normrel(a, b):
    m = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (m if m > 0 else 1.0))


def fields(N, seed, cd):
    rng = np.random.default_rng(seed)
    U = rng.normal(size=N * N) + 1j * rng.normal(size=N * N)
    Uf = 1.0 + 0.2 * (rng.normal(size=N * N) + 1j * rng.normal(size=N * N))
    return torch.as_tensor(U).to(cd).cuda(), torch.as_tensor(Uf).to(cd).cuda()


@pytest.mark.parametrize("prec", [0, 2])
def test_noalloc_detector_matches_the_detector(pkg, prec):
    ctx = pkg.Context(0)
    try:
        cd = torch.complex64 if prec == pkg.FP32_FAST else torch.complex128
        rd = torch.float32 if prec == pkg.FP32_FAST else torch.float64
        tol = 1e-6 if prec == pkg.FP32_FAST else 1e-12
        # (N, padding, Nd, ss): a large call, then a smaller one reusing the memory
        for N, pad, Nd, ss, seed in [(512, 0, 512, 1, 1), (192, 1, 32, 2, 2), (128, 0, 128, 1, 3)]:
            U, Uf = fields(N, seed, cd)
            U0, Uf0 = U.clone(), Uf.clone()
            out = []
            for fn in (ctx.detector, ctx.detector_noalloc):
                D = torch.full((Nd * Nd,), 0.25, dtype=rd, device="cuda")
                Di = torch.full_like(D, 0.5)
                Ab = torch.empty_like(D)
                kw = dict(write_filtered=False) if fn == ctx.detector else {}
                fn(U, Uf, N, N, 1.0 / 30.0, 1.0 / 30.0, 0.1, 0.55, 2.0, pad, Nd, ss, D, Di, Ab,
                   prec=prec, **kw)
                torch.cuda.synchronize()
                out.append((D.cpu().numpy(), Di.cpu().numpy(), Ab.cpu().numpy()))
            for a, b in zip(out[1], out[0]):
                assert normrel(a, b) <= tol, (N, normrel(a, b))
            assert torch.equal(U, U0) and torch.equal(Uf, Uf0)
    finally:
        ctx.close()


def test_noalloc_detector_reports_its_limits(pkg):
    ctx = pkg.Context(0)
    try:
        L = pkg.lib()
        n = 100   # no fixed FFT plan
        U = torch.zeros(n * n, dtype=torch.complex128, device="cuda")
        D = torch.zeros(n * n, dtype=torch.float64, device="cuda")
        rc = L.bimsim_detector_noalloc(ctx.h, 0, 1, U.data_ptr(), U.data_ptr(), n, n, 0.01, 0.01,
                                       0.0, 0.5, 2.0, 0, n, 1, D.data_ptr(), D.data_ptr(), None)
        assert rc == pkg.E_UNSUPPORTED
        # 4096 x 4096 in FP64 with a wide pass band exceeds the work memory
        n = 4096
        U = torch.zeros(n * n, dtype=torch.complex128, device="cuda")
        D = torch.zeros(16, dtype=torch.float64, device="cuda")
        rc = L.bimsim_detector_noalloc(ctx.h, 0, 1, U.data_ptr(), U.data_ptr(), n, n, 1e-3, 1e-3,
                                       0.0, 0.9, 1.0, 0, 4, 1, D.data_ptr(), D.data_ptr(), None)
        assert rc in (pkg.E_INVALID, pkg.E_UNSUPPORTED)
        assert "work memory" in L.bimsim_last_error(ctx.h).decode() or rc == pkg.E_UNSUPPORTED
    finally:
        ctx.close()
