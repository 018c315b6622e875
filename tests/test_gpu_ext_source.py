"""The extended-source detector images (rotate and add, ring sums, radial
line image; reference microscope.cu:37-240) through the C ABI against the
oracle on seeded images: FP64 forms <= 1e-12 normwise, FP32 <= 1e-5."""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

TOL = {0: 1e-12, 2: 1e-5}
RT = {0: torch.float64, 2: torch.float32}


def normrel(a, b):
    m = np.max(np.abs(b))
    return float(np.max(np.abs(np.asarray(a) - b)) / (m if m > 0 else 1.0))


def dev(a, prec):
    return torch.as_tensor(np.ascontiguousarray(a).ravel()).to(RT[prec]).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("prec", [0, 2])
@pytest.mark.parametrize("uR,vR,deg", [(32, 32, 0.3), (33, 20, 2.5), (17, 40, -1.2),
                                       (64, 64, float(np.float32(2 * 3.14159 / 7)))])
def test_rotate_add(pkg, ctx, orc, prec, uR, vR, deg):
    rng = np.random.default_rng(uR * 100 + vR)
    src = rng.random((vR, uR))
    D0 = rng.random((vR, uR))
    want = orc.det_rotate_add(src, D0, deg, 1.0, 1 if prec == 2 else 0)
    D = dev(D0, prec)
    ctx.detector_rotate_add(deg, uR, vR, 1.0, dev(src, prec), D, prec=prec)
    assert normrel(host(D).reshape(vR, uR), want) <= TOL[prec]


@pytest.mark.parametrize("prec", [0, 2])
@pytest.mark.parametrize("uR,vR,ns,nr", [(32, 32, 8, 16), (31, 24, 5, 12), (48, 48, 6, 30),
                                         (20, 20, 1, 3), (64, 64, 1023, 40)])
def test_ring_sums(pkg, ctx, orc, prec, uR, vR, ns, nr):
    rng = np.random.default_rng(ns * 1000 + nr)
    src = rng.random((vR, uR))
    line0 = rng.random(nr + 2)
    want = orc.det_ring_sums(src, line0, ns, nr, 1 if prec == 2 else 0)
    line = dev(line0, prec)
    ctx.detector_ring_sums(uR, vR, ns, nr, dev(src, prec), line, prec=prec)
    got = host(line)
    assert normrel(got, want) <= TOL[prec] * (10 if ns > 100 else 1)
    assert np.array_equal(got[nr:], host(dev(line0[nr:], prec)))   # untouched past n_rings


@pytest.mark.parametrize("prec", [0, 2])
@pytest.mark.parametrize("uR,vR,L", [(32, 32, 12), (25, 18, 30), (16, 16, 1)])
def test_line_image(pkg, ctx, orc, prec, uR, vR, L):
    rng = np.random.default_rng(uR + L)
    line = rng.random(L)
    want = orc.det_from_line(line, uR, vR, 1.0, 1 if prec == 2 else 0)
    D = dev(rng.random((vR, uR)), prec)   # overwritten
    ctx.detector_from_line(uR, vR, 1.0, dev(line, prec), D, prec=prec)
    assert normrel(host(D).reshape(vR, uR), want) <= TOL[prec]


def test_ring_sums_reject_sample_counts(pkg, ctx):
    src = torch.zeros(64, dtype=torch.float64, device="cuda")
    line = torch.zeros(4, dtype=torch.float64, device="cuda")
    for ns in (0, 1024):
        with pytest.raises(pkg.BimsimError):
            ctx.detector_ring_sums(8, 8, ns, 4, src, line)
