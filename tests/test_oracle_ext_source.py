"""Pins of the oracle's extended-source detector images (reference
microscope.cu:37-240) by what the geometry fixes: a rotation by 0 is the
identity and a quarter turn is an exact permutation about the integer
centre; bilinear sampling is exact on linear images, so ring sums of a
linear image have closed forms, including the samples the reference's
tree sum leaves out when the count is not a power of two; the radial line
image of a linear line is linear in the float distance."""
import math

import numpy as np
import pytest

PI_F = float(np.float32(3.14159))


@pytest.mark.parametrize("fp32", [0, 1])
@pytest.mark.parametrize("uR,vR", [(16, 16), (15, 12), (9, 14)])
def test_rotation_by_zero_adds_the_image(orc, fp32, uR, vR):
    rng = np.random.default_rng(uR * 31 + vR)
    src = rng.random((vR, uR))
    D = rng.random((vR, uR))
    out = orc.det_rotate_add(src, D, 0.0, 1.0, fp32)
    tol = 1e-6 if fp32 else 0.0
    assert np.max(np.abs(out - (D + src))) <= tol * 2


@pytest.mark.parametrize("uR,vR", [(16, 16), (15, 12), (9, 14)])
def test_quarter_turn_is_a_permutation_about_the_integer_centre(orc, uR, vR):
    """deg = pi/2 in double: the sample of (iu, iv) is (uc - (iv - vc),
    vc + (iu - uc)), zero outside; a scale multiplies it."""
    rng = np.random.default_rng(7 + uR)
    src = rng.random((vR, uR))
    out = orc.det_rotate_add(src, np.zeros((vR, uR)), math.pi / 2, 0.5)
    uc, vc = int(uR / 2.0), int(vR / 2.0)
    want = np.zeros((vR, uR))
    for iv in range(vR):
        for iu in range(uR):
            x, y = uc - (iv - vc), vc + (iu - uc)
            if 0 <= x < uR and 0 <= y < vR:
                want[iv, iu] = 0.5 * src[y, x]
    assert np.max(np.abs(out - want)) <= 1e-15


def kept_samples(ns):
    """Indices the reference's tree (stride ns/2 halving to 1) sums."""
    acc = [{t} for t in range(ns)]
    step = ns // 2
    while step >= 1:
        for t in range(step):
            acc[t] = acc[t] | acc[t + step]
        step >>= 1
    return sorted(acc[0])


def test_tree_keeps_the_largest_power_of_two():
    for ns in range(1, 40):
        k = kept_samples(ns)
        assert len(k) == 1 << (ns.bit_length() - 1)
    assert kept_samples(5) == [0, 1, 2, 3]
    assert kept_samples(6) == [0, 1, 3, 4]


@pytest.mark.parametrize("fp32", [0, 1])
@pytest.mark.parametrize("ns", [1, 4, 5, 6, 8, 12])
def test_ring_sums_of_a_linear_image(orc, fp32, ns):
    """f(x, y) = c + a x + b y sampled at (uc + r cos t, vc - r sin t)
    rounded to float, t = (2 PI_f/ns)(k+1), summed over the kept samples k."""
    uR = vR = 24
    a, b, c = 0.25, -0.5, 3.0
    y, x = np.mgrid[0:vR, 0:uR].astype(np.float64)
    src = c + a * x + b * y
    nrings = 6
    line0 = np.linspace(0.0, 1.0, nrings)
    out = orc.det_ring_sums(src, line0, ns, nrings, fp32)
    uc, vc = int(uR / 2.0), int(vR / 2.0)
    kept = kept_samples(ns)
    for r in range(nrings):
        want = line0[r]
        for k in kept:
            t = (2.0 * PI_F / ns) * (k + 1)
            if fp32:
                t = float(np.float32(t))
            # the sample coordinates are floats (texture coordinates)
            x = float(np.float32(uc + r * math.cos(t)))
            y = float(np.float32(vc - r * math.sin(t)))
            want += c + a * x + b * y
        assert abs(out[r] - want) <= (2e-5 if fp32 else 1e-12) * max(1.0, abs(want)), (r, ns)


def test_ring_sums_leave_out_samples(orc):
    """Five samples: the fifth (angle 2 PI_f, at (uc + r, vc)) is not summed."""
    uR = vR = 20
    src = np.zeros((vR, uR))
    uc, vc = int(uR / 2.0), int(vR / 2.0)
    src[vc, uc + 3] = 1.0   # only the point at angle ~0 on the ring of radius 3
    out = orc.det_ring_sums(src, np.zeros(5), 5, 5)
    assert np.max(np.abs(out)) <= 1e-4   # the angle-2pi sample is left out
    out = orc.det_ring_sums(src, np.zeros(5), 4, 5)
    assert abs(out[3] - 1.0) <= 1e-4 and np.max(np.abs(np.delete(out, 3))) <= 1e-4


@pytest.mark.parametrize("fp32", [0, 1])
@pytest.mark.parametrize("uR,vR", [(16, 16), (13, 10)])
def test_line_image_of_a_linear_line(orc, fp32, uR, vR):
    L = 9
    alpha, beta = 2.0, -0.125
    line = alpha + beta * np.arange(L)
    out = orc.det_from_line(line, uR, vR, 1.5, fp32)
    uc, vc = int(uR / 2.0), int(vR / 2.0)
    for iv in range(vR):
        for iu in range(uR):
            r = float(np.float32(math.sqrt((iu - uc) ** 2 + (iv - vc) ** 2)))
            i0 = math.floor(r)
            w = r - i0
            if i0 + 1 <= L - 1:
                want = alpha + beta * r
            elif i0 == L - 1:
                want = (1 - w) * line[L - 1]
            else:
                want = 0.0
            assert abs(out[iv, iu] - 1.5 * want) <= (1e-6 if fp32 else 1e-14), (iu, iv)
