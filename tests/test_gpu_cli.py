"""The bimsim command line end to end on the GPU (C++ host API over the C
ABI), against the oracle run through the same steps (waves -> Uf -> Mie ->
Us (series or tables) -> sum -> detector).  The CLI writes float32 ENVI
files, so the FP64 forms are compared at float32 output rounding (2e-6
normwise); FP32 absorbance to 1e-4 where the incident image is bright."""
import math
import os
import re
import struct
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "bimsim-cuda_amd", "bimsim")
PI_F = 3.1415901184082031


def f32(x):
    return struct.unpack("f", struct.pack("f", x))[0]


def normrel(a, b):
    m = np.max(np.abs(b))
    return float(np.max(np.abs(np.asarray(a) - b)) / (m if m > 0 else 1.0))


def read_envi(path):
    hdr = open(path + ".hdr").read()
    get = lambda k: int(re.search(rf"^{k} = (\d+)", hdr, re.M).group(1))
    s, l, b = get("samples"), get("lines"), get("bands")
    assert get("data type") == 4 and "interleave = bsq" in hdr
    wl = re.search(r"wavelength = \{\s*([^}]*)\}", hdr).group(1)
    return (np.fromfile(path, dtype=np.float32).reshape(b, l, s),
            [float(x) for x in wl.split(",")])


CASE = dict(lam=2.5, size=30.0, res=32, pad=1, sphere=(1.0, -2.0, 0.0, 4.0), n=(1.4, 0.05),
            cond=0.5, obj=0.5, nk=50, m=40)


def run_cli(tmp_path, *extra):
    c = CASE
    args = [EXE, "--spheres", *c["sphere"], "--materials", *c["n"], "--lambda", c["lam"],
            "--xy", c["size"], 0, "--resolution", c["res"], "--padding", c["pad"],
            "--condenser", c["cond"], "--objective", c["obj"], "--samples", c["nk"],
            "--field-order", c["m"], "--intensity", "I", "--incident", "I0",
            "--absorbance", "A", "--transmittance", "T", *extra]
    p = subprocess.run([str(a) for a in args], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    return {k: read_envi(str(tmp_path / k)) for k in ["I", "I0", "A", "T"]}


def oracle_case(orc, lut, focus=(0.0, 0.0, 0.0), amp=1.0):
    """Detector images (D, Di, absorbance) of one source at `focus`."""
    U, Uf = oracle_fields(orc, lut, focus, amp)
    return oracle_detector(orc, U, Uf)


def oracle_detector(orc, U, Uf):
    c = CASE
    L = c["size"] * (2 * c["pad"] + 1)
    out = orc.detector(U, Uf, 1.0 / L, 1.0 / L, 0.0, f32(c["obj"]), f32(c["lam"]), c["pad"],
                       c["res"], 1)
    return out[2], out[3], out[4]


def oracle_fields(orc, lut, focus=(0.0, 0.0, 0.0), amp=1.0):
    """Near fields (U = Uf + Us outside the spheres, Uf) of one source."""
    import oracle
    c = CASE
    lam = f32(c["lam"])
    kmag = 2.0 * PI_F / lam
    N = c["res"] * (2 * c["pad"] + 1)
    L = c["size"] * (2 * c["pad"] + 1)
    q = oracle.Quad.xy(L, 0.0)
    x, y, z, a = (f32(v) for v in c["sphere"])
    n = complex(f32(c["n"][0]), f32(c["n"][1]))
    Nl = orc.sphere_order(a, lam)
    s = dict(p=(x, y, z), a=a, n=n, Nl=Nl, coeff_off=0)
    cond, obj = f32(c["cond"]), f32(c["obj"])
    ca, cb = 1.0, math.cos(math.asin(cond))
    waves = orc.mc_sample_na(c["nk"], (0.0, 0.0, 1.0), 0.0, cond, 0, 0)
    waves = np.asarray(waves, dtype=np.float64).reshape(-1, 3).astype(np.float32).astype(np.float64)
    subA = (2.0 * PI_F) * amp * ((1.0 - cb) - (1.0 - ca))
    kdir = (0.0, 0.0, 1.0)
    Uf = orc.nf_uf(q, N, N, focus, kdir, kmag, amp, ca, cb, c["m"])
    A, B = orc.mie_coeffs(Nl, kmag, a, n)
    if lut:
        aR, dR, dmin, dmax = orc.sphere_lut_dims(q, s, N)
        Usp = orc.lut_eval(Nl, B, orc.lut_hankel(Nl, kmag, a, dmin, dmax, dR), 1000)
        Uip = orc.lut_eval(Nl, A, orc.lut_bessel(Nl, kmag, a, n, aR), 1000)
        Us = orc.nf_us_lut(q, N, N, focus, kmag, s, Usp, Uip, dmin, dmax, waves, subA)
    else:
        Us = orc.nf_us(q, N, N, focus, kmag, [s], A, B, waves, subA)
    return orc.nf_sum_uf(q, N, N, [s], Uf, Us), Uf


@pytest.fixture(scope="module")
def exe():
    if not os.path.exists(EXE):
        pytest.skip("bimsim not built")
    return EXE


@pytest.mark.parametrize("lut", [False, True])
def test_cli_parity_matches_oracle(exe, orc, tmp_path, lut):
    extra = ["--precision", "parity"] + ([] if lut else ["--recursive-us"])
    out = run_cli(tmp_path, *extra)
    D, Di, Ab = oracle_case(orc, lut)
    assert normrel(out["I"][0][0], D) <= 2e-6
    assert normrel(out["I0"][0][0], Di) <= 2e-6
    assert normrel(out["T"][0][0], D / Di) <= 2e-6
    bright = Di > 1e-3 * Di.max()
    assert np.max(np.abs(out["A"][0][0][bright] - Ab[bright])) <= 2e-6
    assert out["A"][1] == [f32(CASE["lam"])]


@pytest.mark.parametrize("lut", [False, True])
def test_cli_fp32_absorbance(exe, orc, tmp_path, lut):
    out = run_cli(tmp_path, *([] if lut else ["--recursive-us"]))
    D, Di, Ab = oracle_case(orc, lut)
    bright = Di > 1e-3 * Di.max()
    assert np.max(np.abs(out["A"][0][0][bright] - Ab[bright])) <= 1e-4


def test_cli_fp64_and_granular_imaging_agree(exe, orc, tmp_path):
    a = run_cli(tmp_path, "--precision", "fp64", "--recursive-us")
    D, Di, _ = oracle_case(orc, False)
    assert normrel(a["I"][0][0], D) <= 2e-6 and normrel(a["I0"][0][0], Di) <= 2e-6


def test_cli_append_and_images(exe, tmp_path):
    run_cli(tmp_path, "--recursive-us", "--near-field", "n.bmp", "--far-field", "f.bmp")
    out = run_cli(tmp_path, "--recursive-us", "--append", "--lambda", 3.0)
    D, wl = out["I"]
    assert D.shape == (2, CASE["res"], CASE["res"])
    assert wl == [f32(2.5), f32(3.0)]
    for name, side in [("n.bmp", CASE["res"] * 3), ("f.bmp", CASE["res"])]:
        data = (tmp_path / name).read_bytes()
        assert data[:2] == b"BM"
        w, h = struct.unpack("<ii", data[18:26])
        assert (w, h) == (side, side)
        assert len(data) == 54 + ((3 * side + 3) & ~3) * side


def test_cli_extended_source_matches_oracle(exe, orc, tmp_path):
    """--extended-source: one point source per non-zero pixel of a BMP
    (focus pos(x/Rx, y/Ry) on the detector plane, amplitude v/255); the
    detector images are the sums of the sources' images (microscope.cu:873-914)."""
    from test_cli import write_bmp
    px = [[0, 200, 0],
          [0, 0, 90],
          [255, 0, 0]]
    write_bmp(tmp_path / "src.bmp", px)
    out = run_cli(tmp_path, "--precision", "parity", "--recursive-us",
                  "--extended-source", "src.bmp")
    half = CASE["size"] / 2
    D = Di = 0.0
    for x in range(3):
        for y in range(3):
            if not px[y][x]:
                continue
            u, v = f32(x / 3), f32(y / 3)
            f = (-half + CASE["size"] * u, -half + CASE["size"] * v, 0.0)
            d, di, _ = oracle_case(orc, False, f, f32(px[y][x] / 255))
            D = D + d
            Di = Di + di
    assert normrel(out["I"][0][0], D) <= 2e-6
    assert normrel(out["I0"][0][0], Di) <= 2e-6
    bright = Di > 1e-3 * Di.max()
    Ab = -np.log10(D / Di)
    assert np.max(np.abs(out["A"][0][0][bright] - Ab[bright])) <= 2e-6


def ring_sources(npts, radius):
    """generatePointSources (microscope.cu:1298-1325): float angles stepping
    by 2 PI_f/npts from one step, foci (cos, sin) * radius."""
    step = np.float32(2) * np.float32(PI_F) / np.float32(npts)
    theta = step
    out = []
    for _ in range(npts):
        u = np.float32(np.cos(theta)) * np.float32(radius)
        v = np.float32(np.sin(theta)) * np.float32(radius)
        out.append(((float(u), float(v), 0.0), float(theta)))
        theta = np.float32(theta + step)
    return out


def check_images(out, D, Di):
    assert normrel(out["I"][0][0], D) <= 2e-6
    assert normrel(out["I0"][0][0], Di) <= 2e-6
    bright = Di > 1e-3 * Di.max()
    Ab = -np.log10(D / Di)
    assert np.max(np.abs(out["A"][0][0][bright] - Ab[bright])) <= 2e-6


PARITY = ("--precision", "parity", "--recursive-us")


def test_cli_interpolate_at_detector(exe, orc, tmp_path):
    """--interpolate 1: per ring, the first source's detector images rotated
    by each other source's angle are added (microscope.cu:1004-1095)."""
    npts, rings = 4, 3
    out = run_cli(tmp_path, *PARITY, "--interpolate", 1, "--points-rings", npts, rings)
    D, Di, _ = oracle_case(orc, False)
    for radius in range(1, rings):
        src = ring_sources(npts, radius)
        d0, di0, _ = oracle_case(orc, False, src[0][0], 1.0)
        D, Di = D + d0, Di + di0
        for _, ang in src[1:]:
            D = orc.det_rotate_add(d0, D, ang)
            Di = orc.det_rotate_add(di0, Di, ang)
    check_images(out, D, Di)


@pytest.mark.parametrize("npts", [4, 6])
def test_cli_interpolate_line(exe, orc, tmp_path, npts):
    """--interpolate 2: ring sums of each source's detector images (sources
    at the focus and at (j, 0, 0)) build radial lines, and the images are the
    lines at each pixel's distance from the centre (microscope.cu:1097-1187);
    six samples per ring sum four of them, as the reference's tree does."""
    rings = 3
    out = run_cli(tmp_path, *PARITY, "--interpolate", 2, "--points-rings", npts, rings)
    line = np.zeros(rings)
    linei = np.zeros(rings)
    for j in range(0, rings):
        d, di, _ = oracle_case(orc, False, (float(j), 0.0, 0.0), 1.0)
        line = orc.det_ring_sums(d, line, npts, rings)
        linei = orc.det_ring_sums(di, linei, npts, rings)
    n = CASE["res"]
    D = orc.det_from_line(line, n, n)
    Di = orc.det_from_line(linei, n, n)
    assert normrel(out["I"][0][0], D) <= 2e-6
    assert normrel(out["I0"][0][0], Di) <= 2e-6


def test_cli_interpolate_near_field(exe, orc, tmp_path):
    """--interpolate 3: per ring, the first source's near fields rotated by
    each other source's angle (U about the integer centre, Uf about the float
    centre) go through the detector (microscope.cu:1189-1256)."""
    npts, rings = 3, 3
    out = run_cli(tmp_path, *PARITY, "--interpolate", 3, "--points-rings", npts, rings)
    D, Di, _ = oracle_case(orc, False)
    for radius in range(1, rings):
        src = ring_sources(npts, radius)
        U0, Uf0 = oracle_fields(orc, False, src[0][0], 1.0)
        d, di, _ = oracle_detector(orc, U0, Uf0)
        D, Di = D + d, Di + di
        for _, ang in src[1:]:
            U = orc.nf_interpolate(U0, ang, centre_float=0)
            Uf = orc.nf_interpolate(Uf0, ang, centre_float=1)
            d, di, _ = oracle_detector(orc, U, Uf)
            D, Di = D + d, Di + di
    check_images(out, D, Di)


def test_cli_interpolate_gaussian_rings(exe, orc, tmp_path):
    """--interpolate 0: with no rings the source at the focus alone; with one
    ring or more the reference's loop never ends, which is an error here."""
    out = run_cli(tmp_path, *PARITY, "--interpolate", 0, "--points-rings", 2, 0)
    D, Di, _ = oracle_case(orc, False)
    check_images(out, D, Di)
    c = CASE
    args = [exe, "--spheres", *c["sphere"], "--materials", *c["n"], "--lambda", c["lam"],
            "--xy", c["size"], 0, "--resolution", c["res"], "--padding", c["pad"],
            "--samples", c["nk"], "--field-order", c["m"], "--intensity", "I2", *PARITY,
            "--interpolate", 0, "--points-rings", 2, 1]
    p = subprocess.run([str(a) for a in args], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert p.returncode != 0 and "does not terminate" in (p.stdout + p.stderr)
    assert not (tmp_path / "I2").exists()


@pytest.mark.parametrize("prec", ["parity", "fp32"])
def test_cli_vector_field(exe, orc, tmp_path, prec):
    """--vector (nearfieldStruct::vectorUf, nfVectorUf.cu:162-191): the scalar
    focused field in x, y and z zero, and no scattered field (the reference
    computes none, nearfield.cpp:160-165), so the images are those of Uf with
    the spheres' interiors cut out (sumUf, nfSumUf.cu:6-48); the intensity
    and near-field views sum |x|^2 + |y|^2 + |z|^2."""
    import oracle
    out = run_cli(tmp_path, "--vector", "--precision", prec, "--recursive",
                  "--near-field", "n.bmp")
    c = CASE
    lam = f32(c["lam"])
    N = c["res"] * (2 * c["pad"] + 1)
    L = c["size"] * (2 * c["pad"] + 1)
    q = oracle.Quad.xy(L, 0.0)
    x, y, z, a = (f32(v) for v in c["sphere"])
    s = dict(p=(x, y, z), a=a, n=complex(f32(c["n"][0]), f32(c["n"][1])),
             Nl=orc.sphere_order(a, lam), coeff_off=0)
    ca, cb = 1.0, math.cos(math.asin(f32(c["cond"])))
    Uf = orc.nf_uf(q, N, N, (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 2.0 * PI_F / lam, 1.0, ca, cb,
                   c["m"])
    U = orc.nf_sum_uf(q, N, N, [s], Uf, np.zeros_like(Uf))
    D, Di, Ab = oracle_detector(orc, U, Uf)
    assert np.min(D) < 0.5 * np.max(D)            # the cut-out is visible
    bright = Di > 1e-3 * Di.max()
    if prec == "parity":
        check_images(out, D, Di)
    else:
        assert np.max(np.abs(out["A"][0][0][bright] - Ab[bright])) <= 1e-4
    assert (tmp_path / "n.bmp").read_bytes()[:2] == b"BM"


def test_cli_fp32_beyond_the_series_orders(exe, tmp_path):
    """A sphere of Nl = 122 (a = 40 um at 2.5 um, past the FP32 series
    kernels' 112 orders) with the default FP32 arithmetic: the scattered
    field comes from the FP64 form, so the images agree with the parity
    form's to the float32 output rounding."""
    def run(sub, *extra):
        d = tmp_path / sub
        d.mkdir()
        args = [EXE, "--spheres", 3, -2, 0, 40, "--materials", 1.2, 0.01, "--lambda", 2.5,
                "--xy", 100, 0, "--resolution", 32, "--padding", 0, "--condenser", 0.5,
                "--objective", 0.5, "--samples", 16, "--field-order", 40, "--recursive",
                "--intensity", "I", "--incident", "I0", "--absorbance", "A", *extra]
        p = subprocess.run([str(a) for a in args], cwd=d, capture_output=True, text=True,
                           timeout=120)
        assert p.returncode == 0, p.stdout + p.stderr
        return {k: read_envi(str(d / k))[0][0] for k in ["I", "I0", "A"]}
    f32 = run("fp32")
    ref = run("parity", "--precision", "parity")
    assert normrel(f32["I"], ref["I"]) <= 1e-5 and normrel(f32["I0"], ref["I0"]) <= 1e-5
