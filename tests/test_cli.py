"""The bimsim command line on the CPU: option names and defaults of the
reference (arguments.h:27-90, defaults.h), the parameter loading rules
(arguments.h:95-505) and the material-file interpolation, checked through
--print-config, which exits before any device use."""
import math
import os
import struct
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "bimsim-cuda_amd", "bimsim")


def f32(x):
    return struct.unpack("f", struct.pack("f", x))[0]


def run(*args, cwd=None):
    if not os.path.exists(EXE):
        pytest.skip("bimsim not built (make -C bimsim-cuda_amd)")
    return subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, cwd=cwd,
                          timeout=60)


def config(*args, cwd=None):
    p = run("--print-config", *args, cwd=cwd)
    assert p.returncode == 0, p.stdout + p.stderr
    out = {}
    for line in p.stdout.splitlines():
        if "=" in line:
            k, v = line.split("=", 1)
            out[k] = v
    return out


def vec(s):
    return tuple(float(x) for x in s.split(","))


def test_help_lists_the_reference_options():
    p = run("--help")
    assert p.returncode == 1
    for name in ["spheres", "sphere-file", "materials", "material-file", "lambda", "nu", "k",
                 "amplitude", "condenser", "objective", "focus", "plane-wave", "points-rings",
                 "resolution", "plane-lower-left", "plane-upper-right", "plane-normal", "xy",
                 "xz", "yz", "samples", "padding", "supersample", "field-order", "seed",
                 "recursive", "recursive-us", "lut-uf", "intensity", "incident", "absorbance",
                 "absorbance-spec", "transmittance", "far-field", "near-field",
                 "extended-source", "interpolate", "output-type", "colormap", "append",
                 "vector", "verbose"]:
        assert f" --{name}" in p.stdout, name


def test_defaults():
    c = config()
    assert float(c["lambda"]) == 1.0 and int(c["m"]) == 10 and int(c["samples"]) == 400
    assert vec(c["condenser"]) == (0.0, 1.0) and vec(c["objective"]) == (0.0, 1.0)
    assert c["lut_us"] == "1" and c["lut_uf"] == "0" and c["plane_wave"] == "0"
    # resolution 256, padding 1 -> near field 3 x 256 over 3x the plane
    assert vec(c["detector"]) == (256, 256) and int(c["padding"]) == 1
    assert vec(c["nf_res"]) == (768, 768)
    # default plane: lower-left (-5,0,-5), upper-right (5,0,5), normal (0,1,0)
    # -> quad corner A and edges X = (10,0,0)... scaled by 3 about the centre
    A, X, Y = vec(c["nf_pos_A"]), vec(c["nf_pos_X"]), vec(c["nf_pos_Y"])
    assert X == (0.0, 0.0, 30.0) or X == (30.0, 0.0, 0.0)
    assert abs(A[0] + X[0] / 2 + Y[0] / 2) < 1e-12 and abs(A[2] + X[2] / 2 + Y[2] / 2) < 1e-12
    assert vec(c["k"]) == (0.0, 0.0, 1.0)
    # default material 1.4 + 0.05i, no spheres; default files
    assert vec(c["material0"]) == (f32(1.4), f32(0.05))
    assert "sphere0" not in c
    assert c["files"] == "out_i.bmp|out_inc.bmp|out_a.bmp||out_n.bmp|out_f.bmp"
    assert int(c["interpolate"]) == -1


def test_xy_plane_and_supersampling():
    c = config("--xy", 40, 2, "--resolution", 128, "--padding", 0, "--lambda", 2.5)
    assert vec(c["nf_pos_A"]) == (-20.0, -20.0, 2.0)
    assert vec(c["nf_pos_X"]) == (40.0, 0.0, 0.0) and vec(c["nf_pos_Y"]) == (0.0, 40.0, 0.0)
    assert float(c["z_d"]) == 2.0 and int(c["ss"]) == 1
    # dx = 40/16 = 2.5 > lambda/2 = 0.5 -> ss = ceil(2*40/(16*1)) = 5
    c = config("--xy", 40, "--resolution", 16, "--padding", 1, "--lambda", 1)
    assert int(c["ss"]) == 5
    assert vec(c["detector"]) == (16, 16) and vec(c["nf_res"]) == (240, 240)


def test_spheres_and_materials():
    c = config("--spheres", 1, -2, 0.5, 3, 1, 4, 5, 6, 7, 0, "--materials", 1.4, 0.05, 1.5, 0.1)
    assert vec(c["sphere0"]) == (1.0, -2.0, 0.5, 3.0, 1.0)
    assert vec(c["sphere1"]) == (4.0, 5.0, 6.0, 7.0, 0.0)
    assert vec(c["material1"]) == (f32(1.5), f32(0.1))
    # four values: one sphere of material 0
    c = config("--spheres", 0, 0, 0, 5)
    assert vec(c["sphere0"]) == (0.0, 0.0, 0.0, 5.0, 0.0)
    # parameters pass a 6-significant-digit text stream
    c = config("--spheres", 0, 0, 0, 1.23456789)
    assert vec(c["sphere0"])[3] == f32(1.23457)
    # seven values are not a whole number of spheres
    p = run("--print-config", "--spheres", 1, 2, 3, 4, 5, 6, 7)
    assert p.returncode != 0 and "Invalid number of sphere parameters" in p.stdout
    # a sphere needs its material
    p = run("--print-config", "--spheres", 0, 0, 0, 1, 2)
    assert p.returncode != 0 and "material is not loaded" in p.stdout
    # one value: real index, no absorption
    c = config("--materials", 1.33)
    assert vec(c["material0"]) == (f32(1.33), 0.0)


def test_sphere_file(tmp_path):
    (tmp_path / "s.txt").write_text("1 2 3 4\n-1 -2 -3 2 0\n")
    c = config("--sphere-file", "s.txt", cwd=tmp_path)
    assert vec(c["sphere0"]) == (1.0, 2.0, 3.0, 4.0, 0.0)
    assert vec(c["sphere1"]) == (-1.0, -2.0, -3.0, 2.0, 0.0)


def test_material_file_interpolation(tmp_path):
    """rts::function::linear weights the rows the other way round
    (function.h:69-74): n(l) = n_lo * a + n_hi * (1 - a), a = (l - l_lo)/(l_hi - l_lo)."""
    (tmp_path / "m.txt").write_text("#material [microns,n,k]\n1.0 1.4 0.01\n2.0 1.6 0.03\n")
    c = config("--material-file", "m.txt", "--lambda", 1.25, cwd=tmp_path)
    n = vec(c["material0"])
    assert abs(n[0] - (f32(1.4) * 0.25 + f32(1.6) * 0.75)) < 1e-12
    assert abs(n[1] - (f32(0.01) * 0.25 + f32(0.03) * 0.75)) < 1e-12
    # outside the rows: the end values
    c = config("--material-file", "m.txt", "--lambda", 3, cwd=tmp_path)
    assert vec(c["material0"]) == (f32(1.6), f32(0.03))
    # wavenumber format
    (tmp_path / "w.txt").write_text("#[wavenumber,n,k]\n5000 1.4 0.01\n10000 1.6 0.03\n")
    c = config("--material-file", "w.txt", "--lambda", 2.0, cwd=tmp_path)
    assert vec(c["material0"]) == (f32(1.4), f32(0.01))


def test_wavenumber_k_vector_and_flags():
    c = config("--nu", 4000)
    assert float(c["lambda"]) == f32(10000 / f32(4000.0)) and c["wavenumber"] == "1"
    c = config("--k", 0.5, 1.0)
    k = vec(c["k"])
    want = (math.cos(f32(0.5)) * math.sin(1.0), math.sin(f32(0.5)) * math.sin(1.0), math.cos(1.0))
    assert max(abs(a - b) for a, b in zip(k, want)) < 1e-6
    c = config("--recursive-us", "--plane-wave", "--condenser", 0.3)
    assert c["lut_us"] == "0" and c["plane_wave"] == "1" and vec(c["condenser"]) == (0.0, f32(0.3))
    c = config("--recursive")
    assert c["lut_us"] == "0" and c["lut_uf"] == "0"
    c = config("--lut-uf", "--precision", "parity")
    assert c["lut_uf"] == "1" and int(c["precision"]) == 0


def test_missing_source_image_is_an_error(tmp_path):
    p = run("--print-config", "--extended-source", "img.bmp", cwd=tmp_path)
    assert p.returncode != 0 and "cannot read source image" in p.stdout


def write_bmp(path, pixels, bpp=24):
    """pixels[y][x] grey values, row 0 at the top; 24-bit (grey in all three
    channels) or 8-bit with a grey palette."""
    h, w = len(pixels), len(pixels[0])
    row = (bpp * w + 31) // 32 * 4
    pal = b"".join(bytes((i, i, i, 0)) for i in range(256)) if bpp == 8 else b""
    off = 54 + len(pal)
    body = b""
    for y in range(h - 1, -1, -1):     # bottom-up
        line = b"".join(bytes((v,)) if bpp == 8 else bytes((v, v, v)) for v in pixels[y])
        body += line + b"\0" * (row - len(line))
    hdr = b"BM" + struct.pack("<IHHI", off + len(body), 0, 0, off)
    hdr += struct.pack("<IiiHHIIiiII", 40, w, h, 1, bpp, 0, len(body), 2835, 2835,
                       256 if bpp == 8 else 0, 0)
    path.write_bytes(hdr + pal + body)


@pytest.mark.parametrize("bpp", [24, 8])
def test_extended_source_points(tmp_path, bpp):
    """One point source per non-zero pixel (microscope.cu:1258-1296): x outer,
    y inner, focus pos(x/Rx, y/Ry) of the detector plane, amplitude v/255."""
    px = [[0, 0, 255, 0],
          [0, 128, 0, 0],
          [64, 0, 0, 0]]                # 4 wide, 3 high; row 0 at the top
    write_bmp(tmp_path / "src.bmp", px, bpp)
    c = config("--xy", 40, 1.5, "--extended-source", "src.bmp", cwd=tmp_path)
    assert int(c["sources"]) == 3
    want = []
    for x in range(4):
        for y in range(3):
            if px[y][x]:
                u, v = f32(x / 4), f32(y / 3)
                want.append((-20.0 + 40.0 * u, -20.0 + 40.0 * v, 1.5, f32(px[y][x] / 255)))
    for i, w in enumerate(want):
        got = vec(c[f"source{i}"])
        assert all(abs(a - b) <= 1e-5 for a, b in zip(got, w)), (i, got, w)


def test_extended_source_rejects_other_formats(tmp_path):
    (tmp_path / "src.png").write_bytes(b"\x89PNG\r\n\x1a\n")
    p = run("--print-config", "--extended-source", "src.png", cwd=tmp_path)
    assert p.returncode != 0 and "BMP" in p.stdout + p.stderr
