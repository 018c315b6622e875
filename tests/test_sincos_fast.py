"""sincos_reduced (csrc/sincos_fast.h), the FP64 sin and cos of the FP32
harmonic kernels' radial start values, built for the host from the same
source and compared with mpmath's sin and cos of the same doubles over its
range 0 <= x < 2^20: arguments near multiples of pi/4 (the quadrant edges
and the zeros, where a reduction error shows most), small ones, and random
ones.  CPU only: hipcc builds the host side.
"""
import os
import random
import shutil
import subprocess

import pytest

mp = pytest.importorskip("mpmath")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "bimsim-cuda_amd", "csrc")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

DRIVER = r"""
#include "sincos_fast.h"
#include <cstdio>
int main() {
    double x;
    while (std::scanf("%la", &x) == 1) {
        double s, c;
        bimsim::sincos_reduced(x, &s, &c);
        std::printf("%a %a\n", s, c);
    }
    return 0;
}
"""


def arguments():
    rng = random.Random(20)
    xs = [0.0, 1e-300, 1e-8, 0.5, 1.0]
    for n in list(range(1, 64)) + [1000, 4093, 65537, 1 << 19, (1 << 21) - 3]:
        c = n * mp.pi / 4
        if c >= 2 ** 20:
            continue
        base = float(c)
        xs += [base, base * (1 + 2 ** -50), base * (1 - 2 ** -50), base + 1e-9, base - 1e-9]
    xs += [rng.uniform(0, 10) for _ in range(400)]
    xs += [rng.uniform(0, 2 ** 20) for _ in range(400)]
    xs += [2 ** rng.uniform(-20, 20) for _ in range(200)]
    return [x for x in xs if 0 <= x < 2 ** 20]


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not found")
def test_sincos_reduced_against_mpmath(tmp_path):
    src = tmp_path / "drv.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "drv"
    r = subprocess.run([HIPCC, "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17",
                        "-ffp-contract=off", "-I", CSRC, str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    xs = arguments()
    out = subprocess.run([str(exe)], input="".join(x.hex() + "\n" for x in xs),
                         capture_output=True, text=True, check=True).stdout.split("\n")
    mp.mp.prec = 120
    worst = 0.0
    for x, line in zip(xs, out):
        s, c = (float.fromhex(v) for v in line.split())
        for got, ref in ((s, mp.sin(mp.mpf(x))), (c, mp.cos(mp.mpf(x)))):
            # two ulps of the value, or of 2^-60 where the value is that small
            scale = max(abs(float(ref)), 2.0 ** -60)
            err = float(abs(mp.mpf(got) - ref)) / scale
            worst = max(worst, err)
            assert err <= 2 * 2.0 ** -52, (x.hex(), got, float(ref), err)
    assert worst > 0  # the comparison ran
