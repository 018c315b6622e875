"""Write the golden vectors of tests/golden/ from the reference's own host code.

Runs only where the reference tree is present: builds oracle/_ref (the
reference's bessjy.cpp, cbessjy.cpp, gamma.cpp, montecarlo.cpp and its
CUDA_CALLABLE recurrence templates compiled where they lie) and records
inputs and outputs as small .npz fixtures:

  ref_bessel.npz   bessjyv_sph / cbessjyva_sph (j, y, j', y' for l = 0..nl)
  ref_mc.npz       mcSampleNA directions after srand(seed)
  ref_recur.npz    upward j/y recurrences with the convergence thresholds
                   and the Legendre recurrence, as calc_Us drives them

usage: python tests/golden/make_golden.py
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def main():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])
    from oracle import RefChecker
    R = RefChecker()
    rng = np.random.default_rng(20261202)

    # real argument: both sides of x = 12, Miller and upward regimes
    real_cases = [(24, 12.566), (30, 0.7), (90, 71.1), (13, 5.0), (54, 40.0), (8, 12.0),
                  (60, 3.3), (29, 25.0)]
    real_cases += [(int(rng.integers(2, 80)), float(rng.uniform(0.05, 90.0))) for _ in range(24)]
    nl_max = max(c[0] for c in real_cases)
    J = np.zeros((len(real_cases), 4, nl_max + 1))
    for i, (nl, x) in enumerate(real_cases):
        for k, arr in enumerate(R.sph_jy(nl, x)):
            J[i, k, : nl + 1] = arr
    # complex argument: ka (real) and kna with absorbing index, |z| <= 12 and > 12
    cplx_cases = [(24, 12.566 + 0j), (24, 17.59 + 0.628j), (13, 6.0 + 0.2j), (29, 11.9 + 0.5j),
                  (54, 30.0 + 2.0j), (40, 25.1 + 0.0j), (8, 2.0 + 0.1j), (81, 60.0 + 5.0j)]
    cplx_cases += [(int(rng.integers(2, 80)),
                    complex(rng.uniform(0.1, 70.0), rng.uniform(0.0, 6.0))) for _ in range(24)]
    nc_max = max(c[0] for c in cplx_cases)
    Z = np.zeros((len(cplx_cases), 4, nc_max + 1), dtype=np.complex128)
    for i, (nl, z) in enumerate(cplx_cases):
        for k, arr in enumerate(R.sph_jy_c(nl, z)):
            Z[i, k, : nl + 1] = arr
    np.savez_compressed(
        os.path.join(HERE, "ref_bessel.npz"),
        real_nl=np.array([c[0] for c in real_cases]), real_x=np.array([c[1] for c in real_cases]),
        real_out=J,
        cplx_nl=np.array([c[0] for c in cplx_cases]),
        cplx_z=np.array([c[1] for c in cplx_cases]), cplx_out=Z,
        gamma_x=np.array([0.5, 1.5, 2.5, -0.5, 3.7, 7.0]),
        gamma_out=np.array([R.gamma(x) for x in [0.5, 1.5, 2.5, -0.5, 3.7, 7.0]]),
    )

    mc = []
    for (k, na_in, na_out, seed, n) in [((0.0, 0.0, 1.0), 0.0, 0.5, 0, 400),
                                        ((0.0, 0.0, 1.0), 0.2, 0.62, 0, 1000),
                                        ((0.3, -0.2, 0.932737905308882), 0.0, 0.5, 7, 300)]:
        kf = np.array(k, dtype=np.float32).astype(np.float64)
        mc.append((kf, na_in, na_out, seed, R.mc_sample_na(n, kf, na_in, na_out, seed)))
    np.savez_compressed(
        os.path.join(HERE, "ref_mc.npz"),
        k=np.array([m[0] for m in mc]), na=np.array([[m[1], m[2]] for m in mc]),
        seed=np.array([m[3] for m in mc]),
        out0=mc[0][4], out1=mc[1][4], out2=mc[2][4])

    rec_cases = [(40, 12.566), (60, 25.0), (30, 3.0), (90, 71.1), (25, 0.9)]
    rec_cases += [(int(rng.integers(4, 80)), float(rng.uniform(0.2, 60.0))) for _ in range(12)]
    nr = max(c[0] for c in rec_cases)
    out = np.zeros((len(rec_cases), 2, 2, nr + 1))
    for i, (nl, x) in enumerate(rec_cases):
        for fp32 in (0, 1):
            j, y = R.sbessel_up(nl, x, fp32)
            out[i, fp32, 0, : nl + 1] = j
            out[i, fp32, 1, : nl + 1] = y
    leg_x = np.array([-1.0, -0.7, -0.31, 0.0, 0.123, 0.5, 0.9, 1.0])
    leg = np.array([[R.legendre_up(80, x, fp32) for x in leg_x] for fp32 in (0, 1)])
    cz = [(30, 10.0 + 0.5j), (40, 17.6 + 0.63j), (20, 3.0 + 0.1j)]
    jc = np.zeros((len(cz), 41), dtype=np.complex128)
    for i, (nl, z) in enumerate(cz):
        jc[i, : nl + 1] = R.sbessel_up_c32(nl, z)
    np.savez_compressed(
        os.path.join(HERE, "ref_recur.npz"),
        nl=np.array([c[0] for c in rec_cases]), x=np.array([c[1] for c in rec_cases]), jy=out,
        leg_x=leg_x, leg=leg, cz_nl=np.array([c[0] for c in cz]),
        cz=np.array([c[1] for c in cz]), jc32=jc)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
