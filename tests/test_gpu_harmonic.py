"""The harmonic (addition-theorem) form of the FP32 scattered-field series
against the FP64 oracle and against the direct FP32 form, on geometries that
exercise every branch: slices through and off the sphere centres, pixels at
a centre, tilted illumination (waves far from the z axis), orders up to the
largest bucket, several spheres.  Tolerance as for the direct form: field
<= 2e-5 normwise."""
import math

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


@pytest.fixture(scope="module")
def W(pkg):
    import bimsim_amd.workloads as W
    return W


def cases(W):
    out = {}
    out["centre"] = W.config(1, N=40, nk=40)                    # pixel at the centre
    out["three"] = W.config(2, N=48, nk=64)
    t = W.make_instance([dict(p=(1.0, -2.0, -1.5), a=4.0, n=complex(1.45, 0.03)),
                         dict(p=(-6.0, 5.0, 2.0), a=2.5, n=complex(1.6, 0.0))], 2.0, 40, 30.0,
                        (0.1, 0.7), (0.0, 0.7), 48, 21)
    t["A"] = (-15.0, -15.0, 1.0)                                  # slice off the centres
    th = 0.6                                                      # tilted illumination
    t["kdir"] = (math.sin(th), 0.0, math.cos(th))
    out["tilted"] = t
    big = W.make_instance([dict(p=(0.0, 0.0, 0.0), a=9.0, n=complex(1.3, 0.01))], 1.2, 32, 30.0,
                          (0.0, 0.5), (0.0, 0.5), 24, 5)
    out["large_order"] = big                                      # Nl ~ 60
    return out


@pytest.mark.parametrize("name", ["centre", "three", "tilted", "large_order"])
def test_harmonic_matches_oracle_and_direct(pkg, ctx, orc, W, name):
    import bimsim_amd.pipeline as P
    inst = cases(W)[name]
    q = P.quad_of(inst)
    N = inst["N"]
    waves = W.mc_waves(inst, pkg.mc_sample_na)
    A = torch.zeros(inst["ncoef"], dtype=torch.complex128, device="cuda")
    B = torch.zeros_like(A)
    ctx.mie_coeffs(inst["spheres"], inst["kmag"], A, B)
    wd = torch.as_tensor(waves).cuda()
    ref = orc.nf_us(q, N, N, inst["focus"], inst["kmag"], inst["spheres"], host(A), host(B), waves,
                    inst["subA"]).ravel()
    out = {}
    for algo in (pkg.SERIES_HARMONIC, pkg.SERIES_DIRECT):
        ctx.set_series_algorithm(algo)
        U = torch.empty(N * N, dtype=torch.complex64, device="cuda")
        ctx.nf_us(q, N, N, inst["focus"], inst["kmag"], inst["spheres"], A, B, wd, inst["subA"], U,
                  prec=pkg.FP32_FAST)
        out[algo] = host(U).astype(np.complex128)
    ctx.set_series_algorithm(pkg.SERIES_AUTO)
    assert normrel(out[pkg.SERIES_HARMONIC], ref) <= 2e-5
    assert normrel(out[pkg.SERIES_DIRECT], ref) <= 2e-5
    assert normrel(out[pkg.SERIES_HARMONIC], out[pkg.SERIES_DIRECT]) <= 2e-5
    if name == "large_order":
        assert max(s["Nl"] for s in inst["spheres"]) >= 49   # the 64-order bucket


def test_harmonic_accumulates(pkg, ctx, orc, W):
    import bimsim_amd.pipeline as P
    inst = W.config(2, N=32, nk=32)
    q = P.quad_of(inst)
    N = inst["N"]
    waves = torch.as_tensor(W.mc_waves(inst, pkg.mc_sample_na)).cuda()
    A = torch.zeros(inst["ncoef"], dtype=torch.complex128, device="cuda")
    B = torch.zeros_like(A)
    ctx.mie_coeffs(inst["spheres"], inst["kmag"], A, B)
    ctx.set_series_algorithm(pkg.SERIES_HARMONIC)
    U = torch.empty(N * N, dtype=torch.complex64, device="cuda")
    args = (q, N, N, inst["focus"], inst["kmag"], inst["spheres"], A, B, waves, inst["subA"], U)
    ctx.nf_us(*args, prec=pkg.FP32_FAST)
    once = host(U).copy()
    ctx.nf_us(*args, prec=pkg.FP32_FAST, accumulate=True)
    ctx.set_series_algorithm(pkg.SERIES_AUTO)
    assert normrel(host(U), 2 * once) <= 1e-6


@pytest.mark.parametrize("a,bucket", [(9.0, 80), (11.0, 96), (13.5, 112)])
def test_harmonic_high_order_buckets(pkg, ctx, orc, W, a, bucket):
    """Spheres of order 74, 88 and 105 (buckets 80, 96, 112: the kernels
    compiled for two and one waves per SIMD) against the oracle."""
    import bimsim_amd.pipeline as P
    inst = W.make_instance([dict(p=(3.0, -2.0, 0.5), a=a, n=complex(1.2, 0.01))], 1.0, 24, 40.0,
                           (0.0, 0.5), (0.0, 0.5), 24, 11)
    Nl = inst["spheres"][0]["Nl"]
    assert bucket - 16 < Nl + 1 <= bucket
    q = P.quad_of(inst)
    N = inst["N"]
    waves = W.mc_waves(inst, pkg.mc_sample_na)
    A = torch.zeros(inst["ncoef"], dtype=torch.complex128, device="cuda")
    B = torch.zeros_like(A)
    ctx.mie_coeffs(inst["spheres"], inst["kmag"], A, B)
    ref = orc.nf_us(q, N, N, inst["focus"], inst["kmag"], inst["spheres"], host(A), host(B), waves,
                    inst["subA"]).ravel()
    ctx.set_series_algorithm(pkg.SERIES_HARMONIC)
    U = torch.empty(N * N, dtype=torch.complex64, device="cuda")
    ctx.nf_us(q, N, N, inst["focus"], inst["kmag"], inst["spheres"], A, B,
              torch.as_tensor(waves).cuda(), inst["subA"], U, prec=pkg.FP32_FAST)
    ctx.set_series_algorithm(pkg.SERIES_AUTO)
    assert normrel(host(U).astype(np.complex128), ref) <= 2e-5


def _harmonic_field(pkg, ctx, orc, W, inst):
    import bimsim_amd.pipeline as P
    q = P.quad_of(inst)
    N = inst["N"]
    waves = W.mc_waves(inst, pkg.mc_sample_na)
    A = torch.zeros(inst["ncoef"], dtype=torch.complex128, device="cuda")
    B = torch.zeros_like(A)
    ctx.mie_coeffs(inst["spheres"], inst["kmag"], A, B)
    ref = orc.nf_us(q, N, N, inst["focus"], inst["kmag"], inst["spheres"], host(A), host(B), waves,
                    inst["subA"]).ravel()
    ctx.set_series_algorithm(pkg.SERIES_HARMONIC)
    U = torch.empty(N * N, dtype=torch.complex64, device="cuda")
    ctx.nf_us(q, N, N, inst["focus"], inst["kmag"], inst["spheres"], A, B,
              torch.as_tensor(waves).cuda(), inst["subA"], U, prec=pkg.FP32_FAST)
    ctx.set_series_algorithm(pkg.SERIES_AUTO)
    return host(U).astype(np.complex128), ref


def test_planar_and_general_forms_agree(pkg, ctx, orc, W):
    """Spheres centred in the slice plane take the planar form (Pn_lm(0) in
    the records, l + m even only); the same image with the slice moved
    1e-9 um off the centres takes the general form.  Both against the
    oracle and against each other."""
    sph = [dict(p=(2.0, -3.0, 0.0), a=4.5, n=complex(1.45, 0.04)),
           dict(p=(-7.0, 6.0, 0.0), a=2.0, n=complex(1.35, 0.0))]
    inst = W.make_instance(sph, 2.0, 40, 30.0, (0.1, 0.6), (0.0, 0.6), 64, 17)
    planar, ref_p = _harmonic_field(pkg, ctx, orc, W, inst)
    off = dict(inst)
    off["A"] = (inst["A"][0], inst["A"][1], 1e-9)
    general, ref_g = _harmonic_field(pkg, ctx, orc, W, off)
    assert normrel(planar, ref_p) <= 2e-5
    assert normrel(general, ref_g) <= 2e-5
    assert normrel(planar, general) <= 2e-5


def test_mixed_planar_and_off_plane_spheres(pkg, ctx, orc, W):
    """One sphere in the slice plane and one above it in one instance: the
    kernel takes the planar form per sphere."""
    sph = [dict(p=(1.0, 2.0, 0.0), a=3.5, n=complex(1.5, 0.02)),
           dict(p=(-5.0, -4.0, 1.5), a=2.5, n=complex(1.3, 0.05))]
    inst = W.make_instance(sph, 2.0, 40, 30.0, (0.0, 0.5), (0.0, 0.5), 48, 23)
    got, ref = _harmonic_field(pkg, ctx, orc, W, inst)
    assert normrel(got, ref) <= 2e-5


@pytest.mark.parametrize("variant", [0, 3, 4, 5, 6, 7, 8])
def test_register_variants_give_the_same_bits(pkg, W, variant, monkeypatch):
    """Every register-budget variant of the harmonic kernel, the chunked
    planar ones (orders summed in chunks of 16 or 8 by VALU FMAs, or in
    chunks of 16, 32 or half the orders on the matrix cores, whose FP32 products
    round as fmaf) included, adds the same terms in the same order: the
    images equal the default variant's bit for bit, on a planar sweep whose
    orders span several buckets and whose 50 x 50 images end in a partial
    wave."""
    import bimsim_amd.pipeline as P
    insts = W.config4(n_nu=256, N=50, ns=20, nk=24, stride=51)
    # 50 x 50 = 2500 pixels: the last wave of each instance has 4 of its 64
    # lanes on pixels (the others repeat the last pixel and store nothing)
    assert (50 * 50) % 64 == 4
    # orders in every bucket from 16 to 64 (all the chunk sizes of the
    # matrix-core variants: 12, 16, 20, 24 and 32)
    assert max(s["Nl"] + 1 for i in insts for s in i["spheres"]) > 48
    # planar: the slice is normal to z through every centre
    assert all(i["X"][2] == 0 and i["Y"][2] == 0 and i["A"][2] == s["p"][2]
               for i in insts for s in i["spheres"])
    waves = [W.mc_waves(i, pkg.mc_sample_na) for i in insts]
    arr, spheres, Wv, n_img = P.batch_tables(insts, waves)
    wd = torch.as_tensor(Wv.ravel()).cuda()

    def image(v):
        monkeypatch.setenv("BIMSIM_HARM_VARIANT", str(v))
        c = pkg.Context(0)
        try:
            D = torch.empty(n_img, dtype=torch.float32, device="cuda")
            Di, Ab = torch.empty_like(D), torch.empty_like(D)
            b = c.batch(arr, spheres, wd, D, Di, Ab, prec=pkg.FP32_FAST, use_graph=False)
            out = host(D).copy()
            b.close()
            return out
        finally:
            c.close()

    assert np.array_equal(image(variant), image(1))
