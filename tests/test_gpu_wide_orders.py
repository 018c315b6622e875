"""Complex64 calls whose spheres need more orders than the FP32 series kernels
hold (Nl + 1 > 112): the scattered field then comes from the FP64 form,
rounded to complex64, through the per-instance entry point and inside FP32
batches (eager and captured), mixed with instances the FP32 series serves.

Tolerances: the per-instance field equals the FP64 oracle to 1e-6 normwise
(only the final rounding differs); batch images equal the per-instance
pipeline to 2e-6 and the FP64 oracle to 1e-4 in absorbance where the
incident image is above 1e-3 of its peak."""
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


def wide_instance(W):
    # a = 40 um at 2.5 um: Nl = 122
    return W.make_instance([dict(p=(3.0, -2.0, 0.0), a=40.0, n=complex(1.2, 0.01))], 2.5, 24, 120.0,
                           (0.0, 0.5), (0.0, 0.5), 8, 3)


def test_per_instance_field_beyond_the_fp32_orders(pkg, ctx, orc, W):
    import bimsim_amd.pipeline as P
    inst = wide_instance(W)
    assert inst["spheres"][0]["Nl"] + 1 > 112
    q = P.quad_of(inst)
    N = inst["N"]
    waves = W.mc_waves(inst, pkg.mc_sample_na)
    A = torch.zeros(inst["ncoef"], dtype=torch.complex128, device="cuda")
    B = torch.zeros_like(A)
    ctx.mie_coeffs(inst["spheres"], inst["kmag"], A, B)
    ref = orc.nf_us(q, N, N, inst["focus"], inst["kmag"], inst["spheres"], host(A), host(B), waves,
                    inst["subA"]).ravel()
    wd = torch.as_tensor(waves).cuda()
    U = torch.empty(N * N, dtype=torch.complex64, device="cuda")
    ctx.nf_us(q, N, N, inst["focus"], inst["kmag"], inst["spheres"], A, B, wd, inst["subA"], U,
              prec=pkg.FP32_FAST)
    once = host(U).astype(np.complex128)
    assert normrel(once, ref) <= 1e-6
    ctx.nf_us(q, N, N, inst["focus"], inst["kmag"], inst["spheres"], A, B, wd, inst["subA"], U,
              prec=pkg.FP32_FAST, accumulate=True)
    assert normrel(host(U).astype(np.complex128), 2 * once) <= 1e-6


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_with_wide_and_series_instances(pkg, ctx, orc, W, use_graph):
    import bimsim_amd.pipeline as P
    small = W.config(1, N=32, nk=16)
    insts = [small, wide_instance(W), W.config(2, N=32, nk=12)]
    waves = [W.mc_waves(i, pkg.mc_sample_na) for i in insts]
    arr, spheres, Wv, n_img = P.batch_tables(insts, waves)
    D = torch.empty(n_img, dtype=torch.float32, device="cuda")
    Di = torch.empty_like(D)
    Ab = torch.empty_like(D)
    wd = torch.as_tensor(Wv.ravel()).cuda()
    b = ctx.batch(arr, spheres, wd, D, Di, Ab, prec=pkg.FP32_FAST, use_graph=use_graph)
    for rep in range(2 if use_graph else 1):
        if rep:
            D.zero_()
            Ab.zero_()
            b.launch()
        off = 0
        for inst, w in zip(insts, waves):
            n = inst["Nd"] * inst["Nd"]
            d, ab = host(D)[off:off + n], host(Ab)[off:off + n]
            off += n
            one = P.simulate(ctx, inst, torch.as_tensor(w).cuda(), pkg.FP32_FAST)
            assert normrel(d, host(one.D)) <= 2e-6
            Dw, Diw, Abw = oracle_absorbance(orc, inst, w)
            bright = Diw > 1e-3 * Diw.max()
            assert np.max(np.abs(ab[bright] - Abw[bright])) <= 1e-4
    b.close()


def oracle_absorbance(orc, inst, waves):
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
    out = orc.detector(U, Uf, inst["du"], inst["dv"], inst["obj"][0], inst["obj"][1], inst["lam"],
                       inst["padding"], inst["Nd"], inst["ss"])
    return out[2].ravel(), out[3].ravel(), out[4].ravel()
