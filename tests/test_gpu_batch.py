"""Batches (bimsim_batch_*): one pass over many instances must give each
instance's image as the single-instance pipeline does, eagerly and through
a captured graph, and match the oracle on the spectral-sweep, ragged and
many-sphere workload shapes (reduced sizes).

Tolerances: FP64_PARITY batches equal the per-instance pipeline bit for bit
and the oracle to 1e-10 (normwise) and 1e-9 in absorbance; FP32 batches
equal the per-instance pipeline to 2e-6 (the wave split of the series
differs with the batch size) and the FP64 oracle to 1e-4 in absorbance where
the incident image is above 1e-3 of its peak."""
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


def waves_of(pkg, W, insts):
    return [None if i.get("plane_wave") else W.mc_waves(i, pkg.mc_sample_na) for i in insts]


def run_batch(pkg, ctx, insts, waves, prec, use_graph=False):
    import bimsim_amd.pipeline as P
    arr, spheres, Wv, n_img = P.batch_tables(insts, waves)
    rt = torch.float32 if prec == pkg.FP32_FAST else torch.float64
    D = torch.empty(n_img, dtype=rt, device="cuda")
    Di = torch.empty_like(D)
    Ab = torch.empty_like(D)
    wd = torch.as_tensor(Wv.ravel()).cuda()
    b = ctx.batch(arr, spheres, wd, D, Di, Ab, prec=prec, use_graph=use_graph)
    return b, D, Di, Ab, wd


def split(insts, x):
    out, off = [], 0
    for i in insts:
        n = i["Nd"] * i["Nd"]
        out.append(x[off:off + n])
        off += n
    return out


def single(pkg, ctx, inst, waves, prec):
    import bimsim_amd.pipeline as P
    w = torch.as_tensor(np.asarray(waves if waves is not None else [[0.0, 0.0, 1.0]],
                                   dtype=np.float64)).cuda()
    b = P.simulate(ctx, inst, w, prec)
    return host(b.D), host(b.Di), host(b.Abs)


def ragged(W):
    a = W.config(1, N=48, nk=24)
    b = W.config(2, N=64, nk=33)
    c = W.make_instance([dict(p=(2.0, -1.0, 0.0), a=3.0, n=complex(1.5, 0.02))], 2.8, 48, 30.0,
                        (0.0, 0.5), (0.0, 0.5), 1, 7, plane_wave=True)
    d = W.make_instance([], 3.0, 32, 20.0, (0.0, 0.4), (0.0, 0.4), 8, 9)
    e = W.make_instance([dict(p=(-1.0, 2.0, 0.0), a=2.5, n=complex(1.45, 0.04))], 2.6, 40, 24.0,
                        (0.1, 0.5), (0.0, 0.5), 37, 10, uf_mc=True)
    return [a, b, c, d, e]


@pytest.mark.parametrize("prec", [0, 2])
def test_batch_of_one_equals_the_pipeline(pkg, ctx, W, prec):
    inst = W.config(2, N=64, nk=40)
    waves = waves_of(pkg, W, [inst])
    b, D, Di, Ab, _ = run_batch(pkg, ctx, [inst], waves, prec)
    D1, Di1, Ab1 = single(pkg, ctx, inst, waves[0], prec)
    assert np.array_equal(host(D), D1) and np.array_equal(host(Di), Di1)
    assert np.array_equal(host(Ab), Ab1)
    b.close()


@pytest.mark.parametrize("prec", [0, 2])
def test_mc_focused_batch_of_one_equals_the_pipeline(pkg, ctx, W, prec):
    """An instance whose incident field is the Monte-Carlo sum of its waves:
    the batch and the stepwise entry points run the same per-pixel (FP64)
    or tiled (FP32) code, so the images agree bit for bit."""
    inst = W.make_instance([dict(p=(1.5, -2.0, 0.0), a=3.0, n=complex(1.4, 0.03))], 2.9, 96,
                           30.0, (0.1, 0.5), (0.0, 0.5), 70, 21, uf_mc=True)
    waves = waves_of(pkg, W, [inst])
    b, D, Di, Ab, _ = run_batch(pkg, ctx, [inst], waves, prec)
    D1, Di1, Ab1 = single(pkg, ctx, inst, waves[0], prec)
    assert np.array_equal(host(Di), Di1)
    if prec == 0:
        assert np.array_equal(host(D), D1) and np.array_equal(host(Ab), Ab1)
    else:
        assert normrel(host(D), D1) <= 2e-6
    b.close()


@pytest.mark.parametrize("prec", [0, 2])
def test_zero_filled_detector_batch_equals_the_pipeline(pkg, ctx, W, prec):
    """Instances with the 2x zero-filled detector beside reference-detector
    ones in one batch."""
    a = W.config(1, N=64, nk=20)
    a["zero_fill"] = 2
    b_ = W.config(2, N=32, nk=24)
    insts = [a, b_]
    waves = waves_of(pkg, W, insts)
    b, D, Di, Ab, _ = run_batch(pkg, ctx, insts, waves, prec)
    parts = zip(split(insts, host(D)), split(insts, host(Di)), split(insts, host(Ab)))
    for inst, w, (d, di, ab) in zip(insts, waves, parts):
        D1, Di1, Ab1 = single(pkg, ctx, inst, w, prec)
        if prec == 0:
            assert np.array_equal(d, D1) and np.array_equal(di, Di1)
        else:
            assert normrel(d, D1) <= 2e-6 and normrel(di, Di1) <= 2e-6
    b.close()


@pytest.mark.parametrize("prec", [0, 2])
def test_plane_wave_batch_equals_the_pipeline(pkg, ctx, W, prec):
    """A batch of plane-wave instances only (the FP32 form runs its one
    plane-wave series plan with the coefficients and incident fields in the
    same launch as the records)."""
    insts = [W.make_instance([dict(p=(2.0, -1.0, 0.0), a=3.0, n=complex(1.5, 0.02)),
                              dict(p=(-6.0, 4.0, 0.0), a=2.2, n=complex(1.35, 0.07))],
                             2.8, 48, 30.0, (0.0, 0.5), (0.0, 0.5), 1, 7, plane_wave=True),
             W.make_instance([dict(p=(0.5, 0.5, 0.0), a=4.0, n=complex(1.6, 0.0))],
                             3.1, 32, 24.0, (0.0, 0.4), (0.0, 0.4), 1, 8, plane_wave=True)]
    waves = waves_of(pkg, W, insts)
    b, D, Di, Ab, _ = run_batch(pkg, ctx, insts, waves, prec)
    parts = zip(split(insts, host(D)), split(insts, host(Di)), split(insts, host(Ab)))
    for inst, w, (d, di, ab) in zip(insts, waves, parts):
        D1, Di1, Ab1 = single(pkg, ctx, inst, w, prec)
        if prec == 0:
            assert np.array_equal(d, D1) and np.array_equal(di, Di1)
        else:
            assert normrel(d, D1) <= 2e-6 and normrel(di, Di1) <= 2e-6
    b.close()


@pytest.mark.parametrize("prec", [0, 2])
def test_ragged_batch_equals_per_instance_runs(pkg, ctx, W, prec):
    insts = ragged(W)
    waves = waves_of(pkg, W, insts)
    b, D, Di, Ab, _ = run_batch(pkg, ctx, insts, waves, prec)
    parts = zip(split(insts, host(D)), split(insts, host(Di)), split(insts, host(Ab)))
    for inst, w, (d, di, ab) in zip(insts, waves, parts):
        D1, Di1, Ab1 = single(pkg, ctx, inst, w, prec)
        if prec == 0:
            assert np.array_equal(d, D1) and np.array_equal(di, Di1)
        else:
            assert normrel(d, D1) <= 2e-6 and normrel(di, Di1) <= 2e-6
    b.close()


def test_graph_replays_and_reads_current_inputs(pkg, ctx, W):
    insts = ragged(W)[:3]
    waves = waves_of(pkg, W, insts)
    b, D, Di, Ab, wd = run_batch(pkg, ctx, insts, waves, pkg.FP32_FAST, use_graph=True)
    first = host(D).copy()
    for _ in range(3):
        D.zero_()
        Di.zero_()
        b.launch()
        torch.cuda.synchronize()
        assert np.array_equal(host(D), first)
    # new wave directions in place: the replay sees them
    rng = np.random.default_rng(1)
    w2 = [None if w is None else w[rng.permutation(len(w))][: len(w)] * np.array([-1.0, 1.0, 1.0])
          for w in waves]
    import bimsim_amd.pipeline as P
    _, _, W2, _ = P.batch_tables(insts, w2)
    wd.copy_(torch.as_tensor(W2.ravel()))
    b.launch()
    b2, D2, _, _, _ = run_batch(pkg, ctx, insts, w2, pkg.FP32_FAST)
    assert np.array_equal(host(D), host(D2))
    b.close()
    b2.close()


@pytest.mark.parametrize("prec,shape", [(2, "sweep"), (2, "one"), (0, "one")])
def test_graph_of_the_headline_path(pkg, ctx, W, prec, shape):
    """The headline's path as a captured graph: a spectral sweep (batched
    detector per FFT size, Uf + Us formed by its first row pass) and one
    image (the pre-pass launch with coefficients, records and incident
    field); FP64 batches captured too.  Every replay, with eager work and
    synchronisations between replays, equals the eager pass bit for bit."""
    if shape == "sweep":
        insts = W.config4(n_nu=256, N=64, ns=5, nk=32, stride=37)
    else:
        insts = [W.config(2, N=64, nk=40)]
    waves = waves_of(pkg, W, insts)
    be, De, Die, Abe, _ = run_batch(pkg, ctx, insts, waves, prec)
    be.launch()
    want = host(Abe).copy(), host(De).copy()
    bg, D, Di, Ab, _ = run_batch(pkg, ctx, insts, waves, prec, use_graph=True)
    for _ in range(3):
        D.fill_(7.0)
        Di.fill_(7.0)
        be.launch()          # eager work between the replays
        torch.cuda.synchronize()
        bg.launch()
        torch.cuda.synchronize()
        assert np.array_equal(host(Ab), want[0]) and np.array_equal(host(D), want[1])
    be.close()
    bg.close()


def test_graph_replays_after_the_scratch_grows(pkg, W):
    """A captured batch keeps its image after a larger call on the same
    context grows the context's detector scratch (outgrown buffers stay
    allocated until the context ends)."""
    c = pkg.Context(0)
    try:
        insts = [W.config(2, N=32, nk=16)]
        waves = waves_of(pkg, W, insts)
        b, D, Di, Ab, _ = run_batch(pkg, c, insts, waves, pkg.FP32_FAST, use_graph=True)
        first = host(D).copy()
        big = W.config(2, N=256, nk=16)
        single(pkg, c, big, waves_of(pkg, W, [big])[0], pkg.FP32_FAST)
        for _ in range(2):
            D.zero_()
            Di.zero_()
            b.launch()
            assert np.array_equal(host(D), first)
        b.close()
    finally:
        c.close()


def oracle_image(orc, inst, waves, fp32=0):
    import bimsim_amd.pipeline as P
    q = P.quad_of(inst)
    N = inst["N"]
    if inst.get("uf_mc"):
        Uf = orc.nf_uf_mc(q, N, N, inst["focus"], inst["kmag"], waves, inst["subA"])
    else:
        Uf = orc.nf_uf(q, N, N, inst["focus"], inst["kdir"], inst["kmag"], inst["amp"],
                       inst["cos_alpha"], inst["cos_beta"], inst["m"],
                       mode=1 if inst.get("plane_wave") else 0)
    if inst["spheres"]:
        s0 = inst["spheres"]
        A = np.zeros(inst["ncoef"], dtype=complex)
        B = np.zeros(inst["ncoef"], dtype=complex)
        for s in s0:
            a, b = orc.mie_coeffs(s["Nl"], inst["kmag"], s["a"], s["n"])
            A[s["coeff_off"]:s["coeff_off"] + s["Nl"] + 1] = a
            B[s["coeff_off"]:s["coeff_off"] + s["Nl"] + 1] = b
        if inst.get("plane_wave"):
            w, subA = np.array([list(inst["kdir"])]), inst["amp"]
        else:
            w, subA = waves, inst["subA"]
        Us = orc.nf_us(q, N, N, inst["focus"], inst["kmag"], s0, A, B, w, subA)
        U = orc.nf_sum_uf(q, N, N, s0, Uf, Us)
    else:
        U = Uf.copy()
    out = orc.detector(U, Uf, inst["du"], inst["dv"], inst["obj"][0], inst["obj"][1], inst["lam"],
                       inst["padding"], inst["Nd"], inst["ss"])
    return out[2], out[3], out[4]


@pytest.mark.parametrize("shape", ["sweep", "ragged", "many"])
def test_batch_matches_oracle(pkg, ctx, orc, W, shape):
    if shape == "sweep":
        insts = W.config4(n_nu=256, N=32, ns=4, nk=24, stride=51)
    elif shape == "ragged":
        insts = W.config5(count=6, sizes=(32, 48), nk_range=(20, 60))
    else:
        insts = [W.config3(N=48, ns=8, nk=16, L=60.0)]
    waves = waves_of(pkg, W, insts)
    want = [oracle_image(orc, i, w) for i, w in zip(insts, waves)]
    for prec in (0, 2):
        b, D, Di, Ab, _ = run_batch(pkg, ctx, insts, waves, prec)
        got = zip(split(insts, host(D)), split(insts, host(Di)), split(insts, host(Ab)))
        for (d, di, ab), (Dw, Diw, Abw) in zip(got, want):
            bright = Diw.ravel() > 1e-3 * Diw.max()
            err = np.max(np.abs(ab[bright] - Abw.ravel()[bright]))
            if prec == 0:
                assert normrel(d, Dw.ravel()) <= 1e-10 and normrel(di, Diw.ravel()) <= 1e-10
                assert err <= 1e-9
            else:
                assert err <= 1e-4, (shape, err)
        b.close()


@pytest.mark.parametrize("prec", [0, 2])
def test_extended_source_as_one_batch(pkg, ctx, W, prec):
    """An extended source as one batch: point sources (focus, amplitude) as
    instances sharing one detector image, which then holds the sum of the
    sources' images, as SimulateExtendedSource accumulates them."""
    base = W.config(2, N=48, nk=24)
    insts = []
    for f, amp in [((0.0, 0.0, 0.0), 1.0), ((3.0, -2.0, 0.0), 0.5), ((-4.0, 5.0, 0.0), 0.25)]:
        i = dict(base)
        i["focus"] = f
        i["amp"] = amp
        i["subA"] = (2.0 * W.PI_F) * amp * ((1.0 - i["cos_beta"]) - (1.0 - i["cos_alpha"]))
        insts.append(i)
    waves = waves_of(pkg, W, insts)
    singles = [single(pkg, ctx, i, w, prec) for i, w in zip(insts, waves)]
    import bimsim_amd.pipeline as P
    arr, spheres, Wv, n_img = P.batch_tables(insts, waves)
    for k in range(len(insts)):
        arr[k].image_off = 0
    n = insts[0]["Nd"] ** 2
    rt = torch.float32 if prec == pkg.FP32_FAST else torch.float64
    D = torch.empty(n, dtype=rt, device="cuda")
    Di = torch.empty_like(D)
    Ab = torch.empty_like(D)
    b = ctx.batch(arr, spheres, torch.as_tensor(Wv.ravel()).cuda(), D, Di, Ab, prec=prec)
    want_D = singles[0][0] + singles[1][0] + singles[2][0]
    want_Di = singles[0][1] + singles[1][1] + singles[2][1]
    tol = 1e-13 if prec == 0 else 2e-6
    assert normrel(host(D), want_D) <= tol and normrel(host(Di), want_Di) <= tol
    bright = want_Di > 1e-3 * want_Di.max()
    assert np.max(np.abs(host(Ab)[bright] - (-np.log10(want_D / want_Di))[bright])) <= 100 * tol
    b.close()
