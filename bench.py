"""Headline benchmark: Mie series terms/s and absorbance images/s on one
MI355X, batched FP32 form vs the plain FP64 form.

A step is one absorbance image of the workload (incident field, FP64 Mie
coefficients, scattered field of every sphere under every Monte-Carlo wave,
Uf + Us, detector FFT / band-pass / intensity / -log10) with inputs resident
in device memory, run as a batch of one instance (bimsim_batch_launch), its
kernels enqueued on one stream per step; the same batch replayed as a
captured HIP graph is timed beside it.  The workload is the 3-sphere, 1000-wave, 512 x 512
Monte-Carlo focusing configuration (cfg2 in bimsim-cuda_amd/workloads.py).
The same image through the stepwise entry points (eager launches) and in
the plain FP64 form are reported beside it.

A second, reported-beside measurement ("spectral") times the batched path on
a spectral sweep: every 16th wavenumber of the 256-point 900-4000 cm^-1
grid (16 images), 20 spheres, 400 waves, 512 x 512 (cfg4 in
bimsim-cuda_amd/workloads.py), as one graph-captured batch in FP32 against
the plain FP64 form on two of its images.

A third section ("lut_path") times the sphere-table path on one sphere of
the 50-sphere 1024 x 1024 table workload (cfg3); two more time that whole
image through the harmonic series ("cfg3_image", with its FP64 batch pass
and absorbance error) and the 1024-instance ragged batch ("cfg5_batch"),
each as one graph-captured FP32 batch.  --no-spectral skips all of these
reported-beside sections.

usage: python bench.py [--gpus 1] [--steps 200] [--warmup 20] [--no-cpu] [--no-spectral]
Prints one JSON line.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "Mie series terms/s + absorbance images/s, batched FP32 vs plain FP64, 1 GPU"
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector, AMD published
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "series_harm_cfg2_traffic.md")


def load_pkg():
    d = os.path.join(ROOT, "bimsim-cuda_amd")
    if "bimsim_amd" in sys.modules:
        return sys.modules["bimsim_amd"]
    spec = importlib.util.spec_from_file_location(
        "bimsim_amd", os.path.join(d, "__init__.py"), submodule_search_locations=[d])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["bimsim_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(pkg, W, inst, waves, A, B, budget_s=12.0):
    """The oracle (reference algorithm in FP64, OpenMP over rows) on a band of
    rows of the same workload, scaled to terms/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle

    orc = Oracle()
    threads = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    threads = max(1, min(threads, os.cpu_count() or 1, 16))
    orc.nthreads = threads
    N = inst["N"]
    per_pix = inst["nk"] * sum(s["Nl"] + 1 for s in inst["spheres"])

    def run(rows, r0):
        q = pkg.Quad()
        q.A[:] = [inst["A"][0], inst["A"][1] + inst["L"] * r0 / N, inst["A"][2]]
        q.X[:] = list(inst["X"])
        q.Y[:] = [0.0, inst["L"] * rows / N, 0.0]
        t = time.perf_counter()
        orc.nf_us(q, N, rows, inst["focus"], inst["kmag"], inst["spheres"], A, B, waves,
                  inst["subA"])
        return time.perf_counter() - t

    rows = max(threads, 8)
    dt = run(rows, N // 2)
    rows = int(min(N, max(rows, rows * budget_s / max(dt, 1e-3))))
    rows -= rows % threads if rows > threads else 0
    r0 = max(0, N // 2 - rows // 2)
    dt = run(rows, r0)
    tps = rows * N * per_pix / dt
    return {
        "value": tps,
        "unit": "terms/s",
        "cores": threads,
        "kind": "port",
        "sample": f"scattered-field series (reference FP64 algorithm) of {rows} of {N} rows "
                  f"(rows {r0}..{r0 + rows - 1}), all {len(inst['spheres'])} spheres and "
                  f"{inst['nk']} waves, {dt:.1f} s",
        "images_per_s_est": tps / W.terms(inst),
    }


def harmonic_flops(Nl, planar=False):
    """FLOPs of the harmonic series kernel per (pixel, sphere).  General form,
    P = (Nl+1)(Nl+2)/2 pairs (m, l >= m): per pair the scaled Legendre step
    (3), the rotation by e^{i m phi} (2) and the accumulation (8); per order
    the FP64 radial steps of j_l and y_l (6), B_l h_l (6) and its product with
    the order sum (8); per row the next diagonal value and phase (8).
    Planar form (the slice contains the sphere centre and is normal to z, so
    x = 0 and Pn_lm(0) sits in the records): only the pairs with l + m even,
    8 each (cos and sin times the two complex records); per row the next
    phase (6); per order as above (28)."""
    n = Nl + 1
    if planar:
        pairs = sum((n - m + 1) // 2 for m in range(n))
        return 8 * pairs + 6 * n + 28 * n
    return 13 * n * (n + 1) // 2 + 28 * n


def is_planar(inst, s):
    """The harmonic kernel's planar case: X.z = Y.z = 0 and A.z = the centre's z."""
    return inst["X"][2] == 0.0 and inst["Y"][2] == 0.0 and inst["A"][2] == s["p"][2]


def direct_flops_per(Nl, nk):
    """FLOPs of the direct series kernel per (pixel, sphere): per (wave, l)
    the Legendre step and complex accumulation (8), per wave the phase and
    cos(theta) (13)."""
    return 8 * (Nl + 1) * nk + 13 * nk


def series_kernel_ms(ctx, P, pkg, inst, waves, buf, steps):
    ctx.set_timing(True)
    ctx.timings()
    for _ in range(steps):
        P.simulate(ctx, inst, waves, pkg.FP32_FAST, buf)
    tm = ctx.timings()
    ctx.set_timing(False)
    ms, n = tm["us_series"]
    return ms / max(n, 1), n


def lut_path(pkg, P, W, ctx, torch, steps=3):
    """The sphere-table path (sphere::calcUp + gpuScalarUpLut) on one sphere of
    the 50-sphere 1024 x 1024 table workload (cfg3: lambda 4 um, 200 um field,
    400 waves, tables of 4096 radii x 1000 angles), FP32 tables and lookups."""
    inst = W.make_instance([dict(p=(3.0, -2.0, 0.0), a=6.0, n=complex(1.45, 0.05))], 4.0, 1024,
                           200.0, (0.0, 0.5), (0.0, 0.5), 400, 11)
    s = dict(inst["spheres"][0], coeff_off=0)
    q = P.quad_of(inst)
    N, thetaR = inst["N"], 1000
    aR, dR, dmin, dmax = pkg.sphere_lut_dims(q, s, N, 4096)
    A = torch.zeros(s["Nl"] + 1, dtype=torch.complex128, device="cuda")
    B = torch.zeros_like(A)
    ctx.mie_coeffs([s], inst["kmag"], A, B)
    waves = torch.as_tensor(W.mc_waves(inst, pkg.mc_sample_na)).cuda()
    Usp = torch.empty(thetaR * dR, dtype=torch.complex64, device="cuda")
    Uip = torch.empty(thetaR * aR, dtype=torch.complex64, device="cuda")
    U = torch.zeros(N * N, dtype=torch.complex64, device="cuda")
    t_tab = t_look = 0.0
    for rep in range(steps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.sphere_lut(s, A, B, inst["kmag"], dmin, dmax, dR, aR, thetaR, Usp, Uip,
                       prec=pkg.FP32_FAST)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ctx.nf_us_lut(q, N, N, inst["focus"], inst["kmag"], s, Usp, dR, Uip, aR, thetaR, dmin,
                      dmax, waves, inst["subA"], U, prec=pkg.FP32_FAST)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if rep:   # the first pass warms up
            t_tab += (t1 - t0) / steps
            t_look += (t2 - t1) / steps
    lookups = N * N * inst["nk"]
    log(f"bench: sphere tables {t_tab * 1e3:.2f} ms, table lookups {t_look * 1e3:.2f} ms per sphere")
    return {
        "workload": "one sphere (a 6 um, Nl %d) of cfg3: 1024x1024 over 200 um, lambda 4 um, "
                    "400 waves, tables 4096 radii x 1000 angles, FP32" % s["Nl"],
        "table_ms": t_tab * 1e3,
        "table_terms_per_s": dR * thetaR * (s["Nl"] + 1) / t_tab,
        "lookup_ms": t_look * 1e3,
        "lookups_per_s": lookups / t_look,
        "image_ms_est_50_spheres": 50 * (t_tab + t_look) * 1e3,
    }


def graph_batch(pkg, P, W, ctx, torch, insts, prec):
    """A graph-captured batch of `insts` with its detector buffers."""
    waves = [None if i.get("plane_wave") else W.mc_waves(i, pkg.mc_sample_na) for i in insts]
    arr, spheres, Wv, n_img = P.batch_tables(insts, waves)
    wd = torch.as_tensor(np.ascontiguousarray(Wv).ravel()).cuda()
    rt = torch.float32 if prec == pkg.FP32_FAST else torch.float64
    D = torch.zeros(n_img, dtype=rt, device="cuda")
    Di = torch.zeros_like(D)
    Ab = torch.zeros_like(D)
    b = ctx.batch(arr, spheres, wd, D, Di, Ab, prec=prec, use_graph=prec == pkg.FP32_FAST)
    return b, Di, Ab


def time_batch(b, torch, steps):
    b.launch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.launch()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def large_workloads(pkg, P, W, ctx, torch, steps=20):
    """The 50-sphere 1024 x 1024 image (cfg3) and the 1024-instance ragged
    batch (cfg5), each as one graph-captured FP32 batch; cfg3 also as an FP64
    batch pass for its absorbance error."""
    out = {}
    inst = W.config3()
    b, Di, Ab = graph_batch(pkg, P, W, ctx, torch, [inst], pkg.FP32_FAST)
    t = time_batch(b, torch, steps)
    b6, Di6, Ab6 = graph_batch(pkg, P, W, ctx, torch, [inst], pkg.FP64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b6.launch()
    torch.cuda.synchronize()
    t6 = time.perf_counter() - t0
    di = Di6.cpu().numpy()
    bright = di > 1e-3 * di.max()
    err = float(np.max(np.abs(Ab.cpu().numpy()[bright] - Ab6.cpu().numpy()[bright])))
    b.close()
    b6.close()
    terms = W.terms(inst)
    log(f"bench: cfg3 image {t * 1e3:.2f} ms FP32 ({terms / t:.3e} terms/s), "
        f"FP64 batch pass {t6 * 1e3:.1f} ms, absorbance error {err:.1e}")
    out["cfg3_image"] = {
        "workload": "cfg3: 50 non-overlapping spheres a~U(1,10) um, lambda 4 um, 1024x1024 over "
                    "200 um, 400 waves, harmonic series (no sphere tables), one graph-captured "
                    "instance",
        "ms_per_image": t * 1e3, "terms_per_s": terms / t, "images_per_s": 1.0 / t,
        "fp64_batch_ms_per_image": t6 * 1e3, "absorbance_max_abs_err": err,
    }
    insts = W.config5()
    b, Di, Ab = graph_batch(pkg, P, W, ctx, torch, insts, pkg.FP32_FAST)
    t = time_batch(b, torch, max(1, steps // 2))
    b.close()
    terms = sum(W.terms(i) for i in insts)
    log(f"bench: cfg5 batch of {len(insts)} images {t * 1e3:.1f} ms ({len(insts) / t:.0f} images/s)")
    out["cfg5_batch"] = {
        "workload": "cfg5: 1024 single-sphere instances, a log-U(0.5,25) um, lambda U(2.5,11) um, "
                    "N in {256..1024} over 4a, half plane-wave, half 200-2000 Monte-Carlo waves; "
                    "one graph-captured batch",
        "ms_per_batch": t * 1e3, "images_per_s": len(insts) / t, "terms_per_s": terms / t,
    }
    return out


def spectral(pkg, P, W, ctx, torch, stride=16, steps=20, fp64_images=2):
    """Batched FP32 spectral imaging (one captured batch) against the plain
    FP64 form, images/s and absorbance error on the same instances."""
    insts = W.config4(stride=stride)
    waves = [W.mc_waves(i, pkg.mc_sample_na) for i in insts]
    arr, spheres, Wv, n_img = P.batch_tables(insts, waves)
    wd = torch.as_tensor(Wv.ravel()).cuda()
    D = torch.empty(n_img, dtype=torch.float32, device="cuda")
    Di = torch.empty_like(D)
    Ab = torch.empty_like(D)
    log(f"bench: spectral batch of {len(insts)} images, creating the graph")
    b = ctx.batch(arr, spheres, wd, D, Di, Ab, prec=pkg.FP32_FAST, use_graph=True)
    b.launch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.launch()
    torch.cuda.synchronize()
    t32 = (time.perf_counter() - t0) / steps
    Ab32 = Ab.cpu().numpy().reshape(len(insts), -1)
    b.close()
    # plain FP64 form on evenly spaced instances of the same sweep
    picks = [round(k * (len(insts) - 1) / max(1, fp64_images - 1)) for k in range(fp64_images)]
    t64 = 0.0
    err = 0.0
    for k in picks:
        w = torch.as_tensor(waves[k]).cuda()
        P.simulate(ctx, insts[k], w, pkg.FP64_PARITY)   # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        buf = P.simulate(ctx, insts[k], w, pkg.FP64_PARITY)
        torch.cuda.synchronize()
        t64 += time.perf_counter() - t0
        Di64 = buf.Di.cpu().numpy()
        bright = Di64 > 1e-3 * Di64.max()
        err = max(err, float(np.max(np.abs(Ab32[k][bright] - buf.Abs.cpu().numpy()[bright]))))
    t64 /= len(picks)
    terms = sum(W.terms(i) for i in insts)
    log(f"bench: spectral FP32 batch {t32 * 1e3:.1f} ms ({len(insts) / t32:.1f} images/s), "
        f"FP64 plain {t64 * 1e3:.1f} ms/image")
    return {
        "workload": f"cfg4 sweep: {len(insts)} wavenumbers (every {stride}th of 256, 900-4000 "
                    f"cm^-1), 20 spheres a~U(2,15), 400 waves, 512x512 over 100 um, graph-captured "
                    f"batch",
        "images_per_s_fp32_batch": len(insts) / t32,
        "terms_per_s_fp32_batch": terms / t32,
        "images_per_s_fp64_plain": 1.0 / t64,
        "speedup_vs_fp64_plain": t64 * len(insts) / t32,
        "absorbance_max_abs_err": err,
        "fp64_images_compared": picks,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-spectral", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one device-to-device copy of 256 MiB (the reference "
                         "dispatch for scaling HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch

    pkg = load_pkg()
    import bimsim_amd.pipeline as P
    import bimsim_amd.workloads as W

    if args.calibrate_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("bench: calibration copy of 256 MiB done")
        del src, dst

    inst = W.config(2)
    waves_h = W.mc_waves(inst, pkg.mc_sample_na)
    waves = torch.as_tensor(waves_h).cuda()
    ctx = pkg.Context(0)
    b32 = P.Buffers(inst, pkg.FP32_FAST)
    b64 = P.Buffers(inst, pkg.FP64_PARITY)
    terms = W.terms(inst)
    n2 = inst["N"] ** 2
    series_flops = n2 * sum(harmonic_flops(s["Nl"], is_planar(inst, s)) for s in inst["spheres"])
    planar = all(is_planar(inst, s) for s in inst["spheres"])
    direct_flops = n2 * sum(direct_flops_per(s["Nl"], inst["nk"]) for s in inst["spheres"])
    # the harmonic form is the choice the library makes for Monte-Carlo
    # illumination (every instance with >= 16 waves); set it explicitly
    ctx.set_series_algorithm(pkg.SERIES_HARMONIC)

    log(f"bench: workload cfg2 N={inst['N']} nk={inst['nk']} "
        f"Nl={[s['Nl'] for s in inst['spheres']]} m={inst['m']} terms/image={terms:.3e}")

    # headline: the image as a batch of one instance (the Legendre tables,
    # one launch with the Mie coefficients, records and incident field, the
    # harmonic sums, the detector), its kernels enqueued on one stream per
    # step: the host enqueues an image in ~0.02 ms, well inside its GPU
    # time, and back-to-back steps then run without gaps; a captured graph
    # of the same batch replays with a few microseconds between images and
    # is timed beside it
    arr, spheres, Wv, n_img = P.batch_tables([inst], [waves_h])
    wd = torch.as_tensor(Wv.ravel()).cuda()
    D = torch.empty(n_img, dtype=torch.float32, device="cuda")
    Di = torch.empty_like(D)
    Ab = torch.empty_like(D)
    batch = ctx.batch(arr, spheres, wd, D, Di, Ab, prec=pkg.FP32_FAST, use_graph=False)
    for _ in range(args.warmup):
        batch.launch()
    torch.cuda.synchronize()
    log("bench: warmup done")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.launch()
    torch.cuda.synchronize()
    t32 = (time.perf_counter() - t0) / args.steps
    log(f"bench: FP32 batch {t32 * 1e3:.3f} ms/image")
    Dg, Dig, Abg = torch.empty_like(D), torch.empty_like(D), torch.empty_like(D)
    gbatch = ctx.batch(arr, spheres, wd, Dg, Dig, Abg, prec=pkg.FP32_FAST, use_graph=True)
    t_graph = time_batch(gbatch, torch, args.steps)
    graph_same = bool(torch.equal(Ab, Abg))
    gbatch.close()
    log(f"bench: FP32 graph batch {t_graph * 1e3:.3f} ms/image (same image: {graph_same})")

    # the same image through the stepwise entry points (eager launches)
    for _ in range(args.warmup):
        P.simulate(ctx, inst, waves, pkg.FP32_FAST, b32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        P.simulate(ctx, inst, waves, pkg.FP32_FAST, b32)
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / args.steps
    # the series kernel's own duration: HIP events on the library's stream
    # around its launches, in a separate pass so the timed loops carry none
    avg_series_ms, ser_n = series_kernel_ms(ctx, P, pkg, inst, waves, b32, args.steps)
    log(f"bench: FP32 stepwise {t_step * 1e3:.3f} ms/image, harmonic series kernel "
        f"{avg_series_ms:.3f} ms/launch")
    # the direct form beside it (same image, fewer steps)
    ctx.set_series_algorithm(pkg.SERIES_DIRECT)
    P.simulate(ctx, inst, waves, pkg.FP32_FAST, b32)
    direct_ms, _ = series_kernel_ms(ctx, P, pkg, inst, waves, b32, max(1, args.steps // 3))
    ctx.set_series_algorithm(pkg.SERIES_HARMONIC)
    P.simulate(ctx, inst, waves, pkg.FP32_FAST, b32)
    torch.cuda.synchronize()
    log(f"bench: direct series kernel {direct_ms:.3f} ms/launch")

    # plain FP64 form (reference structure), fewer steps
    P.simulate(ctx, inst, waves, pkg.FP64_PARITY, b64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.fp64_steps):
        P.simulate(ctx, inst, waves, pkg.FP64_PARITY, b64)
    torch.cuda.synchronize()
    t64 = (time.perf_counter() - t0) / args.fp64_steps
    log(f"bench: FP64 plain {t64 * 1e3:.3f} ms/image")

    # accuracy of the FP32 images against the FP64 form on the full workload
    # (the detector works on copies, so U still holds the near field Us + Uf)
    Abs64 = b64.Abs.cpu().numpy()
    Di64 = b64.Di.cpu().numpy()
    bright = Di64 > 1e-3 * Di64.max()
    abs_err = float(np.max(np.abs(Ab.cpu().numpy().astype(np.float64)[bright] - Abs64[bright])))
    abs_err_step = float(np.max(np.abs(b32.Abs.cpu().numpy().astype(np.float64)[bright] -
                                       Abs64[bright])))
    U32 = b32.U.cpu().numpy().astype(np.complex128)
    U64 = b64.U.cpu().numpy()
    field_err = float(np.max(np.abs(U32 - U64)) / np.max(np.abs(U64)))

    # HBM bytes per launch of the series kernel from the committed PMC summary
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        with open(PROFILE_TRAFFIC) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch ="):
                    traffic = int(line.split("=")[1])

    achieved = series_flops / (avg_series_ms * 1e-3) / 1e12 if avg_series_ms > 0 else None
    out = {
        "metric": METRIC,
        "value": terms / t32,
        "unit": "terms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: SplitMix64-seeded spheres and Monte-Carlo waves, float32-rounded",
        "config": {
            "workload": "cfg2: 3 spheres a~U(2,8) um, n~U(1.3,1.6)+iU(0,0.1), 1000 Monte-Carlo "
                        "waves over condenser NA 0.2-0.62, lambda 3 um, 512x512 slice over "
                        "60 um, objective NA 0.62, Legendre incident field; one image per step "
                        "as a batch of one instance enqueued on one stream",
            "N": inst["N"], "nk": inst["nk"], "spheres": len(inst["spheres"]),
            "Nl": [s["Nl"] for s in inst["spheres"]], "m": inst["m"],
            "terms_per_image": terms,
        },
        "images_per_s": 1.0 / t32,
        "graph_batch": {"ms_per_step": t_graph * 1e3, "images_per_s": 1.0 / t_graph,
                        "same_image": graph_same},
        "stepwise_fp32": {"ms_per_step": t_step * 1e3, "terms_per_s": terms / t_step,
                          "absorbance_max_abs_err": abs_err_step},
        "fp64_plain": {"terms_per_s": terms / t64, "images_per_s": 1.0 / t64,
                       "ms_per_step": t64 * 1e3, "steps": args.fp64_steps},
        "speedup_vs_fp64_plain": t64 / t32,
        "fp32_error_vs_fp64": {"field_max_rel": field_err, "absorbance_max_abs": abs_err,
                               "absorbance_pixels": "incident image >= 1e-3 of its peak"},
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS if achieved else None,
            "traffic": traffic,
            "kernel": "series_harm_kernel<32> (harmonic FP32 scattered-field series, "
                      + ("planar form: the slice contains the sphere centres)" if planar
                         else "general form)"),
            "flops_per_launch": series_flops,
            "avg_launch_ms": avg_series_ms,
            "launches": ser_n,
            "peak_note": "157.3 is AMD's FP32 vector peak with packed FMA; this build compiles "
                         "without packed FP32, whose ceiling is half (78.6)",
        },
        "series_direct_form": {
            "kernel": "series_f32_kernel<32>",
            "avg_launch_ms": direct_ms,
            "flops_per_launch": direct_flops,
            "achieved_tflops": direct_flops / (direct_ms * 1e-3) / 1e12 if direct_ms > 0 else None,
        },
    }
    if not args.no_spectral:
        out["spectral"] = spectral(pkg, P, W, ctx, torch)
        out["lut_path"] = lut_path(pkg, P, W, ctx, torch)
        out.update(large_workloads(pkg, P, W, ctx, torch))
    if not args.no_cpu:
        log("bench: CPU baseline")
        A = b64.A.cpu().numpy()
        B = b64.B.cpu().numpy()
        out["cpu_baseline"] = cpu_baseline(pkg, W, inst, waves_h, A, B)
    batch.close()
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
