"""Headline benchmark: absorbance images/s (and Mie series terms/s) of the
batched FP32 form against the plain FP64 form on one MI355X.

Headline workload (cfg4 in bimsim-cuda_amd/workloads.py): the spectral sweep
of 256 wavenumbers evenly spaced over 900-4000 cm^-1, 20 non-overlapping
spheres a~U(2,15) um with a Lorentzian index, 400 Monte-Carlo waves, a
512 x 512 slice over 100 um, condenser and objective NA 0.2-0.5.  A step is
one pass of the whole sweep: all 256 x 20 (wavenumber, sphere) instances as
one FP32 batch (Mie coefficients in FP64, wave tables, incident fields,
harmonic scattered-field sums, Uf + Us, detector FFT / band-pass /
-log10), replayed as one captured HIP graph, inputs resident in HBM.  The
plain FP64 form (the reference's per-wave series in double) is timed on
evenly spaced instances of the same sweep, and the FP32 absorbance is
compared with it there.

Reported beside the headline (--no-beside skips them): the 3-sphere,
1000-wave 512 x 512 image (cfg2) as a batch of one; the one-sphere 128 x 128
image (cfg1) with the 2x zero-filled detector; the sphere-table path of the
50-sphere 1024 x 1024 image (cfg3) and that image by the harmonic series;
the 1024-instance ragged batch (cfg5).  The CPU baseline is the oracle (the
reference algorithm in FP64, OpenMP) on a band of rows of one headline image.

usage: python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--no-cpu] [--no-beside]
Prints one JSON line.
"""
import argparse
import importlib.util
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "Mie series terms/s + absorbance images/s, batched FP32 vs plain FP64, 1 GPU"
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector, AMD published (packed FMA)
PEAK_FP64_TFLOPS = 78.6    # MI355X FP64 vector, AMD published
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "series_harm_cfg4_traffic.md")


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


# ------------------------------------------------------------ work counts
def harmonic_flops(Nl, planar=False):
    """FLOPs of the harmonic series kernel per (pixel, sphere), with the
    coefficients B_l (A_l inside) already in the records.  General form,
    P = (Nl+1)(Nl+2)/2 pairs (m, l >= m): per pair the scaled Legendre step
    (3), the rotation by e^{i m phi} (2) and the accumulation (8); per row
    the next diagonal value and phase (8).  Planar form (the slice contains
    the sphere centre and is normal to z, so x = 0 and Pn_lm(0) sits in the
    records): only the pairs with l + m even, 8 each (cos and sin times the
    two complex records); per row the next phase (6).  Both: per order the
    FP64 steps of j_l and y_l (6) and the product of j_l + i y_l with the
    order sum, accumulated (8).  The chunked variants repeat the phase steps
    of the rows below each chunk; that extra work is not counted."""
    n = Nl + 1
    if planar:
        pairs = sum((n - m + 1) // 2 for m in range(n))
        return 8 * pairs + 6 * n + 14 * n
    return 13 * n * (n + 1) // 2 + 8 * n + 14 * n


def is_planar(inst, s):
    """The harmonic kernel's planar case: X.z = Y.z = 0 and A.z = the centre's z."""
    return inst["X"][2] == 0.0 and inst["Y"][2] == 0.0 and inst["A"][2] == s["p"][2]


def series_flops(inst):
    n2 = inst["N"] ** 2
    return n2 * sum(harmonic_flops(s["Nl"], is_planar(inst, s)) for s in inst["spheres"])


def direct_flops(inst):
    """The reference form's count (per (pixel, sphere, wave, l) 8, per
    (pixel, sphere, l) 12, per (pixel, sphere, wave) 13)."""
    n2, nk = inst["N"] ** 2, inst["nk"]
    return n2 * sum((s["Nl"] + 1) * (8 * nk + 12) + 13 * nk for s in inst["spheres"])


def fp64_parity_flops(inst):
    """FLOPs of the plain FP64 form per image, the reference's per-wave loop
    (nfScalarUs.cu:449-491): per (pixel, sphere, wave, l) the j and y steps
    (6), the Legendre step (4), B_l (j + i y) (6) times P_l and the complex
    add (4); per (pixel, sphere, wave) x_w (5) and the weighted phase MAC (8)."""
    n2, nk = inst["N"] ** 2, inst["nk"]
    return n2 * sum(nk * ((s["Nl"] + 1) * 20 + 13) for s in inst["spheres"])


# ------------------------------------------------------------ helpers
def cpu_info():
    """Cores this process may use on the host, and what they are."""
    info = {"nproc": os.cpu_count()}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        info["affinity"] = info["nproc"]
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as fh:
            q, period = fh.read().split()[:2]
            if q != "max":
                quota = float(q) / float(period)
    except (OSError, ValueError):
        pass
    info["cgroup_cpus"] = quota
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core"):
                info[k] = v
    except (OSError, subprocess.SubprocessError):
        pass
    usable = info["affinity"] or 1
    if quota:
        usable = min(usable, max(1, int(quota)))
    tpc = int(info.get("Thread(s) per core", "1") or 1)
    info["threads_used"] = max(1, usable // max(1, tpc) if usable >= 2 * tpc else usable)
    return info


def graph_batch(pkg, P, W, ctx, torch, insts, prec, use_graph=True):
    """A batch of `insts` with its detector buffers (and the wave list)."""
    waves = [None if i.get("plane_wave") else W.mc_waves(i, pkg.mc_sample_na) for i in insts]
    arr, spheres, Wv, n_img = P.batch_tables(insts, waves)
    wd = torch.as_tensor(np.ascontiguousarray(Wv).ravel()).cuda()
    rt = torch.float32 if prec == pkg.FP32_FAST else torch.float64
    D = torch.zeros(n_img, dtype=rt, device="cuda")
    Di = torch.zeros_like(D)
    Ab = torch.zeros_like(D)
    b = ctx.batch(arr, spheres, wd, D, Di, Ab, prec=prec, use_graph=use_graph)
    return b, Di, Ab, waves


def time_batch(b, torch, steps):
    b.launch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.launch()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def abs_error(Ab32, Ab64, Di64):
    """Max |dA| over pixels whose incident image is >= 1e-3 of its peak."""
    bright = Di64 > 1e-3 * Di64.max()
    return float(np.max(np.abs(Ab32.astype(np.float64)[bright] - Ab64[bright])))


def fp64_plain(pkg, P, ctx, torch, inst, w, prec):
    """One image in a FP64 form through the stepwise entry points: wall
    time (after a warm run), the scattered-field stage's time, the buffers."""
    wd = torch.as_tensor(w).cuda()
    P.simulate(ctx, inst, wd, prec)
    torch.cuda.synchronize()
    ctx.set_timing(True)
    ctx.timings()
    t0 = time.perf_counter()
    buf = P.simulate(ctx, inst, wd, prec)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    tm = ctx.timings()
    ctx.set_timing(False)
    return t, tm["us"][0] * 1e-3, buf


# ------------------------------------------------------------ headline
def headline(pkg, P, W, ctx, torch, args):
    insts = W.config4()
    n = len(insts)
    terms = sum(W.terms(i) for i in insts)
    flops = sum(series_flops(i) for i in insts)
    log(f"bench: cfg4 sweep, {n} instances x {len(insts[0]['spheres'])} spheres, N "
        f"{insts[0]['N']}, nk {insts[0]['nk']}, series FLOP per pass {flops:.3e}")
    t0 = time.perf_counter()
    b, Di, Ab, waves = graph_batch(pkg, P, W, ctx, torch, insts, pkg.FP32_FAST)
    log(f"bench: batch created and captured in {time.perf_counter() - t0:.1f} s")
    for _ in range(args.warmup):
        b.launch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.launch()
    torch.cuda.synchronize()
    t32 = (time.perf_counter() - t0) / args.steps
    log(f"bench: FP32 graph pass {t32 * 1e3:.2f} ms ({n / t32:.1f} images/s)")
    Ab32 = Ab.cpu().numpy().reshape(n, -1)
    b.close()

    # the same batch enqueued eagerly with the stage timers on: the series
    # kernels' own time from HIP events on the stream that launches them
    be, _, Abe, _ = graph_batch(pkg, P, W, ctx, torch, insts, pkg.FP32_FAST, use_graph=False)
    be.launch()
    torch.cuda.synchronize()
    ctx.set_timing(True)
    ctx.timings()
    reps = max(2, min(5, args.steps))
    t0 = time.perf_counter()
    for _ in range(reps):
        be.launch()
    torch.cuda.synchronize()
    t_eager = (time.perf_counter() - t0) / reps
    tm = ctx.timings()
    ctx.set_timing(False)
    series_ms, spans = tm["us_series"]
    series_ms_pass = series_ms / reps
    same = bool(np.array_equal(Abe.cpu().numpy().reshape(n, -1), Ab32))
    be.close()
    log(f"bench: FP32 eager pass {t_eager * 1e3:.2f} ms; series kernels {series_ms_pass:.2f} ms "
        f"per pass ({spans // reps} spans); graph and eager images equal: {same}")

    # plain FP64 form (and the hoisted FP64 form) on evenly spaced instances
    picks = [round(k * (n - 1) / max(1, args.fp64_images - 1)) for k in range(args.fp64_images)]
    t64 = t64h = us64 = 0.0
    f64 = 0
    err = 0.0
    for k in picks:
        t, us, buf = fp64_plain(pkg, P, ctx, torch, insts[k], waves[k], pkg.FP64_PARITY)
        t64 += t
        us64 += us
        f64 += fp64_parity_flops(insts[k])
        err = max(err, abs_error(Ab32[k], buf.Abs.cpu().numpy(), buf.Di.cpu().numpy()))
        th, _, _ = fp64_plain(pkg, P, ctx, torch, insts[k], waves[k], pkg.FP64)
        t64h += th
    t64 /= len(picks)
    t64h /= len(picks)
    log(f"bench: FP64 plain {t64 * 1e3:.1f} ms/image, hoisted FP64 {t64h * 1e3:.1f} ms/image, "
        f"absorbance error {err:.2e}")

    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        with open(PROFILE_TRAFFIC) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_pass ="):
                    traffic = int(line.split("=")[1])
    achieved = flops / (series_ms_pass * 1e-3) / 1e12 if series_ms_pass > 0 else None
    ach64 = f64 / us64 / 1e12 if us64 > 0 else None
    return {
        "value": n / t32,
        "ms_per_step": t32 * 1e3,
        "terms_per_s": terms / t32,
        "config": {
            "workload": "cfg4 spectral sweep: 256 wavenumbers 900-4000 cm^-1, 20 non-overlapping "
                        "spheres a~U(2,15) um, n = 1.5 + Lorentz band (1650 cm^-1, 30 cm^-1, "
                        "0.08), 400 Monte-Carlo waves over NA 0.2-0.5, 512x512 slice over 100 um, "
                        "objective NA 0.2-0.5; all 256 x 20 instances as one FP32 batch, one "
                        "captured HIP graph per pass",
            "n_instances": n, "spheres": len(insts[0]["spheres"]), "N": insts[0]["N"],
            "nk": insts[0]["nk"], "Nl_range": [min(s["Nl"] for i in insts for s in i["spheres"]),
                                               max(s["Nl"] for i in insts for s in i["spheres"])],
            "terms_per_pass": terms,
        },
        "eager_pass": {"ms": t_eager * 1e3, "images_per_s": n / t_eager, "same_images": same},
        "fp64_plain": {
            "images_per_s": 1.0 / t64, "ms_per_image": t64 * 1e3, "instances": picks,
            "terms_per_s": sum(W.terms(insts[k]) for k in picks) / (t64 * len(picks)),
            "us_kernel_ms_per_image": us64 * 1e3 / len(picks),
            "roofline": {"bound": "valu_fp64", "achieved": ach64, "peak": PEAK_FP64_TFLOPS,
                         "unit": "TFLOP/s", "frac": ach64 / PEAK_FP64_TFLOPS if ach64 else None,
                         "kernel": "us_parity_kernel (reference per-wave series, FP64)",
                         "flops_per_image": f64 / len(picks)},
        },
        "fp64_hoisted": {"images_per_s": 1.0 / t64h, "ms_per_image": t64h * 1e3},
        "speedup_vs_fp64_plain": (n / t32) * t64,
        "absorbance_max_abs_err_vs_fp64": err,
        "roofline": {
            "bound": "mfma",
            "bound_note": "compute-bound on the SIMDs, not HBM: the order-sum products run on "
                          "the matrix cores (v_mfma_f32_4x4x1_16b_f32) in the variants the "
                          "tuner picks, the FP64 radial recurrences and phase steps on the "
                          "VALU, and the two do not issue beside each other on MI355X",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS if achieved else None,
            "traffic": traffic,
            "traffic_note": "HBM bytes per pass (the six bucket launches) from rocprofv3 "
                            "FETCH_SIZE x 2 + WRITE_SIZE, scaled by a 256 MiB copy; "
                            "algorithmic bytes per pass 5.68e8 (fields once, records, B_l)",
            "kernel": "series_harm_kernel (harmonic FP32 scattered-field sums, planar form in "
                      "chunks of orders, products on VALU FMAs or MFMA; one persistent launch "
                      "per order bucket)",
            "flops_per_pass": flops,
            "series_ms_per_pass": series_ms_pass,
            "peak_note": "157.3 is AMD's published FP32 peak (vector and matrix); this build "
                         "compiles without packed FP32 instructions: plain v_fma_f32 issues "
                         "at about 4.7 cycles per wave, MFMA 4x4x1 at about 10 for four FMAs "
                         "per lane, and the two hardly overlap (profiles/valu_mfma_probe.md)",
        },
    }


# ------------------------------------------------------------ beside
def cfg2_image(pkg, P, W, ctx, torch, steps):
    """The 3-sphere, 1000-wave 512 x 512 image as a batch of one (eager and
    graph), with the plain FP64 form and the direct series form beside."""
    inst = W.config(2)
    waves_h = W.mc_waves(inst, pkg.mc_sample_na)
    arr, spheres, Wv, n_img = P.batch_tables([inst], [waves_h])
    wd = torch.as_tensor(Wv.ravel()).cuda()
    D = torch.empty(n_img, dtype=torch.float32, device="cuda")
    Di, Ab = torch.empty_like(D), torch.empty_like(D)
    b = ctx.batch(arr, spheres, wd, D, Di, Ab, prec=pkg.FP32_FAST, use_graph=False)
    t_eager = time_batch(b, torch, steps)
    Dg, Dig, Abg = torch.empty_like(D), torch.empty_like(D), torch.empty_like(D)
    g = ctx.batch(arr, spheres, wd, Dg, Dig, Abg, prec=pkg.FP32_FAST, use_graph=True)
    t_graph = time_batch(g, torch, steps)
    same = bool(torch.equal(Ab, Abg))
    g.close()
    b.close()
    t64, us64, buf = fp64_plain(pkg, P, ctx, torch, inst, waves_h, pkg.FP64_PARITY)
    err = abs_error(Ab.cpu().numpy(), buf.Abs.cpu().numpy(), buf.Di.cpu().numpy())
    terms = W.terms(inst)
    # the direct (per-wave) FP32 form's series kernel beside the harmonic one
    w = torch.as_tensor(waves_h).cuda()
    b32 = P.Buffers(inst, pkg.FP32_FAST)
    ctx.set_timing(True)
    res = {}
    for name, algo in (("harmonic", pkg.SERIES_HARMONIC), ("direct", pkg.SERIES_DIRECT)):
        ctx.set_series_algorithm(algo)
        P.simulate(ctx, inst, w, pkg.FP32_FAST, b32)
        torch.cuda.synchronize()
        ctx.timings()
        reps = 5 if algo == pkg.SERIES_DIRECT else steps
        for _ in range(reps):
            P.simulate(ctx, inst, w, pkg.FP32_FAST, b32)
        ms, cnt = ctx.timings()["us_series"]
        res[name] = ms / max(cnt, 1)
    ctx.set_series_algorithm(pkg.SERIES_AUTO)
    ctx.set_timing(False)
    dfl = direct_flops(inst)
    log(f"bench: cfg2 image {t_eager * 1e3:.3f} ms eager, {t_graph * 1e3:.3f} ms graph, FP64 "
        f"plain {t64 * 1e3:.1f} ms; series kernels harmonic {res['harmonic']:.4f} ms, direct "
        f"{res['direct']:.3f} ms")
    return {
        "workload": "cfg2: 3 spheres a~U(2,8) um, 1000 Monte-Carlo waves over NA 0.2-0.62, "
                    "lambda 3 um, 512x512 over 60 um; one image as a batch of one instance",
        "ms_per_image_eager": t_eager * 1e3, "ms_per_image_graph": t_graph * 1e3,
        "images_per_s": 1.0 / t_eager, "terms_per_s": terms / t_eager, "graph_same": same,
        "fp64_plain_ms_per_image": t64 * 1e3, "speedup_vs_fp64_plain": t64 / t_eager,
        "absorbance_max_abs_err_vs_fp64": err,
        "harmonic_series_kernel_ms": res["harmonic"],
        "harmonic_series_tflops": series_flops(inst) / (res["harmonic"] * 1e-3) / 1e12,
        "direct_series_kernel_ms": res["direct"],
        "direct_series_flops": dfl,
        "direct_series_tflops_by_reference_count": dfl / (res["direct"] * 1e-3) / 1e12,
    }


def cfg1_image(pkg, P, W, ctx, torch, steps):
    """The one-sphere 128 x 128 image (FP64 plain form, direct series) with the
    reference detector and with the field zero-filled to 256 x 256 for the FFT."""
    out = {}
    for pad in (0, 2):
        inst = W.config(1)
        inst["zero_fill"] = pad
        w = W.mc_waves(inst, pkg.mc_sample_na)
        t64, _, buf = fp64_plain(pkg, P, ctx, torch, inst, w, pkg.FP64_PARITY)
        b, Di, Ab, _ = graph_batch(pkg, P, W, ctx, torch, [inst], pkg.FP32_FAST)
        t32 = time_batch(b, torch, steps)
        err = abs_error(Ab.cpu().numpy(), buf.Abs.cpu().numpy(), buf.Di.cpu().numpy())
        b.close()
        out["zero_filled_2x" if pad else "reference_detector"] = {
            "fp64_plain_ms": t64 * 1e3, "fp32_batch_ms": t32 * 1e3,
            "absorbance_max_abs_err_vs_fp64": err}
    log(f"bench: cfg1 {out}")
    out["workload"] = ("cfg1: one sphere a 5 um, n 1.4+0.05i, lambda 2.5 um, 128x128 over 40 um, "
                       "400 waves, condenser/objective NA 0-0.5; detector FFT at 128 (reference) "
                       "and with the field zero-filled to 256")
    return out


def lut_path(pkg, P, W, ctx, torch, steps=3):
    """The sphere-table path (sphere::calcUp + gpuScalarUpLut) on the whole
    50-sphere 1024 x 1024 image (cfg3: tables of 4096 radii x 1000 angles per
    sphere, 400 waves), FP32 tables and lookups, one sphere at a time."""
    inst = W.config3()
    q = P.quad_of(inst)
    N, thetaR = inst["N"], inst["lut"]["thetaR"]
    A = torch.zeros(inst["ncoef"], dtype=torch.complex128, device="cuda")
    B = torch.zeros_like(A)
    ctx.mie_coeffs(inst["spheres"], inst["kmag"], A, B)
    waves = torch.as_tensor(W.mc_waves(inst, pkg.mc_sample_na)).cuda()
    U = torch.zeros(N * N, dtype=torch.complex64, device="cuda")
    dims = [pkg.sphere_lut_dims(q, s, N, inst["lut"]["dR"]) for s in inst["spheres"]]
    dR_max = max(d[1] for d in dims)
    aR_max = max(d[0] for d in dims)
    Usp = torch.empty(thetaR * dR_max, dtype=torch.complex64, device="cuda")
    Uip = torch.empty(thetaR * aR_max, dtype=torch.complex64, device="cuda")

    def image():
        U.zero_()
        for s, (aR, dR, dmin, dmax) in zip(inst["spheres"], dims):
            ctx.sphere_lut(s, A, B, inst["kmag"], dmin, dmax, dR, aR, thetaR, Usp, Uip,
                           prec=pkg.FP32_FAST)
            ctx.nf_us_lut(q, N, N, inst["focus"], inst["kmag"], s, Usp, dR, Uip, aR, thetaR, dmin,
                          dmax, waves, inst["subA"], U, prec=pkg.FP32_FAST)

    image()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        image()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / steps
    lookups = N * N * inst["nk"] * len(inst["spheres"])
    log(f"bench: cfg3 scattered field by sphere tables {t * 1e3:.1f} ms")
    return {"workload": "cfg3 scattered field by the sphere tables: 50 spheres, 1024x1024, tables "
                        "of 4096 radii x 1000 angles each, 400 waves, FP32",
            "ms_per_image": t * 1e3, "lookups_per_s": lookups / t}


def cfg3_image(pkg, P, W, ctx, torch, steps):
    inst = W.config3()
    b, Di, Ab, _ = graph_batch(pkg, P, W, ctx, torch, [inst], pkg.FP32_FAST)
    t = time_batch(b, torch, steps)
    b6, Di6, Ab6, _ = graph_batch(pkg, P, W, ctx, torch, [inst], pkg.FP64, use_graph=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b6.launch()
    torch.cuda.synchronize()
    t6 = time.perf_counter() - t0
    err = abs_error(Ab.cpu().numpy(), Ab6.cpu().numpy(), Di6.cpu().numpy())
    b.close()
    b6.close()
    terms = W.terms(inst)
    log(f"bench: cfg3 image {t * 1e3:.2f} ms FP32, FP64 batch pass {t6 * 1e3:.1f} ms, "
        f"absorbance error {err:.1e}")
    return {
        "workload": "cfg3: 50 non-overlapping spheres a~U(1,10) um, lambda 4 um, 1024x1024 over "
                    "200 um, 400 waves, harmonic series, one graph-captured instance",
        "ms_per_image": t * 1e3, "terms_per_s": terms / t, "images_per_s": 1.0 / t,
        "fp64_batch_ms_per_image": t6 * 1e3, "absorbance_max_abs_err": err,
    }


def cfg5_batch(pkg, P, W, ctx, torch, steps):
    insts = W.config5()
    b, Di, Ab, _ = graph_batch(pkg, P, W, ctx, torch, insts, pkg.FP32_FAST)
    t = time_batch(b, torch, steps)
    b.close()
    terms = sum(W.terms(i) for i in insts)
    n_mc = sum(1 for i in insts if i.get("uf_mc"))
    log(f"bench: cfg5 batch of {len(insts)} images {t * 1e3:.1f} ms ({len(insts) / t:.0f} images/s)")
    return {
        "workload": "cfg5: 1024 single-sphere instances, a log-U(0.5,25) um, lambda U(2.5,11) um, "
                    "N in {256..1024} over 4a, 200-2000 Monte-Carlo waves each; incident field by "
                    "the Legendre series or (%d instances) the Monte-Carlo sum of the same waves; "
                    "one graph-captured batch" % n_mc,
        "ms_per_batch": t * 1e3, "images_per_s": len(insts) / t, "terms_per_s": terms / t,
    }


# ------------------------------------------------------------ CPU baseline
def cpu_baseline(pkg, W, inst, waves, A, B, sweep_terms, n_images, budget_s=15.0):
    """The oracle (reference algorithm in FP64, OpenMP over rows) on a band of
    rows of one headline image (the sweep's middle wavenumber), scaled to
    terms/s and, by the sweep's mean terms per image, to images/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle

    info = cpu_info()
    orc = Oracle()
    orc.nthreads = info["threads_used"]
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

    rows = max(orc.nthreads, 4)
    dt = run(rows, N // 2)
    rows = int(min(N, max(rows, rows * budget_s / max(dt, 1e-3))))
    rows -= rows % orc.nthreads if rows > orc.nthreads else 0
    r0 = max(0, N // 2 - rows // 2)
    dt = run(rows, r0)
    tps = rows * N * per_pix / dt
    return {
        "value": tps,
        "unit": "terms/s",
        "cores": orc.nthreads,
        "kind": "port",
        "sample": f"scattered-field series (reference FP64 algorithm) of rows {r0}..{r0 + rows - 1} "
                  f"of {N}, all {len(inst['spheres'])} spheres and {inst['nk']} waves of one "
                  f"headline image (lambda {inst['lam']:.3f} um), {dt:.1f} s",
        "images_per_s_est": tps * n_images / sweep_terms,
        "images_per_s_note": "terms/s divided by the sweep's mean terms per image",
        "host": info,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--fp64-images", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-beside", action="store_true")
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

    ctx = pkg.Context(0)
    h = headline(pkg, P, W, ctx, torch, args)
    out = {
        "metric": METRIC,
        "value": h["value"],
        "unit": "images/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: SplitMix64-seeded spheres and Monte-Carlo waves, float32-rounded",
        "config": h["config"],
        "reference_equivalent_terms_per_s": h["terms_per_s"],
        "terms_note": "Mie series terms counted as the reference's per-wave series would evaluate "
                      "them (N^2 nk sum(Nl+1) + N^2 (m+1) per image); the harmonic kernel "
                      "reaches them with fewer operations (roofline.flops_per_pass is its own "
                      "count), so this is a reference-equivalent rate",
        "eager_pass": h["eager_pass"],
        "fp64_plain": h["fp64_plain"],
        "fp64_hoisted": h["fp64_hoisted"],
        "speedup_vs_fp64_plain": h["speedup_vs_fp64_plain"],
        "absorbance_max_abs_err_vs_fp64": h["absorbance_max_abs_err_vs_fp64"],
        "roofline": h["roofline"],
    }
    if not args.no_beside:
        out["beside"] = {
            "cfg2_image": cfg2_image(pkg, P, W, ctx, torch, args.steps),
            "cfg1_image": cfg1_image(pkg, P, W, ctx, torch, args.steps),
            "cfg3_lut_path": lut_path(pkg, P, W, ctx, torch),
            "cfg3_image": cfg3_image(pkg, P, W, ctx, torch, max(3, args.steps // 2)),
            "cfg5_batch": cfg5_batch(pkg, P, W, ctx, torch, max(2, args.steps // 4)),
        }
    if not args.no_cpu:
        log("bench: CPU baseline")
        sweep = W.config4()
        inst = sweep[len(sweep) // 2]
        waves = W.mc_waves(inst, pkg.mc_sample_na)
        A = torch.zeros(inst["ncoef"], dtype=torch.complex128, device="cuda")
        B = torch.zeros_like(A)
        ctx.mie_coeffs(inst["spheres"], inst["kmag"], A, B)
        out["cpu_baseline"] = cpu_baseline(pkg, W, inst, waves, A.cpu().numpy(), B.cpu().numpy(),
                                           h["config"]["terms_per_pass"], len(sweep))
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
