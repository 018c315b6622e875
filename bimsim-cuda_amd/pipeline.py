"""microscopeStruct::Simulate over the C ABI (reference microscope.cu:818-857,
nearfield.cpp:194-218): incident field, Mie coefficients, scattered field,
Uf + Us sum, detector.  Buffers are torch device tensors; all compute is in
libbimsim_hip.so on the context stream."""
import numpy as np
import torch

from . import FP32_FAST, QUIRKS_REF, UF_MC, UF_PLANE, UF_SERIES, Instance, Quad


def quad_of(inst):
    q = Quad()
    q.A[:] = list(inst["A"])
    q.X[:] = list(inst["X"])
    q.Y[:] = list(inst["Y"])
    return q


class Buffers:
    """Device buffers of one instance at one precision."""

    def __init__(self, inst, prec, device="cuda"):
        n = inst["N"] * inst["N"]
        cd = torch.complex64 if prec == FP32_FAST else torch.complex128
        rd = torch.float32 if prec == FP32_FAST else torch.float64
        self.Uf = torch.empty(n, dtype=cd, device=device)
        self.U = torch.empty(n, dtype=cd, device=device)
        nd = inst["Nd"] * inst["Nd"]
        self.D = torch.zeros(nd, dtype=rd, device=device)
        self.Di = torch.zeros(nd, dtype=rd, device=device)
        self.Abs = torch.empty(nd, dtype=rd, device=device)
        self.A = torch.empty(inst["ncoef"], dtype=torch.complex128, device=device)
        self.B = torch.empty(inst["ncoef"], dtype=torch.complex128, device=device)


def simulate(ctx, inst, waves, prec, buf=None, quirks=QUIRKS_REF, write_filtered=False):
    """One absorbance image.  `waves` is a (nk, 3) float64 device tensor."""
    if buf is None:
        buf = Buffers(inst, prec)
    q = quad_of(inst)
    N = inst["N"]
    mode = uf_mode_of(inst)
    if mode == UF_MC:
        # the incident field as the Monte-Carlo sum of the same waves as Us
        ctx.nf_uf_mc(q, N, N, inst["focus"], inst["kmag"], waves, inst["subA"], buf.Uf,
                     prec=prec)
    else:
        ctx.nf_uf(q, N, N, inst["focus"], inst["kdir"], inst["kmag"], inst["amp"],
                  inst["cos_alpha"], inst["cos_beta"], inst["m"], buf.Uf, mode=mode, prec=prec,
                  quirks=quirks)
    ctx.mie_coeffs(inst["spheres"], inst["kmag"], buf.A, buf.B, quirks=quirks)
    subA = inst["subA"]
    if inst.get("plane_wave"):
        # a single wave along k with amplitude A (nfScalarUs.cu:547-561)
        waves = torch.tensor([list(inst["kdir"])], dtype=torch.float64, device=buf.U.device)
        subA = inst["amp"]
    ctx.nf_us(q, N, N, inst["focus"], inst["kmag"], inst["spheres"], buf.A, buf.B, waves,
              subA, buf.U, prec=prec, quirks=quirks)
    ctx.nf_sum_uf(q, N, N, inst["spheres"], buf.Uf, buf.U, prec=prec)
    buf.D.zero_()
    buf.Di.zero_()
    if inst.get("zero_fill", 0) >= 2:
        ctx.detector_zero_filled(buf.U, buf.Uf, N, N, inst["zero_fill"], inst["du"], inst["dv"],
                                 inst["obj"][0], inst["obj"][1], inst["lam"], buf.D, buf.Di,
                                 buf.Abs, prec=prec, quirks=quirks)
        return buf
    ctx.detector(buf.U, buf.Uf, N, N, inst["du"], inst["dv"], inst["obj"][0], inst["obj"][1],
                 inst["lam"], inst["padding"], inst["Nd"], inst["ss"], buf.D, buf.Di, buf.Abs,
                 prec=prec, quirks=quirks, write_filtered=write_filtered)
    return buf


def uf_mode_of(inst):
    """Incident-field form of an instance: a plane wave, the Monte-Carlo sum
    of its waves (inst["uf_mc"]) or the Legendre series."""
    if inst.get("plane_wave"):
        return UF_PLANE
    return UF_MC if inst.get("uf_mc") else UF_SERIES


def batch_tables(insts, waves):
    """bimsim_instance array, sphere list and stacked waves for a batch.
    `waves[i]` is the (nk, 3) wave array of instance i (ignored for plane
    waves).  Images are packed instance after instance."""
    arr = (Instance * len(insts))()
    spheres, rows = [], []
    wb = 0
    img = 0
    for i, inst in enumerate(insts):
        I = arr[i]
        I.pos.A[:] = list(inst["A"])
        I.pos.X[:] = list(inst["X"])
        I.pos.Y[:] = list(inst["Y"])
        I.N = inst["N"]
        I.m = inst["m"]
        I.focus[:] = list(inst["focus"])
        I.kdir[:] = list(inst["kdir"])
        I.lambda_ = inst["lam"]
        I.amp = inst["amp"]
        I.condenser[:] = list(inst["cond"])
        I.objective[:] = list(inst["obj"])
        I.uf_mode = uf_mode_of(inst)
        I.sphere_begin = len(spheres)
        I.ns = len(inst["spheres"])
        spheres.extend(inst["spheres"])
        I.wave_begin = wb
        if I.uf_mode == UF_PLANE:
            I.nk = 1
        else:
            w = np.asarray(waves[i], dtype=np.float64).reshape(-1, 3)
            I.nk = w.shape[0]
            rows.append(w)
            wb += w.shape[0]
        I.padding = inst["padding"]
        I.Nd = inst["Nd"]
        I.ss = inst["ss"]
        I.image_off = img
        I.zero_fill = inst.get("zero_fill", 0)
        img += inst["Nd"] * inst["Nd"]
    W = np.concatenate(rows) if rows else np.zeros((1, 3))
    return arr, spheres, W, img

