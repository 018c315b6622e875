"""microscopeStruct::Simulate over the C ABI (reference microscope.cu:818-857,
nearfield.cpp:194-218): incident field, Mie coefficients, scattered field,
Uf + Us sum, detector.  Buffers are torch device tensors; all compute is in
libbimsim_hip.so on the context stream."""
import torch

from . import FP32_FAST, QUIRKS_REF, UF_PLANE, UF_SERIES, Quad


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
    mode = UF_PLANE if inst.get("plane_wave") else UF_SERIES
    ctx.nf_uf(q, N, N, inst["focus"], inst["kdir"], inst["kmag"], inst["amp"], inst["cos_alpha"],
              inst["cos_beta"], inst["m"], buf.Uf, mode=mode, prec=prec, quirks=quirks)
    ctx.mie_coeffs(inst["spheres"], inst["kmag"], buf.A, buf.B, quirks=quirks)
    ctx.nf_us(q, N, N, inst["focus"], inst["kmag"], inst["spheres"], buf.A, buf.B, waves,
              inst["subA"], buf.U, prec=prec, quirks=quirks)
    ctx.nf_sum_uf(q, N, N, inst["spheres"], buf.Uf, buf.U, prec=prec)
    buf.D.zero_()
    buf.Di.zero_()
    ctx.detector(buf.U, buf.Uf, N, N, inst["du"], inst["dv"], inst["obj"][0], inst["obj"][1],
                 inst["lam"], inst["padding"], inst["Nd"], inst["ss"], buf.D, buf.Di, buf.Abs,
                 prec=prec, quirks=quirks, write_filtered=write_filtered)
    return buf
