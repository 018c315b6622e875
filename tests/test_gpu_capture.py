"""Per-call entry points under stream capture (torch.cuda.graph on the
context's stream): after one eager call of the same shapes, the whole FP32
image (incident field, coefficients, scattered field, sum, detector) is
captured and each replay gives the eager image bit for bit, with eager work
between replays; a scattered-field call whose plan is not cached fails
with BIMSIM_E_UNSUPPORTED during a capture instead of allocating in it."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.fixture(scope="module")
def W(pkg):
    import bimsim_amd.workloads as W
    return W


@pytest.mark.parametrize("algo", [1, 2])
def test_captured_image_replays(pkg, W, algo):
    import bimsim_amd.pipeline as P
    s = torch.cuda.Stream()
    c = pkg.Context(0, stream=s.cuda_stream)
    try:
        c.set_series_algorithm(algo)
        inst = W.config(2, N=48, nk=32)
        wd = torch.as_tensor(W.mc_waves(inst, pkg.mc_sample_na)).cuda()
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            buf = P.Buffers(inst, pkg.FP32_FAST)
            P.simulate(c, inst, wd, pkg.FP32_FAST, buf)
        want = host(buf.Abs).copy()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            P.simulate(c, inst, wd, pkg.FP32_FAST, buf)
        other = W.config(1, N=48, nk=16)
        for _ in range(3):
            with torch.cuda.stream(s):
                buf.Abs.zero_()
            torch.cuda.synchronize()
            g.replay()
            assert np.array_equal(host(buf.Abs), want)
            # eager work on the same context between replays
            with torch.cuda.stream(s):
                P.simulate(c, other, torch.as_tensor(W.mc_waves(other, pkg.mc_sample_na)).cuda(),
                           pkg.FP32_FAST)
            torch.cuda.synchronize()
        del g
    finally:
        c.close()


def test_uncached_plan_is_an_error_under_capture(pkg, W):
    import bimsim_amd.pipeline as P
    s = torch.cuda.Stream()
    c = pkg.Context(0, stream=s.cuda_stream)
    try:
        inst = W.config(2, N=32, nk=24)
        q = P.quad_of(inst)
        wd = torch.as_tensor(W.mc_waves(inst, pkg.mc_sample_na)).cuda()
        A = torch.zeros(inst["ncoef"], dtype=torch.complex128, device="cuda")
        B = torch.zeros_like(A)
        U = torch.zeros(32 * 32, dtype=torch.complex64, device="cuda")
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            c.mie_coeffs(inst["spheres"], inst["kmag"], A, B)
            c.nf_us(q, 32, 32, inst["focus"], inst["kmag"], inst["spheres"], A, B, wd,
                    inst["subA"], U, prec=pkg.FP32_FAST)
        torch.cuda.synchronize()
        moved = [dict(sp) for sp in inst["spheres"]]
        moved[0]["p"] = (moved[0]["p"][0] + 1.0, moved[0]["p"][1], moved[0]["p"][2])
        g = torch.cuda.CUDAGraph()
        with pytest.raises(pkg.BimsimError) as e:
            with torch.cuda.graph(g, stream=s):
                c.nf_us(q, 32, 32, inst["focus"], inst["kmag"], moved, A, B, wd, inst["subA"], U,
                        prec=pkg.FP32_FAST)
        assert e.value.code == pkg.E_UNSUPPORTED
        assert "outside the capture" in str(e.value)
        torch.cuda.synchronize()
        # the context still works eagerly, and the plan is made there
        with torch.cuda.stream(s):
            c.nf_us(q, 32, 32, inst["focus"], inst["kmag"], moved, A, B, wd, inst["subA"], U,
                    prec=pkg.FP32_FAST)
        torch.cuda.synchronize()
        assert np.isfinite(host(U)).all()
    finally:
        c.close()
