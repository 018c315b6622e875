"""C-ABI checks that need no GPU: the library loads, exports every function
include/bimsim_hip.h declares, its host helpers (sphere order, Monte-Carlo
directions) match the oracle bit for bit, and entry points reject bad calls
with error codes instead of exiting."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:int|const char\*|void)\s+(bimsim_\w+)\s*\(",
                                    text, re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol(pkg):
    L = pkg.lib()
    names = declared_functions()
    assert len(names) >= 14
    for n in names:
        assert hasattr(L, n), n
    assert set(pkg.EXPORTS) <= set(names)
    assert L.bimsim_abi_version() == 2


def test_sphere_order_matches_oracle(pkg, orc):
    rng = np.random.default_rng(2)
    for _ in range(200):
        a = float(np.float32(rng.uniform(0.3, 30.0)))
        lam = float(np.float32(rng.uniform(1.0, 12.0)))
        assert pkg.sphere_order(a, lam) == orc.sphere_order(a, lam)
    # config-1 sphere: ka = 12.566 -> Nl = 24
    assert pkg.sphere_order(5.0, 2.5) == 24


@pytest.mark.parametrize("rng_kind", [0, 1])
def test_mc_sampler_matches_oracle(pkg, orc, rng_kind):
    for k, nin, nout, seed in [((0.0, 0.0, 1.0), 0.0, 0.5, 0),
                               ((0.0, 0.0, 1.0), 0.2, 0.62, 3),
                               ((0.6, 0.0, 0.8), 0.1, 0.7, 12345)]:
        got = np.array(pkg.mc_sample_na(333, k, nin, nout, rng=rng_kind, seed=seed))
        want = orc.mc_sample_na(333, k, nin, nout, rng_kind, seed)
        assert np.array_equal(got, want)


def test_mc_sampler_distribution(pkg):
    # uniform in cos(phi) over [cos asin NAout, cos asin NAin], uniform azimuth
    w = np.array(pkg.mc_sample_na(20000, (0, 0, 1), 0.2, 0.62, rng=1, seed=99))
    assert np.allclose(np.linalg.norm(w, axis=1), 1.0, atol=1e-12)
    lo, hi = np.cos(np.arcsin(0.62)), np.cos(np.arcsin(0.2))
    assert w[:, 2].min() >= lo - 1e-12 and w[:, 2].max() <= hi + 1e-12
    hist, _ = np.histogram(w[:, 2], bins=10, range=(lo, hi))
    assert hist.min() > 0.85 * hist.mean()
    az = np.arctan2(w[:, 1], w[:, 0])
    hist, _ = np.histogram(az, bins=8, range=(-np.pi, np.pi))
    assert hist.min() > 0.85 * hist.mean()


def test_errors_are_codes_not_exits(pkg):
    L = pkg.lib()
    assert L.bimsim_sync(None) == pkg.E_INVALID
    assert L.bimsim_ctx_destroy(None) == pkg.E_INVALID
    assert L.bimsim_ctx_create(0, None, None) == pkg.E_INVALID
    q = pkg.Quad.xy(10.0)
    f = (ctypes.c_double * 3)(0, 0, 0)
    assert L.bimsim_nf_uf(None, ctypes.byref(q), 8, 8, f, f, 1.0, 1.0, 1.0, 0.5, 10, 0, 0, 1,
                          None) == pkg.E_INVALID
    out = (ctypes.c_double * 3)()
    assert L.bimsim_mc_sample_na(out, 1, (ctypes.c_double * 3)(0, 0, 1), 0.6, 0.2, 0,
                                 0) == pkg.E_INVALID
    assert L.bimsim_mc_sample_na(out, 1, (ctypes.c_double * 3)(0, 0, 1), 0.0, 0.5, 7,
                                 0) == pkg.E_INVALID


def test_context_without_device(pkg):
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    h = ctypes.c_void_p()
    rc = pkg.lib().bimsim_ctx_create(0, None, ctypes.byref(h))
    assert rc == pkg.E_NODEVICE and not h.value


def test_sphere_lut_dims_match_oracle(pkg, orc):
    import oracle
    rng = np.random.default_rng(5)
    for _ in range(50):
        size, z = float(rng.uniform(5, 60)), float(rng.uniform(-3, 3))
        q, qo = pkg.Quad.xy(size, z), oracle.Quad.xy(size, z)
        s = dict(p=tuple(rng.uniform(-10, 10, 3)), a=float(rng.uniform(0.5, 8)),
                 n=complex(1.4, 0.01), Nl=10, coeff_off=0)
        R = int(rng.integers(16, 300))
        assert pkg.sphere_lut_dims(q, s, R) == orc.sphere_lut_dims(qo, s, R)


def test_buffer_checks_reject_mismatched_tensors(pkg):
    """The binding checks every buffer against the call before it reaches the
    library: dtype for the precision, contiguity, size, device."""
    torch = pytest.importorskip("torch")
    c64 = torch.zeros(16, dtype=torch.complex64)
    with pytest.raises(ValueError, match="complex128"):
        pkg.check_field(c64, pkg.FP64_PARITY, 16, "U")
    with pytest.raises(ValueError, match="contiguous"):
        pkg.check_field(torch.zeros(32, dtype=torch.complex64)[::2], pkg.FP32_FAST, 16, "U")
    with pytest.raises(ValueError, match="needs 17"):
        pkg.check_field(c64, pkg.FP32_FAST, 17, "U")
    with pytest.raises(ValueError, match="GPU"):
        pkg.check_field(c64, pkg.FP32_FAST, 16, "U")
    with pytest.raises(ValueError, match="float32"):
        pkg.check_real(torch.zeros(4, dtype=torch.float64), pkg.FP32_FAST, 4, "D")
    pkg.check_field(None, pkg.FP32_FAST, 16, "Uip", optional=True)
    with pytest.raises(ValueError, match="required"):
        pkg.check_field(None, pkg.FP32_FAST, 16, "U")
    assert pkg.coeff_extent([dict(Nl=5, coeff_off=3), dict(Nl=2, coeff_off=0)]) == 9


def test_batch_checks_use_the_instance_table(pkg):
    """Images and waves must cover what the instance table addresses."""
    torch = pytest.importorskip("torch")
    arr = (pkg.Instance * 2)()
    arr[0].Nd, arr[0].image_off, arr[0].uf_mode, arr[0].ns = 8, 0, pkg.UF_SERIES, 0
    arr[1].Nd, arr[1].image_off, arr[1].uf_mode, arr[1].ns = 4, 64, pkg.UF_MC, 0
    arr[1].wave_begin, arr[1].nk = 10, 5
    D = torch.zeros(79, dtype=torch.float32)
    with pytest.raises(ValueError, match="D holds 79 elements, the call needs 80"):
        pkg.Context.check_batch(arr, [], None, D, D, None, pkg.FP32_FAST)
    with pytest.raises(ValueError, match="float32"):
        pkg.Context.check_batch(arr, [], None, D.double(), D, None, pkg.FP32_FAST)
