"""ctypes binding of libbimsim_hip.so, the gfx950 BIMSim near-field and
imaging library (C ABI in include/bimsim_hip.h).

Device buffers are torch tensors on the GPU (torch supplies device memory and
the stream); every compute call goes to the HIP library and raises
BimsimError when the library reports a failure.  There is no CPU fallback:
loading fails loudly when libbimsim_hip.so has not been built.
"""
import ctypes
import os
import weakref

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbimsim_hip.so")

c_int, c_double, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p

# bimsim_status / enums (include/bimsim_hip.h)
OK, E_INVALID, E_HIP, E_NOMEM, E_UNSUPPORTED, E_NODEVICE = 0, -1, -2, -3, -4, -5
FP64_PARITY, FP64, FP32_FAST = 0, 1, 2
QUIRKS_OFF, QUIRKS_REF = 0, 1
UF_SERIES, UF_PLANE, UF_LUT, UF_MC = 0, 1, 2, 3
SLICE_INTENSITY, SLICE_MAGNITUDE, SLICE_REAL, SLICE_IMAG = 0, 1, 2, 3
RATIO_ABSORBANCE, RATIO_TRANSMITTANCE = 0, 1
SERIES_AUTO, SERIES_DIRECT, SERIES_HARMONIC = 0, 1, 2
US_SERIES, US_LUT = 0, 1
RNG_GLIBC, RNG_SPLITMIX64 = 0, 1
PI_F = 3.1415901184082031   # float(3.14159f), reference dataTypes.h:17

EXPORTS = (
    "bimsim_abi_version", "bimsim_ctx_create", "bimsim_ctx_destroy", "bimsim_sync",
    "bimsim_last_error", "bimsim_set_timing", "bimsim_get_timings", "bimsim_sphere_order",
    "bimsim_mc_sample_na", "bimsim_mie_coeffs", "bimsim_nf_uf", "bimsim_nf_us",
    "bimsim_nf_sum_uf", "bimsim_detector", "bimsim_uf_bessel_lut", "bimsim_nf_uf_lut",
    "bimsim_sphere_lut_dims", "bimsim_sphere_lut", "bimsim_nf_us_lut", "bimsim_nf_interpolate",
    "bimsim_sphere_radial_lut", "bimsim_field_fft", "bimsim_field_bandpass",
    "bimsim_field_to_scalar", "bimsim_field_scale", "bimsim_field_crop",
    "bimsim_detector_resample", "bimsim_field3_to_scalar", "bimsim_detector_resample3",
    "bimsim_slice_ratio", "bimsim_batch_create",
    "bimsim_batch_launch", "bimsim_batch_destroy", "bimsim_batch_run", "bimsim_set_autotune",
    "bimsim_set_series_algorithm", "bimsim_detector_rotate_add", "bimsim_detector_ring_sums",
    "bimsim_detector_from_line", "bimsim_field_backpropagate", "bimsim_nf_uf_mc",
    "bimsim_detector_zero_filled", "bimsim_detector_noalloc",
)


class BimsimError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bimsim error {code}: {msg}")
        self.code = code


class Quad(ctypes.Structure):
    """bimsim_quad (rts::quad): corner A, edges X and Y."""
    _fields_ = [("A", c_double * 3), ("X", c_double * 3), ("Y", c_double * 3)]

    @classmethod
    def xy(cls, size, z=0.0):
        """The --xy size z slice (arguments.h:340-360 via rts::quad(pMin, pMax, n))."""
        q = cls()
        q.A[:] = [-size / 2.0, -size / 2.0, z]
        q.X[:] = [size, 0.0, 0.0]
        q.Y[:] = [0.0, size, 0.0]
        return q


class Sphere(ctypes.Structure):
    """bimsim_sphere."""
    _fields_ = [("p", c_double * 3), ("a", c_double), ("n", c_double * 2),
                ("Nl", c_int), ("coeff_off", c_int)]


class Instance(ctypes.Structure):
    """bimsim_instance: one absorbance image of a batch."""
    _fields_ = [("pos", Quad), ("N", c_int), ("m", c_int), ("focus", c_double * 3),
                ("kdir", c_double * 3), ("lambda_", c_double), ("amp", c_double),
                ("condenser", c_double * 2), ("objective", c_double * 2),
                ("uf_mode", c_int), ("sphere_begin", c_int), ("ns", c_int),
                ("wave_begin", c_int), ("nk", c_int), ("padding", c_int), ("Nd", c_int),
                ("ss", c_int), ("image_off", ctypes.c_longlong), ("zero_fill", c_int)]


def sphere_array(spheres):
    arr = (Sphere * max(len(spheres), 1))()
    for i, s in enumerate(spheres):
        arr[i].p[:] = list(s["p"])
        arr[i].a = s["a"]
        n = s["n"]
        arr[i].n[:] = [n.real, n.imag] if isinstance(n, complex) else list(n)
        arr[i].Nl = s["Nl"]
        arr[i].coeff_off = s.get("coeff_off", 0)
    return arr


_lib = None


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `make -C bimsim-cuda_amd` "
                              "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.bimsim_last_error.restype = ctypes.c_char_p
        L.bimsim_last_error.argtypes = [c_void_p]
        L.bimsim_ctx_create.argtypes = [c_int, c_void_p, ctypes.POINTER(c_void_p)]
        for name in ("bimsim_ctx_destroy", "bimsim_sync"):
            getattr(L, name).argtypes = [c_void_p]
        L.bimsim_set_timing.argtypes = [c_void_p, c_int]
        L.bimsim_set_autotune.argtypes = [c_void_p, c_int]
        L.bimsim_set_series_algorithm.argtypes = [c_void_p, c_int]
        L.bimsim_get_timings.argtypes = [c_void_p, ctypes.POINTER(ctypes.c_float),
                                         ctypes.POINTER(c_int)]
        L.bimsim_sphere_order.argtypes = [c_double, c_double]
        L.bimsim_mc_sample_na.argtypes = [c_void_p, c_int, c_double * 3, c_double, c_double,
                                          c_int, ctypes.c_ulonglong]
        L.bimsim_mie_coeffs.argtypes = [c_void_p, c_void_p, c_int, c_double, c_int, c_void_p,
                                        c_void_p]
        L.bimsim_nf_uf.argtypes = [c_void_p, ctypes.POINTER(Quad), c_int, c_int, c_double * 3,
                                   c_double * 3, c_double, c_double, c_double, c_double, c_int,
                                   c_int, c_int, c_int, c_void_p]
        L.bimsim_nf_uf_mc.argtypes = [c_void_p, ctypes.POINTER(Quad), c_int, c_int,
                                      c_double * 3, c_double, c_void_p, c_int, c_double, c_int,
                                      c_void_p]
        L.bimsim_nf_us.argtypes = [c_void_p, ctypes.POINTER(Quad), c_int, c_int, c_double * 3,
                                   c_double, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int,
                                   c_double, c_int, c_int, c_void_p, c_int]
        L.bimsim_nf_sum_uf.argtypes = [c_void_p, ctypes.POINTER(Quad), c_int, c_int, c_void_p,
                                       c_int, c_int, c_void_p, c_void_p]
        L.bimsim_detector.argtypes = [c_void_p, c_int, c_int, c_void_p, c_void_p, c_int, c_int,
                                      c_double, c_double, c_double, c_double, c_double, c_int,
                                      c_int, c_int, c_void_p, c_void_p, c_void_p, c_int]
        L.bimsim_detector_zero_filled.argtypes = [c_void_p, c_int, c_int, c_void_p, c_void_p,
                                                  c_int, c_int, c_int, c_double, c_double,
                                                  c_double, c_double, c_double, c_void_p,
                                                  c_void_p, c_void_p]
        L.bimsim_detector_noalloc.argtypes = [c_void_p, c_int, c_int, c_void_p, c_void_p,
                                              c_int, c_int, c_double, c_double, c_double,
                                              c_double, c_double, c_int, c_int, c_int, c_void_p,
                                              c_void_p, c_void_p]
        L.bimsim_uf_bessel_lut.argtypes = [c_void_p, c_int, c_double, c_double, c_double, c_int,
                                           c_int, c_void_p]
        L.bimsim_nf_uf_lut.argtypes = [c_void_p, ctypes.POINTER(Quad), c_int, c_int,
                                       c_double * 3, c_double * 3, c_double, c_double, c_double,
                                       c_int, c_void_p, c_int, c_double, c_double, c_int,
                                       c_void_p]
        ip, dp = ctypes.POINTER(c_int), ctypes.POINTER(c_double)
        L.bimsim_sphere_lut_dims.argtypes = [ctypes.POINTER(Quad), c_void_p, c_int, c_int, ip, ip,
                                             dp, dp]
        L.bimsim_sphere_lut.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_double,
                                        c_double, c_double, c_int, c_int, c_int, c_int, c_void_p,
                                        c_void_p]
        L.bimsim_nf_us_lut.argtypes = [c_void_p, ctypes.POINTER(Quad), c_int, c_int,
                                       c_double * 3, c_double, c_void_p, c_void_p, c_int,
                                       c_void_p, c_int, c_int, c_double, c_double, c_void_p,
                                       c_int, c_double, c_int, c_void_p]
        L.bimsim_nf_interpolate.argtypes = [c_void_p, c_int, c_double, c_int, c_int, c_void_p,
                                            c_void_p, c_void_p, c_void_p]
        L.bimsim_sphere_radial_lut.argtypes = [c_void_p, c_void_p, c_double, c_double, c_double,
                                               c_int, c_int, c_void_p, c_void_p]
        L.bimsim_field_fft.argtypes = [c_void_p, c_int, c_void_p, c_int, c_int, c_int]
        L.bimsim_field_bandpass.argtypes = [c_void_p, c_int, c_int, c_void_p, c_int, c_int,
                                            c_double, c_double, c_double, c_double, c_double]
        L.bimsim_field_to_scalar.argtypes = [c_void_p, c_int, c_int, c_void_p, ctypes.c_longlong,
                                             c_void_p]
        L.bimsim_field_scale.argtypes = [c_void_p, c_int, c_void_p, ctypes.c_longlong, c_double]
        L.bimsim_field_crop.argtypes = [c_void_p, c_int, c_void_p, c_int, c_int, c_int, c_int,
                                        c_int, c_int, c_void_p]
        L.bimsim_detector_resample.argtypes = [c_void_p, c_int, c_void_p, c_int, c_int, c_int,
                                               c_double, c_int, c_void_p]
        L.bimsim_slice_ratio.argtypes = [c_void_p, c_int, c_int, c_void_p, c_void_p,
                                         ctypes.c_longlong, c_void_p]
        L.bimsim_field_backpropagate.argtypes = [c_void_p, c_int, c_int, c_void_p, c_int, c_int,
                                                 c_double, c_double, c_double, c_double]
        L.bimsim_detector_rotate_add.argtypes = [c_void_p, c_int, c_double, c_int, c_int,
                                                 c_double, c_void_p, c_void_p]
        L.bimsim_detector_ring_sums.argtypes = [c_void_p, c_int, c_int, c_int, c_int, c_int,
                                                c_void_p, c_void_p]
        L.bimsim_detector_from_line.argtypes = [c_void_p, c_int, c_int, c_int, c_double,
                                                c_void_p, c_int, c_void_p]
        L.bimsim_batch_create.argtypes = [c_void_p, ctypes.POINTER(Instance), c_int, c_void_p,
                                          c_int, c_void_p, c_int, c_int, c_void_p, c_void_p,
                                          c_void_p, c_int, ctypes.POINTER(c_void_p)]
        L.bimsim_batch_launch.argtypes = [c_void_p]
        L.bimsim_batch_destroy.argtypes = [c_void_p]
        L.bimsim_batch_run.argtypes = [c_void_p, ctypes.POINTER(Instance), c_int, c_void_p, c_int,
                                       c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]
        _lib = L
    return _lib


def sphere_order(a, lam):
    """sphere::calcNl (sphere.h:71-74)."""
    return lib().bimsim_sphere_order(a, lam)


def mc_sample_na(N, k, NAin, NAout, rng=RNG_GLIBC, seed=0):
    """mcSampleNA (montecarlo.cpp:9-71); returns a list of N (x, y, z)."""
    out = (c_double * (3 * N))()
    rc = lib().bimsim_mc_sample_na(out, N, (c_double * 3)(*k), NAin, NAout, rng, seed)
    if rc != OK:
        raise BimsimError(rc, "invalid Monte-Carlo sampling arguments")
    return [tuple(out[3 * i:3 * i + 3]) for i in range(N)]


def sphere_lut_dims(q, sphere, R, dR_override=0):
    """sphere::calcUp table sizes -> (aR, dR, d_min, d_max)."""
    S = sphere_array([sphere])
    aR, dR = c_int(), c_int()
    dmin, dmax = c_double(), c_double()
    rc = lib().bimsim_sphere_lut_dims(ctypes.byref(q), S, R, dR_override, ctypes.byref(aR),
                                      ctypes.byref(dR), ctypes.byref(dmin), ctypes.byref(dmax))
    if rc != OK:
        raise BimsimError(rc, "bad sphere LUT arguments")
    return aR.value, dR.value, dmin.value, dmax.value


def _ptr(t):
    return c_void_p(t.data_ptr()) if t is not None else None


def _dtypes(prec):
    """(complex dtype, real dtype) of fields and images at precision `prec`."""
    import torch
    if prec == FP32_FAST:
        return torch.complex64, torch.float32
    if prec in (FP64_PARITY, FP64):
        return torch.complex128, torch.float64
    raise ValueError(f"unknown precision {prec}")


def check_tensor(t, dtype, n, name, optional=False):
    """Raise ValueError unless `t` is a contiguous device tensor of `dtype`
    with at least `n` elements (None passes when `optional`)."""
    import torch
    if t is None:
        if optional:
            return
        raise ValueError(f"{name} is required")
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name} must be a torch tensor")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype} (got {t.dtype})")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if t.numel() < n:
        raise ValueError(f"{name} holds {t.numel()} elements, the call needs {n}")
    if not t.is_cuda:
        raise ValueError(f"{name} must be on the GPU (got {t.device})")


def check_field(t, prec, n, name, optional=False):
    check_tensor(t, _dtypes(prec)[0], n, name, optional)


def check_real(t, prec, n, name, optional=False):
    check_tensor(t, _dtypes(prec)[1], n, name, optional)


def coeff_extent(spheres):
    """Elements of the packed A_l / B_l arrays the spheres address."""
    return max((s.get("coeff_off", 0) + s["Nl"] + 1 for s in spheres), default=0)


def check_waves(waves, name="waves"):
    """Monte-Carlo directions: a float64 device tensor of rows (x, y, z)."""
    import torch
    check_tensor(waves, torch.float64, 3, name)
    if waves.numel() % 3:
        raise ValueError(f"{name} must hold rows of 3 doubles (got {waves.numel()} values)")


class Context:
    """A bimsim_ctx bound to a device and a stream (torch's current stream by default)."""

    def __init__(self, device=0, stream=None):
        L = lib()
        self.L = L
        h = c_void_p()
        if stream is None:
            import torch
            stream = torch.cuda.current_stream(device).cuda_stream
        rc = L.bimsim_ctx_create(device, c_void_p(stream), ctypes.byref(h))
        if rc != OK:
            raise BimsimError(rc, "bimsim_ctx_create failed (no gfx950 device?)")
        self.h = h
        # batches created on this context: destroyed before it is
        self._batches = weakref.WeakSet()

    def close(self):
        if getattr(self, "h", None):
            for b in list(getattr(self, "_batches", ())):
                b.close()
            self.L.bimsim_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != OK:
            raise BimsimError(rc, self.L.bimsim_last_error(self.h).decode())

    def sync(self):
        self._check(self.L.bimsim_sync(self.h))

    def set_series_algorithm(self, algo):
        """SERIES_AUTO, SERIES_DIRECT or SERIES_HARMONIC (FP32 scattered field)."""
        self._check(self.L.bimsim_set_series_algorithm(self.h, algo))

    def set_autotune(self, on):
        self._check(self.L.bimsim_set_autotune(self.h, 1 if on else 0))

    def set_timing(self, on):
        self._check(self.L.bimsim_set_timing(self.h, 1 if on else 0))

    STAGES = ("uf", "us", "det", "us_series")

    def timings(self):
        """{stage: (summed ms, spans)} since the previous call (waits for the events)."""
        ms = (ctypes.c_float * 4)()
        n = (c_int * 4)()
        self._check(self.L.bimsim_get_timings(self.h, ms, n))
        return {s: (ms[i], n[i]) for i, s in enumerate(self.STAGES)}

    # ---- entry points (device tensors; shapes, dtypes and sizes are
    # checked here, so that a mismatched buffer raises ValueError instead of
    # reaching the device)
    def mie_coeffs(self, spheres, kmag, A, B, quirks=QUIRKS_REF):
        import torch
        n = coeff_extent(spheres)
        check_tensor(A, torch.complex128, n, "A")
        check_tensor(B, torch.complex128, n, "B")
        S = sphere_array(spheres)
        self._check(self.L.bimsim_mie_coeffs(self.h, S, len(spheres), kmag, quirks, _ptr(A),
                                             _ptr(B)))

    def nf_uf(self, q, uR, vR, focus, kdir, kmag, amp, ca, cb, m, Uf, mode=UF_SERIES,
              prec=FP64_PARITY, quirks=QUIRKS_REF):
        check_field(Uf, prec, uR * vR, "Uf")
        self._check(self.L.bimsim_nf_uf(self.h, ctypes.byref(q), uR, vR, (c_double * 3)(*focus),
                                        (c_double * 3)(*kdir), kmag, amp, ca, cb, m, mode, prec,
                                        quirks, _ptr(Uf)))

    def nf_uf_mc(self, q, uR, vR, focus, kmag, waves, subA, Uf, prec=FP64_PARITY):
        """Incident field as the Monte-Carlo sum of the plane waves `waves`."""
        check_waves(waves)
        check_field(Uf, prec, uR * vR, "Uf")
        self._check(self.L.bimsim_nf_uf_mc(self.h, ctypes.byref(q), uR, vR,
                                           (c_double * 3)(*focus), kmag, _ptr(waves),
                                           waves.numel() // 3, subA, prec, _ptr(Uf)))

    def nf_us(self, q, uR, vR, focus, kmag, spheres, A, B, waves, subA, U, prec=FP64_PARITY,
              quirks=QUIRKS_REF, accumulate=False):
        import torch
        n = coeff_extent(spheres)
        check_tensor(A, torch.complex128, n, "A")
        check_tensor(B, torch.complex128, n, "B")
        check_waves(waves)
        check_field(U, prec, uR * vR, "U")
        S = sphere_array(spheres)
        nk = waves.numel() // 3
        self._check(self.L.bimsim_nf_us(self.h, ctypes.byref(q), uR, vR, (c_double * 3)(*focus),
                                        kmag, S, len(spheres), _ptr(A), _ptr(B), _ptr(waves), nk,
                                        subA, prec, quirks, _ptr(U), 1 if accumulate else 0))

    def nf_sum_uf(self, q, uR, vR, spheres, Uf, U, prec=FP64_PARITY):
        check_field(Uf, prec, uR * vR, "Uf")
        check_field(U, prec, uR * vR, "U")
        S = sphere_array(spheres)
        self._check(self.L.bimsim_nf_sum_uf(self.h, ctypes.byref(q), uR, vR, S, len(spheres),
                                            prec, _ptr(Uf), _ptr(U)))

    def uf_bessel_lut(self, m, kmag, dmin, dmax, dR, lut, prec=FP64_PARITY):
        check_real(lut, prec, dR * (m + 1), "lut")
        self._check(self.L.bimsim_uf_bessel_lut(self.h, m, kmag, dmin, dmax, dR, prec, _ptr(lut)))

    def nf_uf_lut(self, q, uR, vR, focus, kdir, amp, ca, cb, m, lut, dR, dmin, dmax, Uf,
                  prec=FP64_PARITY):
        check_real(lut, prec, dR * (m + 1), "lut")
        check_field(Uf, prec, uR * vR, "Uf")
        self._check(self.L.bimsim_nf_uf_lut(self.h, ctypes.byref(q), uR, vR,
                                            (c_double * 3)(*focus), (c_double * 3)(*kdir), amp,
                                            ca, cb, m, _ptr(lut), dR, dmin, dmax, prec, _ptr(Uf)))

    def sphere_lut(self, sphere, A, B, kmag, dmin, dmax, dR, aR, thetaR, Usp, Uip,
                   prec=FP64_PARITY):
        import torch
        n = coeff_extent([sphere])
        check_tensor(A, torch.complex128, n, "A")
        check_tensor(B, torch.complex128, n, "B")
        check_field(Usp, prec, thetaR * dR, "Usp")
        check_field(Uip, prec, thetaR * aR, "Uip", optional=True)
        S = sphere_array([sphere])
        self._check(self.L.bimsim_sphere_lut(self.h, S, _ptr(A), _ptr(B), kmag, dmin, dmax, dR,
                                             aR, thetaR, prec, _ptr(Usp), _ptr(Uip)))

    def nf_us_lut(self, q, uR, vR, focus, kmag, sphere, Usp, dR, Uip, aR, thetaR, dmin, dmax,
                  waves, subA, U, prec=FP64_PARITY):
        check_field(Usp, prec, thetaR * dR, "Usp")
        check_field(Uip, prec, thetaR * aR, "Uip", optional=True)
        check_waves(waves)
        check_field(U, prec, uR * vR, "U")
        S = sphere_array([sphere])
        nk = waves.numel() // 3
        self._check(self.L.bimsim_nf_us_lut(self.h, ctypes.byref(q), uR, vR,
                                            (c_double * 3)(*focus), kmag, S, _ptr(Usp), dR,
                                            _ptr(Uip), aR, thetaR, dmin, dmax, _ptr(waves), nk,
                                            subA, prec, _ptr(U)))

    def nf_interpolate(self, deg, uR, vR, U_src, Uf_src, U, Uf, prec=FP64_PARITY):
        n = uR * vR
        if (U_src is None) != (U is None) or (Uf_src is None) != (Uf is None):
            raise ValueError("each field to rotate needs both its source and its output")
        for t, name in ((U_src, "U_src"), (Uf_src, "Uf_src"), (U, "U"), (Uf, "Uf")):
            check_field(t, prec, n, name, optional=True)
        self._check(self.L.bimsim_nf_interpolate(self.h, prec, deg, uR, vR, _ptr(U_src),
                                                 _ptr(Uf_src), _ptr(U), _ptr(Uf)))

    def sphere_radial_lut(self, sphere, kmag, dmin, dmax, dR, aR, h, j):
        import torch
        nl1 = sphere["Nl"] + 1
        check_tensor(h, torch.complex128, dR * nl1, "h", optional=True)
        check_tensor(j, torch.complex128, aR * nl1, "j", optional=True)
        S = sphere_array([sphere])
        self._check(self.L.bimsim_sphere_radial_lut(self.h, S, kmag, dmin, dmax, dR, aR, _ptr(h),
                                                    _ptr(j)))

    def field_fft(self, F, uR, vR, inverse=False, prec=FP64_PARITY):
        check_field(F, prec, uR * vR, "F")
        self._check(self.L.bimsim_field_fft(self.h, prec, _ptr(F), uR, vR, 1 if inverse else 0))

    def field_bandpass(self, F, uR, vR, du, dv, NAin, NAout, lam, prec=FP64_PARITY,
                       quirks=QUIRKS_REF):
        check_field(F, prec, uR * vR, "F")
        self._check(self.L.bimsim_field_bandpass(self.h, prec, quirks, _ptr(F), uR, vR, du, dv,
                                                 NAin, NAout, lam))

    def field_to_scalar(self, F, op, S, prec=FP64_PARITY):
        check_field(F, prec, 1, "F")
        check_real(S, prec, F.numel(), "S")
        self._check(self.L.bimsim_field_to_scalar(self.h, prec, op, _ptr(F), F.numel(), _ptr(S)))

    def field_scale(self, F, v, prec=FP64_PARITY):
        check_field(F, prec, 1, "F")
        self._check(self.L.bimsim_field_scale(self.h, prec, _ptr(F), F.numel(), v))

    def field_crop(self, src, uR, vR, u, v, su, sv, dst, prec=FP64_PARITY):
        check_field(src, prec, uR * vR, "src")
        check_field(dst, prec, su * sv, "dst")
        self._check(self.L.bimsim_field_crop(self.h, prec, _ptr(src), uR, vR, u, v, su, sv,
                                             _ptr(dst)))

    def detector_resample(self, F, Du, Dv, ss, scale, accumulate, D, prec=FP64_PARITY):
        check_field(F, prec, Du * ss * Dv * ss, "F")
        check_real(D, prec, Du * Dv, "D")
        self._check(self.L.bimsim_detector_resample(self.h, prec, _ptr(F), Du, Dv, ss, scale,
                                                    1 if accumulate else 0, _ptr(D)))

    def field_backpropagate(self, F, uR, vR, du, dv, lam, d, prec=FP64_PARITY, quirks=QUIRKS_REF):
        """Angular spectrum F times -exp(i kz d) where propagating, else 0
        (microscope.cu:423-489)."""
        check_field(F, prec, uR * vR, "F")
        self._check(self.L.bimsim_field_backpropagate(self.h, prec, quirks, _ptr(F), uR, vR, du,
                                                      dv, lam, d))

    def detector_rotate_add(self, deg, uR, vR, scale, src, D, prec=FP64_PARITY):
        """D += scale * src rotated by deg (microscope.cu:89-132)."""
        check_real(src, prec, uR * vR, "src")
        check_real(D, prec, uR * vR, "D")
        self._check(self.L.bimsim_detector_rotate_add(self.h, prec, deg, uR, vR, scale,
                                                      _ptr(src), _ptr(D)))

    def detector_ring_sums(self, uR, vR, n_samples, n_rings, src, line, prec=FP64_PARITY):
        """line[r] += the ring sums of src (microscope.cu:138-240)."""
        check_real(src, prec, uR * vR, "src")
        check_real(line, prec, n_rings, "line")
        self._check(self.L.bimsim_detector_ring_sums(self.h, prec, uR, vR, n_samples, n_rings,
                                                     _ptr(src), _ptr(line)))

    def detector_from_line(self, uR, vR, scale, line, D, prec=FP64_PARITY):
        """D = scale * line(|p - centre|) (microscope.cu:37-86)."""
        check_real(line, prec, 1, "line")
        check_real(D, prec, uR * vR, "D")
        self._check(self.L.bimsim_detector_from_line(self.h, prec, uR, vR, scale, _ptr(line),
                                                     line.numel(), _ptr(D)))

    def slice_ratio(self, op, D, Di, out, prec=FP64_PARITY):
        check_real(D, prec, 1, "D")
        check_real(Di, prec, D.numel(), "Di")
        check_real(out, prec, D.numel(), "out")
        self._check(self.L.bimsim_slice_ratio(self.h, prec, op, _ptr(D), _ptr(Di), D.numel(),
                                              _ptr(out)))

    @staticmethod
    def check_batch(instances, spheres, waves, D, Di, Abs, prec):
        """The buffers a batch reads and writes, against its instance table."""
        import torch
        n_img = max((I.image_off + I.Nd * I.Nd for I in instances), default=0)
        check_real(D, prec, n_img, "D")
        check_real(Di, prec, n_img, "Di")
        check_real(Abs, prec, n_img, "Abs", optional=True)
        nw = max((I.wave_begin + I.nk for I in instances
                  if I.uf_mode != UF_PLANE and I.ns > 0 or I.uf_mode == UF_MC), default=0)
        if nw > 0:
            check_tensor(waves, torch.float64, 3 * nw, "waves")
        for I in instances:
            if I.ns < 0 or I.sphere_begin < 0 or I.sphere_begin + I.ns > len(spheres):
                raise ValueError("an instance's sphere range lies outside the sphere list")

    def batch(self, instances, spheres, waves, D, Di, Abs=None, prec=FP32_FAST,
              quirks=QUIRKS_REF, use_graph=False):
        """A bimsim_batch over `instances` (a ctypes Instance array); runs the
        pass once.  Keep the returned Batch (and the tensors) alive while
        launching it again."""
        self.check_batch(instances, spheres, waves, D, Di, Abs, prec)
        S = sphere_array(spheres)
        h = c_void_p()
        self._check(self.L.bimsim_batch_create(self.h, instances, len(instances), S, len(spheres),
                                               _ptr(waves), prec, quirks, _ptr(D), _ptr(Di),
                                               _ptr(Abs), 1 if use_graph else 0, ctypes.byref(h)))
        b = Batch(self, h, (instances, S, waves, D, Di, Abs))
        self._batches.add(b)
        return b

    def batch_run(self, instances, spheres, waves, D, Di, Abs=None, prec=FP32_FAST,
                  quirks=QUIRKS_REF):
        self.check_batch(instances, spheres, waves, D, Di, Abs, prec)
        S = sphere_array(spheres)
        self._check(self.L.bimsim_batch_run(self.h, instances, len(instances), S, len(spheres),
                                            _ptr(waves), prec, quirks, _ptr(D), _ptr(Di),
                                            _ptr(Abs)))

    def detector_zero_filled(self, U, Uf, uR, vR, fill, du, dv, NAin, NAout, lam, D, Di,
                             Abs=None, prec=FP64_PARITY, quirks=QUIRKS_REF):
        """The detector on the field zero-filled to fill x its size for the FFT."""
        check_field(U, prec, uR * vR, "U")
        check_field(Uf, prec, uR * vR, "Uf")
        for t, name in ((D, "D"), (Di, "Di")):
            check_real(t, prec, uR * uR, name)
        check_real(Abs, prec, uR * uR, "Abs", optional=True)
        self._check(self.L.bimsim_detector_zero_filled(self.h, prec, quirks, _ptr(U), _ptr(Uf), uR,
                                                       vR, fill, du, dv, NAin, NAout, lam, _ptr(D),
                                                       _ptr(Di), _ptr(Abs)))

    def detector_noalloc(self, U, Uf, uR, vR, du, dv, NAin, NAout, lam, padding, Nd, ss, D, Di,
                         Abs=None, prec=FP64_PARITY, quirks=QUIRKS_REF):
        """bimsim_detector (U and Uf left unchanged) on the library's fixed
        work memory: no device allocation."""
        check_field(U, prec, uR * vR, "U")
        check_field(Uf, prec, uR * vR, "Uf")
        for t, name in ((D, "D"), (Di, "Di")):
            check_real(t, prec, Nd * Nd, name)
        check_real(Abs, prec, Nd * Nd, "Abs", optional=True)
        self._check(self.L.bimsim_detector_noalloc(self.h, prec, quirks, _ptr(U), _ptr(Uf), uR, vR,
                                                   du, dv, NAin, NAout, lam, padding, Nd, ss,
                                                   _ptr(D), _ptr(Di), _ptr(Abs)))

    def detector(self, U, Uf, uR, vR, du, dv, NAin, NAout, lam, padding, Nd, ss, D, Di,
                 Abs=None, prec=FP64_PARITY, quirks=QUIRKS_REF, write_filtered=True):
        check_field(U, prec, uR * vR, "U")
        check_field(Uf, prec, uR * vR, "Uf")
        for t, name in ((D, "D"), (Di, "Di")):
            check_real(t, prec, Nd * Nd, name)
        check_real(Abs, prec, Nd * Nd, "Abs", optional=True)
        self._check(self.L.bimsim_detector(self.h, prec, quirks, _ptr(U), _ptr(Uf), uR, vR, du,
                                           dv, NAin, NAout, lam, padding, Nd, ss, _ptr(D),
                                           _ptr(Di), _ptr(Abs), 1 if write_filtered else 0))


class Batch:
    """A created bimsim_batch; launch() replays the pass."""

    def __init__(self, ctx, h, keep):
        self.ctx = ctx
        self.h = h
        self._keep = keep

    def launch(self):
        self.ctx._check(self.ctx.L.bimsim_batch_launch(self.h))

    def close(self):
        # a batch lives on its context: it is gone once the context is
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            self.ctx.L.bimsim_batch_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

