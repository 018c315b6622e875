"""ORACLE -- test infrastructure only.

ctypes bindings of the CPU oracle (liboracle.so, built by ``make -C oracle``)
and of the reference checker library (``oracle/_ref/libbimsim_ref.so``, built
from the reference's own sources by ``make -C oracle ref``).  Only tests/,
``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg import this.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libbimsim_ref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
c_int, c_double, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p


class Quad(ctypes.Structure):
    """bimsim_quad: corner A and edges X, Y (rts::quad)."""
    _fields_ = [("A", c_double * 3), ("X", c_double * 3), ("Y", c_double * 3)]

    @classmethod
    def xy(cls, size, z=0.0):
        """--xy size z slice: A = (-s/2, -s/2, z), X = (s,0,0), Y = (0,s,0)."""
        q = cls()
        q.A[:] = [-size / 2.0, -size / 2.0, z]
        q.X[:] = [size, 0.0, 0.0]
        q.Y[:] = [0.0, size, 0.0]
        return q


class Sphere(ctypes.Structure):
    """bimsim_sphere."""
    _fields_ = [("p", c_double * 3), ("a", c_double), ("n", c_double * 2),
                ("Nl", c_int), ("coeff_off", c_int)]


def sphere_array(spheres):
    arr = (Sphere * max(len(spheres), 1))()
    for i, s in enumerate(spheres):
        arr[i].p[:] = list(s["p"])
        arr[i].a = s["a"]
        arr[i].n[:] = [s["n"].real, s["n"].imag] if isinstance(s["n"], complex) else list(s["n"])
        arr[i].Nl = s["Nl"]
        arr[i].coeff_off = s.get("coeff_off", 0)
    return arr


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _c2d(a):
    """complex128 ndarray -> float64 interleaved view."""
    return np.ascontiguousarray(a, dtype=np.complex128).view(np.float64)


class Oracle:
    def __init__(self, path=LIB_PATH):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        self.L = L = ctypes.CDLL(path)
        L.orc_gamma.restype = c_double
        L.orc_gamma.argtypes = [c_double]
        self.nthreads = int(os.environ.get("ORACLE_THREADS", os.cpu_count() or 1))

    # ---- special functions
    def gamma(self, x):
        return self.L.orc_gamma(x)

    def msta1(self, x, mp):
        return self.L.orc_msta1(c_double(x), c_int(mp))

    def msta2(self, x, n, mp):
        return self.L.orc_msta2(c_double(x), c_int(n), c_int(mp))

    def sph_jy(self, nl, x, quirks=1):
        out = [np.zeros(nl + 2) for _ in range(4)]
        self.L.orc_sph_jy(c_int(nl), c_double(x), c_int(quirks), *[_p(o) for o in out])
        return [o[: nl + 1] for o in out]

    def sph_jy_c(self, nl, z, quirks=1):
        out = [np.zeros(2 * (nl + 2)) for _ in range(4)]
        self.L.orc_sph_jy_c(c_int(nl), c_double(z.real), c_double(z.imag), c_int(quirks),
                            *[_p(o) for o in out])
        return [o.view(np.complex128)[: nl + 1] for o in out]

    def sphere_order(self, a, lam):
        return self.L.orc_sphere_order(c_double(a), c_double(lam))

    def mie_coeffs(self, nl, kmag, a, n, quirks=1):
        A = np.zeros(2 * (nl + 1))
        B = np.zeros(2 * (nl + 1))
        self.L.orc_mie_coeffs(c_int(nl), c_double(kmag), c_double(a), c_double(n.real),
                              c_double(n.imag), c_int(quirks), _p(A), _p(B))
        return A.view(np.complex128), B.view(np.complex128)

    def sbessel_up(self, nl, x, fp32=0):
        j = np.zeros(nl + 1)
        y = np.zeros(nl + 1)
        self.L.orc_sbessel_up(c_int(nl), c_double(x), c_int(fp32), _p(j), _p(y))
        return j, y

    def sbessel_up_c(self, nl, z, fp32=0):
        j = np.zeros(2 * (nl + 1))
        self.L.orc_sbessel_up_c(c_int(nl), c_double(z.real), c_double(z.imag), c_int(fp32), _p(j))
        return j.view(np.complex128)

    def legendre_up(self, nl, x, fp32=0):
        P = np.zeros(nl + 1)
        self.L.orc_legendre_up(c_int(nl), c_double(x), c_int(fp32), _p(P))
        return P

    def quad_dist(self, q, p, fp32=0):
        dmin, dmax = c_double(), c_double()
        pa = (c_double * 3)(*p)
        self.L.orc_quad_dist(ctypes.byref(q), pa, c_int(fp32), ctypes.byref(dmin), ctypes.byref(dmax))
        return dmin.value, dmax.value

    # ---- near field
    def uf_bessel_lut(self, m, kmag, dmin, dmax, dR, fp32=0):
        lut = np.zeros(dR * (m + 1))
        self.L.orc_uf_bessel_lut(c_int(m), c_double(kmag), c_double(dmin), c_double(dmax),
                                 c_int(dR), c_int(fp32), _p(lut))
        return lut

    def nf_uf(self, q, uR, vR, focus, kdir, kmag, amp, ca, cb, m, mode=0, fp32=0, quirks=1,
              lut=None, dR=0, dmin=0.0, dmax=0.0):
        Uf = np.zeros(2 * uR * vR)
        f = (c_double * 3)(*focus)
        k = (c_double * 3)(*kdir)
        rc = self.L.orc_nf_uf(ctypes.byref(q), c_int(uR), c_int(vR), f, k, c_double(kmag),
                              c_double(amp), c_double(ca), c_double(cb), c_int(m), c_int(mode),
                              c_int(fp32), c_int(quirks), _p(lut), c_int(dR), c_double(dmin),
                              c_double(dmax), _p(Uf), c_int(self.nthreads))
        assert rc == 0
        return Uf.view(np.complex128).reshape(vR, uR)

    def nf_uf_mc(self, q, uR, vR, focus, kmag, waves, subA, fp32=0):
        """Monte-Carlo incident field: sum of the plane waves `waves`."""
        Uf = np.zeros(2 * uR * vR)
        f = (c_double * 3)(*focus)
        w = np.ascontiguousarray(waves, dtype=np.float64).ravel()
        rc = self.L.orc_nf_uf_mc(ctypes.byref(q), c_int(uR), c_int(vR), f, c_double(kmag),
                                 _p(w), c_int(len(w) // 3), c_double(subA), c_int(fp32), _p(Uf),
                                 c_int(self.nthreads))
        assert rc == 0
        return Uf.view(np.complex128).reshape(vR, uR)

    def nf_us(self, q, uR, vR, focus, kmag, spheres, A, B, waves, subA, fp32=0, quirks=1, U=None):
        acc = U is not None
        Ud = _c2d(U).ravel().copy() if acc else np.zeros(2 * uR * vR)
        S = sphere_array(spheres)
        f = (c_double * 3)(*focus)
        w = np.ascontiguousarray(waves, dtype=np.float64).ravel()
        rc = self.L.orc_nf_us(ctypes.byref(q), c_int(uR), c_int(vR), f, c_double(kmag), S,
                              c_int(len(spheres)), _p(_c2d(A)), _p(_c2d(B)), _p(w),
                              c_int(len(w) // 3), c_double(subA), c_int(fp32), c_int(quirks),
                              _p(Ud), c_int(1 if acc else 0), c_int(self.nthreads))
        assert rc == 0
        return Ud.view(np.complex128).reshape(vR, uR)

    def nf_sum_uf(self, q, uR, vR, spheres, Uf, U, fp32=0):
        Ud = _c2d(U).ravel().copy()
        S = sphere_array(spheres)
        self.L.orc_nf_sum_uf(ctypes.byref(q), c_int(uR), c_int(vR), S, c_int(len(spheres)),
                             c_int(fp32), _p(_c2d(Uf).ravel()), _p(Ud))
        return Ud.view(np.complex128).reshape(vR, uR)

    # ---- LUT path
    def sphere_lut_dims(self, q, sphere, R, fp32=0):
        S = sphere_array([sphere])
        aR, dR = c_int(), c_int()
        dmin, dmax = c_double(), c_double()
        self.L.orc_sphere_lut_dims(ctypes.byref(q), S, c_int(R), c_int(fp32), ctypes.byref(aR),
                                   ctypes.byref(dR), ctypes.byref(dmin), ctypes.byref(dmax))
        return aR.value, dR.value, dmin.value, dmax.value

    def lut_hankel(self, nl, kmag, a, dmin, dmax, dR, fp32=0):
        h = np.zeros(2 * dR * (nl + 1))
        self.L.orc_lut_hankel(c_int(nl), c_double(kmag), c_double(a), c_double(dmin),
                              c_double(dmax), c_int(dR), c_int(fp32), _p(h))
        return h.view(np.complex128).reshape(dR, nl + 1)

    def lut_bessel(self, nl, kmag, a, n, aR, fp32=0):
        j = np.zeros(2 * aR * (nl + 1))
        self.L.orc_lut_bessel(c_int(nl), c_double(kmag), c_double(a), c_double(n.real),
                              c_double(n.imag), c_int(aR), c_int(fp32), _p(j))
        return j.view(np.complex128).reshape(aR, nl + 1)

    def lut_eval(self, nl, coef, radial, thetaR, fp32=0):
        nr = radial.shape[0]
        out = np.zeros(2 * nr * thetaR)
        self.L.orc_lut_eval(c_int(nl), _p(_c2d(coef)), _p(_c2d(radial).ravel()), c_int(nr),
                            c_int(thetaR), c_int(fp32), _p(out))
        return out.view(np.complex128).reshape(thetaR, nr)

    def nf_us_lut(self, q, uR, vR, focus, kmag, sphere, Usp, Uip, dmin, dmax, waves, subA,
                  fp32=0, U=None):
        Ud = _c2d(U).ravel().copy() if U is not None else np.zeros(2 * uR * vR)
        S = sphere_array([sphere])
        f = (c_double * 3)(*focus)
        w = np.ascontiguousarray(waves, dtype=np.float64).ravel()
        thetaR, dR = Usp.shape
        aR = Uip.shape[1] if Uip is not None else 0
        self.L.orc_nf_us_lut(ctypes.byref(q), c_int(uR), c_int(vR), f, c_double(kmag), S,
                             _p(_c2d(Usp).ravel()), c_int(dR),
                             _p(_c2d(Uip).ravel()) if Uip is not None else None, c_int(aR),
                             c_int(thetaR), c_double(dmin), c_double(dmax), _p(w),
                             c_int(len(w) // 3), c_double(subA), c_int(fp32), _p(Ud),
                             c_int(self.nthreads))
        return Ud.view(np.complex128).reshape(vR, uR)

    def nf_interpolate(self, U, deg, centre_float=0, fp32=0):
        vR, uR = U.shape
        out = np.zeros(2 * uR * vR)
        self.L.orc_nf_interpolate(_p(_c2d(U).ravel()), _p(out), c_int(uR), c_int(vR),
                                  c_double(deg), c_int(centre_float), c_int(fp32))
        return out.view(np.complex128).reshape(vR, uR)

    def backpropagate(self, F, du, dv, lam, d, quirks=1, fp32=0):
        """Angular spectrum F (vR x uR complex) propagated by d (microscope.cu:423-489)."""
        vR, uR = F.shape
        out = _c2d(F).ravel().copy()
        self.L.orc_backpropagate(_p(out), c_int(uR), c_int(vR), c_double(du), c_double(dv),
                                 c_double(lam), c_double(d), c_int(quirks), c_int(fp32))
        return out.view(np.complex128).reshape(vR, uR)

    # ---- extended-source detector images (real uR x vR arrays)
    def det_rotate_add(self, src, D, deg, scale=1.0, fp32=0):
        vR, uR = src.shape
        out = np.ascontiguousarray(D, dtype=np.float64).copy()
        self.L.orc_det_rotate_add(_p(np.ascontiguousarray(src, dtype=np.float64).ravel()),
                                  _p(out.ravel()), c_int(uR), c_int(vR), c_double(deg),
                                  c_double(scale), c_int(fp32))
        return out

    def det_ring_sums(self, src, line, n_samples, n_rings, fp32=0):
        vR, uR = src.shape
        out = np.ascontiguousarray(line, dtype=np.float64).copy()
        assert out.size >= n_rings
        rc = self.L.orc_det_ring_sums(_p(np.ascontiguousarray(src, dtype=np.float64).ravel()),
                                      _p(out), c_int(uR), c_int(vR), c_int(n_samples),
                                      c_int(n_rings), c_int(fp32))
        assert rc == 0
        return out

    def det_from_line(self, line, uR, vR, scale=1.0, fp32=0):
        line = np.ascontiguousarray(line, dtype=np.float64)
        out = np.zeros((vR, uR))
        self.L.orc_det_from_line(_p(line), c_int(line.size), _p(out.ravel()), c_int(uR),
                                 c_int(vR), c_double(scale), c_int(fp32))
        return out

    # ---- detector
    def fft2(self, U, inverse=False):
        vR, uR = U.shape
        d = _c2d(U).ravel().copy()
        self.L.orc_fft2(_p(d), c_int(uR), c_int(vR), c_int(1 if inverse else 0))
        return d.view(np.complex128).reshape(vR, uR)

    def detector(self, U, Uf, du, dv, NAin, NAout, lam, padding, Nd, ss, fp32=0, quirks=1,
                 D=None, Di=None):
        vR, uR = U.shape
        Ud = _c2d(U).ravel().copy()
        Ufd = _c2d(Uf).ravel().copy()
        D = np.zeros(Nd * Nd) if D is None else D.ravel().astype(np.float64).copy()
        Di = np.zeros(Nd * Nd) if Di is None else Di.ravel().astype(np.float64).copy()
        Abs = np.zeros(Nd * Nd)
        rc = self.L.orc_detector(c_int(fp32), c_int(quirks), _p(Ud), _p(Ufd), c_int(uR),
                                 c_int(vR), c_double(du), c_double(dv), c_double(NAin),
                                 c_double(NAout), c_double(lam), c_int(padding), c_int(Nd),
                                 c_int(ss), _p(D), _p(Di), _p(Abs))
        assert rc == 0
        sh = (Nd, Nd)
        return (Ud.view(np.complex128).reshape(vR, uR), Ufd.view(np.complex128).reshape(vR, uR),
                D.reshape(sh), Di.reshape(sh), Abs.reshape(sh))

    # ---- sampler
    def mc_sample_na(self, N, k, NAin, NAout, rng=0, seed=0):
        out = np.zeros(3 * N)
        kk = (c_double * 3)(*k)
        self.L.orc_mc_sample_na(_p(out), c_int(N), kk, c_double(NAin), c_double(NAout),
                                c_int(rng), ctypes.c_ulonglong(seed))
        return out.reshape(N, 3)

    def glibc_rand(self, seed, n):
        out = np.zeros(n, dtype=np.int32)
        self.L.orc_glibc_rand(ctypes.c_uint(seed), c_int(n), out.ctypes.data_as(_ip))
        return out


class RefChecker:
    """The reference's own host code (oracle/_ref), for pinning the oracle."""

    def __init__(self, path=REF_PATH):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle ref`")
        self.L = L = ctypes.CDLL(path)
        L.ref_gamma.restype = c_double
        L.ref_gamma.argtypes = [c_double]

    def gamma(self, x):
        return self.L.ref_gamma(x)

    def msta1(self, x, mp):
        return self.L.ref_msta1(c_double(x), c_int(mp))

    def msta2(self, x, n, mp):
        return self.L.ref_msta2(c_double(x), c_int(n), c_int(mp))

    def sph_jy(self, nl, x):
        out = [np.zeros(nl + 2) for _ in range(4)]
        self.L.ref_sph_jy(c_int(nl), c_double(x), *[_p(o) for o in out])
        return [o[: nl + 1] for o in out]

    def sph_jy_c(self, nl, z):
        out = [np.zeros(2 * (nl + 2)) for _ in range(4)]
        self.L.ref_sph_jy_c(c_int(nl), c_double(z.real), c_double(z.imag), *[_p(o) for o in out])
        return [o.view(np.complex128)[: nl + 1] for o in out]

    def mc_sample_na(self, N, k, NAin, NAout, seed):
        out = np.zeros(3 * N)
        self.L.ref_mc_sample_na(_p(out), c_int(N), c_double(k[0]), c_double(k[1]),
                                c_double(k[2]), c_double(NAin), c_double(NAout),
                                ctypes.c_uint(seed))
        return out.reshape(N, 3)

    def sbessel_up(self, nl, x, fp32=0):
        j = np.zeros(max(nl + 1, 2))
        y = np.zeros(max(nl + 1, 2))
        self.L.ref_sbessel_up(c_int(nl), c_double(x), c_int(fp32), _p(j), _p(y))
        return j[: nl + 1], y[: nl + 1]

    def sbessel_up_c32(self, nl, z):
        j = np.zeros(2 * max(nl + 1, 2))
        self.L.ref_sbessel_up_c32(c_int(nl), c_double(z.real), c_double(z.imag), _p(j))
        return j.view(np.complex128)[: nl + 1]

    def legendre_up(self, nl, x, fp32=0):
        P = np.zeros(max(nl + 1, 2))
        self.L.ref_legendre_up(c_int(nl), c_double(x), c_int(fp32), _p(P))
        return P[: nl + 1]
