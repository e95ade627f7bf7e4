"""ctypes bindings of the compiled reference (oracle/_ref/libmsref.so).

TEST INFRASTRUCTURE ONLY: used by tests/golden/make_ref_fixtures.py and by the
live oracle-vs-reference tests, which skip when oracle/_ref is not built.  The
library is compiled from the reference sources by oracle/ref.mk.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(_HERE, "_ref")
_LIB = None
_SHIM = None
_EXT = None


class Cplx(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


def _c(z):
    z = complex(z)
    return Cplx(z.real, z.imag)


def _z(c):
    return complex(c.re, c.im)


class Urandom(ctypes.Structure):          # nru_random_type.h:16-21
    _fields_ = [("u", ctypes.c_uint64), ("v", ctypes.c_uint64), ("w", ctypes.c_uint64), ("seed", ctypes.c_uint64)]


class ScatteringModel(ctypes.Structure):  # scattering_model_type.h:22-34
    _fields_ = [("type", ctypes.c_int), ("np", ctypes.c_int), ("p", ctypes.POINTER(ctypes.c_double)),
                ("dir", ctypes.c_char * 50), ("name", ctypes.c_char * 150),
                ("param_ascii", ctypes.c_char * 150), ("param_latex", ctypes.c_char * 150)]


class Medium(ctypes.Structure):           # medium_type.h:24-36
    _fields_ = [("d", ctypes.c_int), ("n", ctypes.c_int), ("ntot", ctypes.c_int), ("shape", ctypes.c_int),
                ("ratio", ctypes.c_double), ("scmodel", ctypes.POINTER(ScatteringModel)),
                ("alpha", ctypes.c_double), ("rnd", ctypes.POINTER(Urandom)),
                ("pos", ctypes.POINTER(ctypes.c_double))]


def available():
    return os.path.exists(os.path.join(REF_DIR, "libmsref.so")) and \
        os.path.exists(os.path.join(REF_DIR, "libmsref_shim.so"))


def lib():
    global _LIB, _SHIM
    if _LIB is None:
        L = ctypes.CDLL(os.path.join(REF_DIR, "libmsref.so"))
        dp = ctypes.POINTER(ctypes.c_double)
        mp = ctypes.POINTER(Medium)
        sig = {
            "init_random": (None, [ctypes.POINTER(Urandom), ctypes.c_uint64]),
            "next_state": (ctypes.c_uint64, [ctypes.POINTER(Urandom)]),
            "random_double": (ctypes.c_double, [ctypes.POINTER(Urandom)]),
            "random_normal": (ctypes.c_double, [ctypes.POINTER(Urandom)]),
            "init_scmodel": (ctypes.c_int, [ctypes.POINTER(ScatteringModel), ctypes.c_int, ctypes.c_int, dp]),
            "init_medium": (None, [mp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ScatteringModel),
                                   ctypes.c_int, ctypes.c_double]),
            "fill_medium": (None, [mp, ctypes.c_uint64]),
            "radius_of_ball": (ctypes.c_double, [mp]),
            "free_green": (Cplx, [ctypes.c_int, Cplx, ctypes.c_double]),
            "free_green_imag": (Cplx, [ctypes.c_int, Cplx, ctypes.c_double]),
            "dk_free_green": (Cplx, [ctypes.c_int, Cplx, ctypes.c_double]),
            "invf": (Cplx, [mp, Cplx]),
            "build_ms_matrix": (None, [mp, Cplx, dp]),
            "build_dk_ms_matrix": (None, [mp, Cplx, dp]),
            "build_normalized_matrix": (None, [mp, Cplx, dp]),
            "build_plane_wave": (None, [mp, Cplx, ctypes.c_double, dp]),
            "tracelog_lu": (Cplx, [ctypes.c_int, dp]),
            "dk_log_det_ms": (Cplx, [mp, Cplx]),
            "traceinv": (Cplx, [ctypes.c_int, dp]),
            "inverse_iteration": (ctypes.c_int, [ctypes.c_int, dp, dp, dp, ctypes.c_int, ctypes.c_double,
                                                 ctypes.c_int]),
            "total_cross_section": (ctypes.c_double, [mp, ctypes.c_double]),
            "zgesv_": (None, [ctypes.POINTER(ctypes.c_int)] * 2 + [dp, ctypes.POINTER(ctypes.c_int),
                                                                  ctypes.POINTER(ctypes.c_int), dp,
                                                                  ctypes.POINTER(ctypes.c_int),
                                                                  ctypes.POINTER(ctypes.c_int)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        S = ctypes.CDLL(os.path.join(REF_DIR, "libmsref_shim.so"))
        S.shim_aberth.restype = None
        S.shim_aberth.argtypes = [mp] + [ctypes.c_double] * 4 + [ctypes.POINTER(ctypes.c_int), dp,
                                                                ctypes.c_int, ctypes.c_int, ctypes.c_double]
        S.shim_wave_grid.restype = None
        S.shim_wave_grid.argtypes = [mp, ctypes.c_int, ctypes.c_double, ctypes.c_double, dp, ctypes.c_int,
                                     ctypes.c_int] + [ctypes.c_double] * 4 + [dp]
        _LIB, _SHIM = L, S
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def ext():
    """oracle/_ref/libmsref_ext.so: k-grid generators and the [eigenstate_list] driver."""
    global _EXT
    if _EXT is None:
        lib()
        E = ctypes.CDLL(os.path.join(REF_DIR, "libmsref_ext.so"))
        dp = ctypes.POINTER(ctypes.c_double)
        E.shim_kgrid_map.restype = None
        E.shim_kgrid_map.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 4 + [dp]
        E.shim_kgrid_cut.restype = None
        E.shim_kgrid_cut.argtypes = [ctypes.c_int] + [ctypes.c_double] * 4 + [dp]
        E.shim_kgrid_logcut.restype = None
        E.shim_kgrid_logcut.argtypes = [ctypes.c_int] + [ctypes.c_double] * 3 + [dp]
        E.shim_eigenstate_list.restype = ctypes.c_int
        E.shim_eigenstate_list.argtypes = [ctypes.POINTER(Medium), ctypes.c_int, ctypes.c_int] + \
            [ctypes.c_double] * 4 + [ctypes.c_int, ctypes.c_double, dp, dp, dp]
        _EXT = E
    return _EXT


def kgrid_map(nx, ny, xmin, xmax, ymin, ymax):
    k = np.zeros(nx * ny, dtype=np.complex128)
    ext().shim_kgrid_map(nx, ny, xmin, xmax, ymin, ymax, _dp(k))
    return k


def kgrid_cut(ns, zmin, zmax):
    k = np.zeros(3 * (ns + 2), dtype=np.complex128)
    zmin, zmax = complex(zmin), complex(zmax)
    ext().shim_kgrid_cut(ns, zmin.real, zmin.imag, zmax.real, zmax.imag, _dp(k))
    return k


def kgrid_logcut(ns, rek, imkmin, imkmax):
    k = np.zeros(ns, dtype=np.complex128)
    ext().shim_kgrid_logcut(ns, rek, imkmin, imkmax, _dp(k))
    return k


def inverse_iteration(m, v0, maxit, toler):
    """The reference inverse_iteration (ms_matrix_util.c:279-314): (converged, vector, lambda)."""
    a = np.array(m, dtype=np.complex128, order="F", copy=True)
    v = np.array(v0, dtype=np.complex128, copy=True)
    lam = np.zeros(1, dtype=np.complex128)
    conv = lib().inverse_iteration(a.shape[0], _dp(a), _dp(v), _dp(lam), maxit, toler, 0)
    return conv, v, complex(lam[0])


class RefMedium:
    """A reference Medium built by init_medium; keeps its buffers alive."""

    def __init__(self, d, n, model=0, p=(0.1,), shape=1, ratio=1.0):
        L = lib()
        self._p = np.ascontiguousarray(np.atleast_1d(np.asarray(p, dtype=np.float64)))
        sm = ScatteringModel()
        npar = 0 if model == 3 else len(self._p)
        L.init_scmodel(ctypes.byref(sm), model, npar, _dp(self._p))
        self.med = Medium()
        L.init_medium(ctypes.byref(self.med), d, n, ctypes.byref(sm), shape, ratio)
        self.d, self.n = d, n

    def fill(self, seed):
        lib().fill_medium(ctypes.byref(self.med), seed)
        return self

    @property
    def pos(self):
        return np.ctypeslib.as_array(self.med.pos, shape=(self.n * self.med.d,)).reshape(self.n, self.med.d).copy()

    def set_pos(self, pos):
        flat = np.ascontiguousarray(pos, dtype=np.float64).ravel()
        ctypes.memmove(self.med.pos, flat.ctypes.data, flat.nbytes)

    def matrix(self, k, kind="ms"):
        m = np.zeros((self.n, self.n), dtype=np.complex128, order="F")
        fn = {"ms": "build_ms_matrix", "dk": "build_dk_ms_matrix", "normalized": "build_normalized_matrix"}[kind]
        getattr(lib(), fn)(ctypes.byref(self.med), _c(k), _dp(m))
        return m

    def invf(self, k):
        return _z(lib().invf(ctypes.byref(self.med), _c(k)))

    def dk_log_det(self, k):
        return _z(lib().dk_log_det_ms(ctypes.byref(self.med), _c(k)))

    def total_cross_section(self, k):
        return lib().total_cross_section(ctypes.byref(self.med), float(k))

    def plane_wave(self, k, theta):
        v = np.zeros(self.n, dtype=np.complex128)
        lib().build_plane_wave(ctypes.byref(self.med), _c(k), theta, _dp(v))
        return v

    def aberth(self, starts, dom, nexpsup=4, maxit=50, toler=1e-12):
        lib()
        z = np.array(starts, dtype=np.complex128, copy=True)
        nz = ctypes.c_int(z.size)
        _SHIM.shim_aberth(ctypes.byref(self.med), dom[0], dom[1], dom[2], dom[3], ctypes.byref(nz), _dp(z),
                          nexpsup, maxit, toler)
        return z[: nz.value]

    def eigenstate_list(self, method, ntarget, dom, maxit, toler):
        """run_eigenstate_list_serial (ms_core.c:406-430) with method 0 determinant, 1 mineigval,
        2 invtraceinv: the states' roots, eigenvalues mu and eigenvectors (nstate x n)."""
        kroot = np.zeros(ntarget, dtype=np.complex128)
        mu = np.zeros(ntarget, dtype=np.complex128)
        eigv = np.zeros((ntarget, self.n), dtype=np.complex128)
        ns = ext().shim_eigenstate_list(ctypes.byref(self.med), method, ntarget, dom[0], dom[1], dom[2], dom[3],
                                        maxit, toler, _dp(kroot), _dp(mu), _dp(eigv))
        return kroot[:ns], mu[:ns], eigv[:ns]

    def wave_grid(self, a, k, nx, ny, xmin, xmax, ymin, ymax, source=0):
        lib()
        psi = np.zeros(nx * ny, dtype=np.complex128)
        a = np.ascontiguousarray(a, dtype=np.complex128)
        k = complex(k)
        _SHIM.shim_wave_grid(ctypes.byref(self.med), source, k.real, k.imag, _dp(a), nx, ny, xmin, xmax, ymin,
                             ymax, _dp(psi))
        return psi.reshape(ny, nx)


def rng_stream(seed, count):
    st = Urandom()
    lib().init_random(ctypes.byref(st), seed)
    return np.array([lib().next_state(ctypes.byref(st)) for _ in range(count)], dtype=np.uint64)


def tracelog(m):
    a = np.array(m, dtype=np.complex128, order="F", copy=True)
    return _z(lib().tracelog_lu(a.shape[0], _dp(a)))


def traceinv(m):
    a = np.array(m, dtype=np.complex128, order="F", copy=True)
    return _z(lib().traceinv(a.shape[0], _dp(a)))


def zgesv(m, b):
    a = np.array(m, dtype=np.complex128, order="F", copy=True)
    x = np.array(b, dtype=np.complex128, order="F", copy=True)
    n = ctypes.c_int(a.shape[0])
    nrhs = ctypes.c_int(1 if x.ndim == 1 else x.shape[1])
    ipiv = np.zeros(a.shape[0], dtype=np.int32)
    info = ctypes.c_int(0)
    lib().zgesv_(ctypes.byref(n), ctypes.byref(nrhs), _dp(a), ctypes.byref(n),
                 ipiv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dp(x), ctypes.byref(n), ctypes.byref(info))
    return x, a, ipiv, info.value


def green(d, k, r):
    return _z(lib().free_green(d, _c(k), r))


def green_imag(d, k, r):
    return _z(lib().free_green_imag(d, _c(k), r))


def dk_green(d, k, r):
    return _z(lib().dk_free_green(d, _c(k), r))
