"""ctypes bindings of the CPU parity oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker.  The product package
(msmodel_amd) never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

HARDSPHERE, SOFTSPHERE, RESONANT, MAXIMUM = 0, 1, 2, 3
CUBE, BALL, LATTICE, GAUSSIAN = 0, 1, 2, 3
PLANE, SPHERICAL = 0, 1


class _C(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


def _c(z):
    z = complex(z)
    return _C(z.real, z.imag)


def _z(c):
    return complex(c.re, c.im)


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        u64p = ctypes.POINTER(ctypes.c_uint64)
        ip = ctypes.POINTER(ctypes.c_int)
        sig = {
            "orc_rng_init": (None, [u64p, ctypes.c_uint64]),
            "orc_rng_next": (ctypes.c_uint64, [u64p]),
            "orc_rng_double": (ctypes.c_double, [u64p]),
            "orc_rng_normal": (ctypes.c_double, [u64p]),
            "orc_uball_volume": (ctypes.c_double, [ctypes.c_int]),
            "orc_uball_surface": (ctypes.c_double, [ctypes.c_int]),
            "orc_radius_of_ball": (ctypes.c_double, [ctypes.c_int, ctypes.c_int]),
            "orc_fill_medium": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                               ctypes.c_uint64, dp]),
            "orc_bessel_k": (_C, [ctypes.c_int, _C]),
            "orc_hyp0f1_reg": (_C, [ctypes.c_double, _C]),
            "orc_green": (_C, [ctypes.c_int, _C, ctypes.c_double]),
            "orc_green_imag": (_C, [ctypes.c_int, _C, ctypes.c_double]),
            "orc_dk_green": (_C, [ctypes.c_int, _C, ctypes.c_double]),
            "orc_invf": (_C, [ctypes.c_int, dp, ctypes.c_int, _C]),
            "orc_build_matrix": (None, [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, dp, _C, dp]),
            "orc_build_dk_matrix": (None, [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, dp, _C, dp]),
            "orc_build_normalized_matrix": (None, [ctypes.c_int, ctypes.c_int, dp, _C, dp]),
            "orc_build_plane_wave": (None, [ctypes.c_int, ctypes.c_int, dp, _C, ctypes.c_double, dp]),
            "orc_getrf": (ctypes.c_int, [ctypes.c_int, dp, ctypes.c_int, ip]),
            "orc_getrs": (None, [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, ip, dp, ctypes.c_int]),
            "orc_gesv": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, ip, dp, ctypes.c_int]),
            "orc_tracelog": (_C, [ctypes.c_int, dp]),
            "orc_dk_log_det": (_C, [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, dp, _C]),
            "orc_traceinv": (_C, [ctypes.c_int, dp]),
            "orc_inverse_iteration": (ctypes.c_int, [ctypes.c_int, dp, dp, dp, ctypes.c_int, ctypes.c_double]),
            "orc_total_cross_section": (ctypes.c_double, [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, dp,
                                                          ctypes.c_double]),
            "orc_incident": (_C, [ctypes.c_int, _C, ctypes.c_int, dp]),
            "orc_total_wave": (_C, [ctypes.c_int, _C, ctypes.c_int, ctypes.c_int, dp, dp, dp]),
            "orc_wavefunction_grid": (None, [ctypes.c_int, _C, ctypes.c_int, ctypes.c_int, dp, dp,
                                             ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_double, dp]),
            "orc_kgrid_map": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, dp]),
            "orc_kgrid_cut": (None, [ctypes.c_int, _C, _C, dp]),
            "orc_kgrid_logcut": (None, [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, dp]),
            "orc_halton": (None, [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, dp]),
            "orc_aberth": (None, [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, dp, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_double, ip, dp, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_double]),
            "orc_roots_muller": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, dp, ctypes.c_double,
                                                ctypes.c_double, ctypes.c_double, ctypes.c_double, ip, dp,
                                                ctypes.c_int, ctypes.c_double, ctypes.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _params(p):
    arr = np.ascontiguousarray(np.atleast_1d(np.asarray(p, dtype=np.float64)))
    if arr.size == 0:
        arr = np.zeros(1)
    return arr


# ---------------------------------------------------------------- RNG / medium
def rng_stream(seed, count):
    st = (ctypes.c_uint64 * 4)()
    lib().orc_rng_init(st, seed)
    return np.array([lib().orc_rng_next(st) for _ in range(count)], dtype=np.uint64)


def fill_medium(d, n, seed, shape=BALL, ratio=1.0):
    pos = np.zeros(n * d, dtype=np.float64)
    lib().orc_fill_medium(d, n, shape, ratio, seed, _dp(pos))
    return pos.reshape(n, d)


def radius_of_ball(d, n):
    return lib().orc_radius_of_ball(d, n)


# ---------------------------------------------------------------- scalar math
def bessel_k(twonu, z):
    return _z(lib().orc_bessel_k(twonu, _c(z)))


def green(d, k, r):
    return _z(lib().orc_green(d, _c(k), r))


def green_imag(d, k, r):
    return _z(lib().orc_green_imag(d, _c(k), r))


def dk_green(d, k, r):
    return _z(lib().orc_dk_green(d, _c(k), r))


def invf(model, p, d, k):
    pp = _params(p)
    return _z(lib().orc_invf(model, _dp(pp), d, _c(k)))


# ---------------------------------------------------------------- matrices
def build_matrix(pos, k, model=HARDSPHERE, p=(0.1,), kind="ms"):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n, d = pos.shape
    m = np.zeros((n, n), dtype=np.complex128, order="F")
    pp = _params(p)
    mp = m.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    if kind == "ms":
        lib().orc_build_matrix(d, n, _dp(pos), model, _dp(pp), _c(k), mp)
    elif kind == "dk":
        lib().orc_build_dk_matrix(d, n, _dp(pos), model, _dp(pp), _c(k), mp)
    elif kind == "normalized":
        lib().orc_build_normalized_matrix(d, n, _dp(pos), _c(k), mp)
    else:
        raise ValueError(kind)
    return m


def plane_wave(pos, k, theta_deg=0.0):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n, d = pos.shape
    v = np.zeros(n, dtype=np.complex128)
    lib().orc_build_plane_wave(d, n, _dp(pos), _c(k), theta_deg, v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return v


def getrf(a):
    a = np.array(a, dtype=np.complex128, order="F", copy=True)
    n = a.shape[0]
    ipiv = np.zeros(n, dtype=np.int32)
    info = lib().orc_getrf(n, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n,
                           ipiv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return a, ipiv, info


def getrs(lu, ipiv, b):
    lu = np.asfortranarray(lu, dtype=np.complex128)
    n = lu.shape[0]
    bb = np.array(b, dtype=np.complex128, order="F", copy=True)
    nrhs = 1 if bb.ndim == 1 else bb.shape[1]
    lib().orc_getrs(n, nrhs, lu.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n,
                    np.ascontiguousarray(ipiv, dtype=np.int32).ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                    bb.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n)
    return bb


def tracelog(m):
    a = np.array(m, dtype=np.complex128, order="F", copy=True)
    return _z(lib().orc_tracelog(a.shape[0], a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))


def dk_log_det(pos, k, model=HARDSPHERE, p=(0.1,)):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n, d = pos.shape
    pp = _params(p)
    return _z(lib().orc_dk_log_det(d, n, _dp(pos), model, _dp(pp), _c(k)))


def traceinv(m):
    a = np.array(m, dtype=np.complex128, order="F", copy=True)
    return _z(lib().orc_traceinv(a.shape[0], a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))


def total_cross_section(pos, k, model=HARDSPHERE, p=(0.1,)):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n, d = pos.shape
    pp = _params(p)
    return lib().orc_total_cross_section(d, n, _dp(pos), model, _dp(pp), float(k))


def solve(pos, k, model=HARDSPHERE, p=(0.1,), source=PLANE):
    """Amplitudes a = M^-1 phi for the incident wave at the scatterers (ms_core.c:616-620)."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n, d = pos.shape
    m = build_matrix(pos, k, model, p)
    phi = np.array([incident(source, k, pos[i]) for i in range(n)], dtype=np.complex128)
    ipiv = np.zeros(n, dtype=np.int32)
    b = phi.copy()
    info = lib().orc_gesv(n, 1, m.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n,
                          ipiv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                          b.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n)
    return b, info


def incident(source, k, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return _z(lib().orc_incident(source, _c(k), x.size, _dp(x)))


def total_wave(pos, a, k, x, source=PLANE):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n, d = pos.shape
    a = np.ascontiguousarray(a, dtype=np.complex128)
    x = np.ascontiguousarray(x, dtype=np.float64)
    return _z(lib().orc_total_wave(source, _c(k), d, n, _dp(pos), a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                   _dp(x)))


def wavefunction_grid(pos, a, k, nx, ny, xmin, xmax, ymin, ymax, source=PLANE):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n, d = pos.shape
    a = np.ascontiguousarray(a, dtype=np.complex128)
    psi = np.zeros(nx * ny, dtype=np.complex128)
    lib().orc_wavefunction_grid(source, _c(k), d, n, _dp(pos), a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                nx, ny, xmin, xmax, ymin, ymax,
                                psi.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return psi.reshape(ny, nx)


def kgrid_map(nx, ny, xmin, xmax, ymin, ymax):
    k = np.zeros(nx * ny, dtype=np.complex128)
    lib().orc_kgrid_map(nx, ny, xmin, xmax, ymin, ymax, k.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return k


def kgrid_cut(ns, zmin, zmax):
    """The 3 (ns + 2) points of a complex cut, triplet-major (complex_cut_util.c:32-36)."""
    k = np.zeros(3 * (ns + 2), dtype=np.complex128)
    lib().orc_kgrid_cut(ns, _c(zmin), _c(zmax), k.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return k


def kgrid_logcut(ns, rek, imkmin, imkmax):
    k = np.zeros(ns, dtype=np.complex128)
    lib().orc_kgrid_logcut(ns, rek, imkmin, imkmax, k.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return k


def inverse_iteration(m, v0, maxit, toler):
    """(converged, normalised vector, eigenvalue) of inverse_iteration (ms_matrix_util.c:279-314)."""
    a = np.array(m, dtype=np.complex128, order="F", copy=True)
    v = np.array(v0, dtype=np.complex128, copy=True)
    lam = np.zeros(1, dtype=np.complex128)
    conv = lib().orc_inverse_iteration(a.shape[0], a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                       v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                       lam.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), maxit, toler)
    return conv, v, complex(lam[0])


MINEIGVAL, INVTRACEINV, MINEIGVAL_PER_START = 1, 2, 3


def roots_muller(pos, starts, dom, method, model=HARDSPHERE, p=(0.1,), maxit=30, toler=1e-10):
    """Damped Muller roots of det M(k) from the starts in order (find_eigenstates_serial_2/_3)."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n, d = pos.shape
    pp = _params(p)
    z = np.array(starts, dtype=np.complex128, copy=True)
    nz = ctypes.c_int(z.size)
    lib().orc_roots_muller(d, n, _dp(pos), model, _dp(pp), dom[0], dom[1], dom[2], dom[3], ctypes.byref(nz),
                           z.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), maxit, toler, method)
    return z[: nz.value]


def halton(n, xmin, xmax, ymin, ymax):
    z = np.zeros(n, dtype=np.complex128)
    lib().orc_halton(n, xmin, xmax, ymin, ymax, z.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return z


def aberth(pos, starts, dom, model=HARDSPHERE, p=(0.1,), nexpsup=4, maxit=50, toler=1e-12):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n, d = pos.shape
    pp = _params(p)
    z = np.array(starts, dtype=np.complex128, copy=True)
    nz = ctypes.c_int(z.size)
    lib().orc_aberth(d, n, _dp(pos), model, _dp(pp), dom[0], dom[1], dom[2], dom[3], ctypes.byref(nz),
                     z.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), nexpsup, maxit, toler)
    return z[: nz.value]
