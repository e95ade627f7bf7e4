"""msmodel_amd -- MI355X implementation of the MSModel Foldy-Lax hot path.

This package is a thin ctypes binding of libmsgpu.so (HIP kernels for gfx950
behind a C ABI with the reference's signatures, see include/msgpu.h).  There is
no CPU fallback: importing the package fails if the library is not built, and
every entry point raises if the device call fails.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmsgpu.so")

HARDSPHERE, SOFTSPHERE, RESONANT, MAXIMUM = 0, 1, 2, 3
CUBE, BALL, LATTICE, GAUSSIAN = 0, 1, 2, 3
PLANE, SPHERICAL = 0, 1
FP64, MIXED, FP32, MIXED_GATED, FP64_LDLT = 0, 1, 2, 3, 4

if not os.path.exists(LIB_PATH):
    raise ImportError("msmodel_amd: %s is not built; run `python msmodel_amd/build.py` "
                      "(there is no CPU fallback)" % LIB_PATH)


class Cplx(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


def _c(z):
    z = complex(z)
    return Cplx(z.real, z.imag)


def _z(c):
    return complex(c.re, c.im)


class Urandom(ctypes.Structure):
    _fields_ = [("u", ctypes.c_uint64), ("v", ctypes.c_uint64), ("w", ctypes.c_uint64), ("seed", ctypes.c_uint64)]


class ScModel(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("np", ctypes.c_int), ("p", ctypes.POINTER(ctypes.c_double)),
                ("dir", ctypes.c_char * 50), ("name", ctypes.c_char * 150),
                ("param_ascii", ctypes.c_char * 150), ("param_latex", ctypes.c_char * 150)]


class CMedium(ctypes.Structure):
    _fields_ = [("d", ctypes.c_int), ("n", ctypes.c_int), ("ntot", ctypes.c_int), ("shape", ctypes.c_int),
                ("ratio", ctypes.c_double), ("scmodel", ctypes.POINTER(ScModel)), ("alpha", ctypes.c_double),
                ("rnd", ctypes.POINTER(Urandom)), ("pos", ctypes.POINTER(ctypes.c_double))]


class IWave(ctypes.Structure):
    _fields_ = [("source", ctypes.c_int), ("k", Cplx), ("desc_ascii", ctypes.c_char * 80),
                ("desc_latex", ctypes.c_char * 80)]


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_mp = ctypes.POINTER(CMedium)


class Request(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("med", _mp), ("mode", ctypes.c_int),
                ("nk", ctypes.c_int), ("ks", _dp), ("reeval_rel", ctypes.c_double), ("out", _dp),
                ("k", Cplx), ("source", ctypes.c_int), ("tol", ctypes.c_double), ("a", _dp),
                ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("xmin", ctypes.c_double), ("xmax", ctypes.c_double),
                ("ymin", ctypes.c_double), ("ymax", ctypes.c_double), ("psi", _dp),
                ("nz", ctypes.c_int), ("z", _dp), ("dom", ctypes.c_double * 4), ("nexpsup", ctypes.c_int),
                ("maxit", ctypes.c_int), ("toler", ctypes.c_double), ("roots", _dp), ("nroots", _ip)]


class Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("count", ctypes.c_int), ("queue_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double)]

lib = ctypes.CDLL(LIB_PATH)
_SIG = {
    "msgpu_init": (ctypes.c_int, []),
    "msgpu_last_error": (ctypes.c_char_p, []),
    "msgpu_profile_begin": (ctypes.c_int, []),
    "msgpu_profile_end": (ctypes.c_int, [_dp, _dp, _dp, _ip, ctypes.c_int]),
    "msgpu_tracelog_batch": (ctypes.c_int, [_mp, ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_double, _ip]),
    "msgpu_dk_log_det_batch": (ctypes.c_int, [_mp, ctypes.c_int, _dp, _dp, ctypes.c_int]),
    "msgpu_tracelog_batch_noalloc": (ctypes.c_int, [_mp, ctypes.c_int, _dp, _dp, ctypes.c_int]),
    "msgpu_tracelog_stats": (ctypes.c_int, [_mp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]),
    "msgpu_solve": (ctypes.c_int, [_mp, ctypes.POINTER(IWave), _dp, ctypes.c_int, ctypes.c_double]),
    "msgpu_eigenstates": (ctypes.c_int, [_mp, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double, _dp, _dp, _ip,
                                         _ip]),
    "msgpu_scattering_multi": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_mp), ctypes.c_int, _ip, _dp, ctypes.c_int,
                                              ctypes.c_int, _dp, _dp]),
    "msgpu_roots_muller": (ctypes.c_int, [_mp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          _ip, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int]),
    "msgpu_wavefunction_grid": (ctypes.c_int, [_mp, ctypes.POINTER(IWave), _dp, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                               _dp]),
    "msgpu_resonances": (ctypes.c_int, [_mp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        _ip, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int]),
    "msgpu_tracelog_multi": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_mp), ctypes.c_int, _ip, _dp, _dp,
                                            ctypes.c_int, ctypes.c_double, _ip]),
    "msgpu_dk_log_det_multi": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_mp), ctypes.c_int, _ip, _dp, _dp,
                                              ctypes.c_int]),
    "msgpu_resonances_multi": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_mp), _dp, _ip, _dp, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_double, ctypes.c_int]),
    "msgpu_autotune": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]),
    "msgpu_kgrid_map": (ctypes.c_int, [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 4 + [_dp]),
    "msgpu_kgrid_cut": (ctypes.c_int, [ctypes.c_int, Cplx, Cplx, _dp]),
    "msgpu_kgrid_logcut": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]),
    "msgpu_tune_save": (ctypes.c_int, [ctypes.c_char_p]),
    "msgpu_tune_load": (ctypes.c_int, [ctypes.c_char_p]),
    "msgpu_server_create": (ctypes.c_void_p, [ctypes.c_double, ctypes.c_int]),
    "msgpu_server_submit": (ctypes.c_longlong, [ctypes.c_void_p, ctypes.POINTER(Request)]),
    "msgpu_server_wait": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong, ctypes.POINTER(Result)]),
    "msgpu_server_stats": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.c_int]),
    "msgpu_server_destroy": (None, [ctypes.c_void_p]),
    "init_scmodel": (ctypes.c_int, [ctypes.POINTER(ScModel), ctypes.c_int, ctypes.c_int, _dp]),
    "init_medium": (None, [_mp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ScModel), ctypes.c_int, ctypes.c_double]),
    "del_medium": (None, [_mp]),
    "fill_medium": (None, [_mp, ctypes.c_uint64]),
    "radius_of_ball": (ctypes.c_double, [_mp]),
    "atom_distance": (ctypes.c_double, [_mp, ctypes.c_int, ctypes.c_int]),
    "invf": (Cplx, [_mp, Cplx]),
    "build_ms_matrix": (None, [_mp, Cplx, _dp]),
    "build_dk_ms_matrix": (None, [_mp, Cplx, _dp]),
    "build_normalized_matrix": (None, [_mp, Cplx, _dp]),
    "build_plane_wave": (None, [_mp, Cplx, ctypes.c_double, _dp]),
    "tracelog_lu": (Cplx, [ctypes.c_int, _dp]),
    "dk_log_det_ms": (Cplx, [_mp, Cplx]),
    "traceinv": (Cplx, [ctypes.c_int, _dp]),
    "inverse_iteration": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int]),
    "total_cross_section": (ctypes.c_double, [_mp, ctypes.c_double]),
    "eval_incident_wave": (Cplx, [ctypes.POINTER(IWave), ctypes.c_int, _dp]),
    "eval_total_wave": (Cplx, [ctypes.POINTER(IWave), _mp, _dp, _dp]),
    "zgetrf_": (None, [_ip, _ip, _dp, _ip, _ip, _ip]),
    "zgetrs_": (None, [ctypes.c_char_p, _ip, _ip, _dp, _ip, _ip, _dp, _ip, _ip]),
    "zgesv_": (None, [_ip, _ip, _dp, _ip, _ip, _dp, _ip, _ip]),
}
for _name, (_res, _args) in _SIG.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args

EXPORTED = sorted(_SIG)


class MsgpuError(RuntimeError):
    pass


def _check(rc, what):
    if rc < 0:
        raise MsgpuError("%s failed (%d): %s" % (what, rc, lib.msgpu_last_error().decode()))
    return rc


def _cp(a):
    return a.ctypes.data_as(_dp)


def init():
    return _check(lib.msgpu_init(), "msgpu_init")


PROFILE_CATEGORIES = ("panel", "swap_trsm", "gemm", "assembly", "tracelog", "psi", "persistent")


def profile_begin():
    _check(lib.msgpu_profile_begin(), "msgpu_profile_begin")


def profile_end():
    nc = len(PROFILE_CATEGORIES)
    ms, fl, by = np.zeros(nc), np.zeros(nc), np.zeros(nc)
    cnt = np.zeros(nc, dtype=np.int32)
    _check(lib.msgpu_profile_end(_cp(ms), _cp(fl), _cp(by), cnt.ctypes.data_as(_ip), nc), "msgpu_profile_end")
    return {c: {"ms": float(ms[i]), "flops": float(fl[i]), "bytes": float(by[i]), "launches": int(cnt[i])}
            for i, c in enumerate(PROFILE_CATEGORIES)}


class Medium:
    """A reference-layout Medium owned by libmsgpu (init_medium / fill_medium / del_medium)."""

    def __init__(self, d, n, model=HARDSPHERE, p=(0.1,), shape=BALL, ratio=1.0):
        self._p = np.ascontiguousarray(np.atleast_1d(np.asarray(p, dtype=np.float64)))
        if self._p.size == 0:
            self._p = np.zeros(1)
        sm = ScModel()
        npar = {HARDSPHERE: 1, SOFTSPHERE: 1, RESONANT: 2, MAXIMUM: 0}[model]
        lib.init_scmodel(ctypes.byref(sm), model, npar, _cp(self._p))
        self.c = CMedium()
        lib.init_medium(ctypes.byref(self.c), d, n, ctypes.byref(sm), shape, ratio)
        self.d, self.n = d, n

    def __del__(self):
        try:
            lib.del_medium(ctypes.byref(self.c))
        except Exception:
            pass

    @property
    def ref(self):
        return ctypes.byref(self.c)

    def fill(self, seed, radius=None):
        """fill_medium(seed); with radius R, positions are then scaled by R/radius_of_ball."""
        lib.fill_medium(self.ref, seed)
        if radius is not None:
            self.set_pos(self.pos * (radius / lib.radius_of_ball(self.ref)))
        return self

    @property
    def pos(self):
        return np.ctypeslib.as_array(self.c.pos, shape=(self.n * self.c.d,)).reshape(self.n, self.c.d).copy()

    def set_pos(self, pos):
        flat = np.ascontiguousarray(pos, dtype=np.float64).ravel()
        assert flat.size == self.n * self.c.d
        ctypes.memmove(self.c.pos, flat.ctypes.data, flat.nbytes)
        return self


def _iwave(k, source):
    w = IWave()
    w.source = source
    w.k = _c(k)
    return w


# ------------------------------------------------------------------ batched entry points
def tracelog_batch(med, ks, mode=FP64, reeval_rel=1e-6):
    ks = np.ascontiguousarray(ks, dtype=np.complex128)
    out = np.zeros_like(ks)
    flags = np.zeros(ks.size, dtype=np.int32)
    _check(lib.msgpu_tracelog_batch(med.ref, ks.size, _cp(ks), _cp(out), mode, reeval_rel,
                                    flags.ctypes.data_as(_ip)), "msgpu_tracelog_batch")
    return out, flags


def tracelog_batch_noalloc(med, ks, mode=FP64):
    """msgpu_tracelog_batch_noalloc: the trace-log map without device allocation (FP64/FP32)."""
    ks = np.ascontiguousarray(ks, dtype=np.complex128)
    out = np.zeros_like(ks)
    _check(lib.msgpu_tracelog_batch_noalloc(med.ref, ks.size, _cp(ks), _cp(out), mode), "msgpu_tracelog_batch_noalloc")
    return out


def tracelog_stats(med, ks, mode=FP64):
    ks = np.ascontiguousarray(ks, dtype=np.complex128)
    out = np.zeros_like(ks)
    st = np.zeros(2 * ks.size)
    _check(lib.msgpu_tracelog_stats(med.ref, ks.size, _cp(ks), _cp(out), _cp(st), mode), "msgpu_tracelog_stats")
    return out, st.reshape(-1, 2)


def dk_log_det_batch(med, ks, mode=FP64):
    ks = np.ascontiguousarray(ks, dtype=np.complex128)
    out = np.zeros_like(ks)
    _check(lib.msgpu_dk_log_det_batch(med.ref, ks.size, _cp(ks), _cp(out), mode), "msgpu_dk_log_det_batch")
    return out


CRSEC, DOS, DCS = 0, 1, 2
ROOT_MINEIGVAL, ROOT_INVTRACEINV = 1, 2


def roots_muller(med, starts, dom, maxit=30, toler=1e-12, method=ROOT_INVTRACEINV):
    """Roots of det M(k) by damped Muller on 1/tr M^-1 or the smallest eigenvalue of M(k),
    all starts together; returns (roots, rounds)."""
    z = np.array(starts, dtype=np.complex128)
    nz = ctypes.c_int(z.size)
    rounds = _check(lib.msgpu_roots_muller(med.ref, dom[0], dom[1], dom[2], dom[3], ctypes.byref(nz), _cp(z), maxit,
                                           toler, method), "msgpu_roots_muller")
    return z[:nz.value].copy(), rounds


def scattering_multi(meds, med_of, ks, kind=CRSEC, thetas=None):
    """Total cross section (CRSEC), density of states (DOS) or differential cross section
    (DCS, per angle in degrees) of items (meds[med_of[i]], real ks[i]) in one batch."""
    ks = np.ascontiguousarray(ks, dtype=np.float64)
    mi = np.ascontiguousarray(med_of, dtype=np.int32)
    th = np.ascontiguousarray(thetas if thetas is not None else np.zeros(0), dtype=np.float64)
    nang = th.size if kind == DCS else 0
    out = np.zeros(ks.size * nang if kind == DCS else ks.size)
    _check(lib.msgpu_scattering_multi(len(meds), _media(meds), ks.size, mi.ctypes.data_as(_ip), ks.ctypes.data_as(_dp),
                                      kind, nang, th.ctypes.data_as(_dp), out.ctypes.data_as(_dp)),
           "msgpu_scattering_multi")
    return out.reshape(ks.size, nang) if kind == DCS else out


def eigenstates(med, kroot, maxit=50, toler=1e-10):
    """Eigenvectors of M(k_r) at resonances k_r by inverse power iteration (one batch).
    Returns (vecs[nroot, N], mu, conv, suspect) with the check_eigenstate_list flags."""
    kr = np.ascontiguousarray(kroot, dtype=np.complex128)
    vecs = np.zeros((kr.size, med.n), dtype=np.complex128)
    mu = np.zeros(kr.size, dtype=np.complex128)
    conv = np.zeros(kr.size, dtype=np.int32)
    suspect = np.zeros(kr.size, dtype=np.int32)
    _check(lib.msgpu_eigenstates(med.ref, kr.size, _cp(kr), maxit, toler, _cp(vecs), _cp(mu),
                                 conv.ctypes.data_as(_ip), suspect.ctypes.data_as(_ip)), "msgpu_eigenstates")
    return vecs, mu, conv.astype(bool), suspect.astype(bool)


def solve(med, k, source=PLANE, mode=FP64, tol=0.0):
    a = np.zeros(med.n, dtype=np.complex128)
    w = _iwave(k, source)
    steps = _check(lib.msgpu_solve(med.ref, ctypes.byref(w), _cp(a), mode, tol), "msgpu_solve")
    return a, steps


def wavefunction_grid(med, k, a, nx, ny, xmin, xmax, ymin, ymax, source=PLANE):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    psi = np.zeros(nx * ny, dtype=np.complex128)
    w = _iwave(k, source)
    _check(lib.msgpu_wavefunction_grid(med.ref, ctypes.byref(w), _cp(a), nx, ny, xmin, xmax, ymin, ymax, _cp(psi)),
           "msgpu_wavefunction_grid")
    return psi.reshape(ny, nx)


def resonances(med, starts, dom, nexpsup=4, maxit=50, toler=1e-12, mode=FP64):
    z = np.array(starts, dtype=np.complex128, copy=True)
    nz = ctypes.c_int(z.size)
    sweeps = _check(lib.msgpu_resonances(med.ref, dom[0], dom[1], dom[2], dom[3], ctypes.byref(nz), _cp(z), nexpsup,
                                         maxit, toler, mode), "msgpu_resonances")
    return z[: nz.value], sweeps


# ------------------------------------------------------------------ drop-in entry points
def build_matrix(med, k, kind="ms"):
    m = np.zeros((med.n, med.n), dtype=np.complex128, order="F")
    fn = {"ms": lib.build_ms_matrix, "dk": lib.build_dk_ms_matrix, "normalized": lib.build_normalized_matrix}[kind]
    fn(med.ref, _c(k), _cp(m))
    return m


def plane_wave(med, k, theta_deg):
    v = np.zeros(med.n, dtype=np.complex128)
    lib.build_plane_wave(med.ref, _c(k), theta_deg, _cp(v))
    return v


def tracelog_lu(m):
    a = np.array(m, dtype=np.complex128, order="F", copy=True)
    t = _z(lib.tracelog_lu(a.shape[0], _cp(a)))
    return t, a


def dk_log_det_ms(med, k):
    return _z(lib.dk_log_det_ms(med.ref, _c(k)))


def traceinv(m):
    a = np.array(m, dtype=np.complex128, order="F", copy=True)
    return _z(lib.traceinv(a.shape[0], _cp(a)))


def inverse_iteration(m, vec, maxit=50, toler=1e-10):
    a = np.array(m, dtype=np.complex128, order="F", copy=True)
    v = np.array(vec, dtype=np.complex128, copy=True)
    lam = np.zeros(1, dtype=np.complex128)
    conv = lib.inverse_iteration(a.shape[0], _cp(a), _cp(v), _cp(lam), maxit, toler, 0)
    return conv, v, complex(lam[0])


def total_cross_section(med, k):
    return lib.total_cross_section(med.ref, float(k))


def eval_total_wave(med, k, a, x, source=PLANE):
    w = _iwave(k, source)
    a = np.ascontiguousarray(a, dtype=np.complex128)
    x = np.ascontiguousarray(x, dtype=np.float64)
    return _z(lib.eval_total_wave(ctypes.byref(w), med.ref, _cp(a), _cp(x)))


def invf(med, k):
    return _z(lib.invf(med.ref, _c(k)))


def zgesv(a, b):
    a = np.array(a, dtype=np.complex128, order="F", copy=True)
    b = np.array(b, dtype=np.complex128, order="F", copy=True)
    n = ctypes.c_int(a.shape[0])
    nrhs = ctypes.c_int(1 if b.ndim == 1 else b.shape[1])
    ipiv = np.zeros(a.shape[0], dtype=np.int32)
    info = ctypes.c_int(0)
    lib.zgesv_(ctypes.byref(n), ctypes.byref(nrhs), _cp(a), ctypes.byref(n), ipiv.ctypes.data_as(_ip), _cp(b),
               ctypes.byref(n), ctypes.byref(info))
    return b, a, ipiv, info.value


def zgetrf(a):
    a = np.array(a, dtype=np.complex128, order="F", copy=True)
    n = ctypes.c_int(a.shape[0])
    ipiv = np.zeros(a.shape[0], dtype=np.int32)
    info = ctypes.c_int(0)
    lib.zgetrf_(ctypes.byref(n), ctypes.byref(n), _cp(a), ctypes.byref(n), ipiv.ctypes.data_as(_ip),
                ctypes.byref(info))
    return a, ipiv, info.value


def zgetrs(lu, ipiv, b):
    lu = np.asfortranarray(lu, dtype=np.complex128)
    b = np.array(b, dtype=np.complex128, order="F", copy=True)
    n = ctypes.c_int(lu.shape[0])
    nrhs = ctypes.c_int(1 if b.ndim == 1 else b.shape[1])
    info = ctypes.c_int(0)
    ip = np.ascontiguousarray(ipiv, dtype=np.int32)
    lib.zgetrs_(b"N", ctypes.byref(n), ctypes.byref(nrhs), _cp(lu), ctypes.byref(n), ip.ctypes.data_as(_ip), _cp(b),
                ctypes.byref(n), ctypes.byref(info))
    return b, info.value


# ------------------------------------------------------------------ autotuner
def kgrid_map(nx, ny, xmin, xmax, ymin, ymax):
    k = np.zeros(nx * ny, dtype=np.complex128)
    _check(lib.msgpu_kgrid_map(nx, ny, xmin, xmax, ymin, ymax, _cp(k)), "msgpu_kgrid_map")
    return k


def kgrid_cut(ns, zmin, zmax):
    k = np.zeros(3 * (ns + 2), dtype=np.complex128)
    _check(lib.msgpu_kgrid_cut(ns, _c(zmin), _c(zmax), _cp(k)), "msgpu_kgrid_cut")
    return k


def kgrid_logcut(ns, rek, imkmin, imkmax):
    k = np.zeros(ns, dtype=np.complex128)
    _check(lib.msgpu_kgrid_logcut(ns, rek, imkmin, imkmax, _cp(k)), "msgpu_kgrid_logcut")
    return k


def autotune(n, mode=FP64, nitems=512):
    """Pick the fastest swap + TRSM shape, chunk size and stream count for size n."""
    rep = np.zeros(8)
    _check(lib.msgpu_autotune(n, mode, nitems, _cp(rep)), "msgpu_autotune")
    return {"swap_variant": int(rep[0]), "chunk": int(rep[1]), "streams": int(rep[2]), "best_ms": float(rep[3]),
            "row_major": int(rep[4]), "panel_variant": int(rep[5]), "graph": int(rep[6]), "block128": int(rep[7])}


def tuning_table_path():
    """The tuning table loaded at initialisation: MSGPU_TUNE_FILE, else the table shipped
    beside the library (None if neither is there)."""
    env = os.environ.get("MSGPU_TUNE_FILE")
    if env is not None:
        return env or None
    path = os.path.join(os.path.dirname(LIB_PATH), "tuning_gfx950.txt")
    return path if os.path.exists(path) else None


def tune_save(path):
    return _check(lib.msgpu_tune_save(path.encode()), "msgpu_tune_save")


def tune_load(path):
    return _check(lib.msgpu_tune_load(path.encode()), "msgpu_tune_load")


# ------------------------------------------------------------------ batches over several media
def _media(meds):
    arr = (_mp * len(meds))()
    for i, m in enumerate(meds):
        arr[i] = ctypes.pointer(m.c)
    return arr


def tracelog_multi(meds, med_of, ks, mode=FP64, reeval_rel=1e-6):
    """Trace-log of items (meds[med_of[i]], ks[i]); returns (out, flags)."""
    ks = np.ascontiguousarray(ks, dtype=np.complex128)
    mi = np.ascontiguousarray(med_of, dtype=np.int32)
    assert mi.size == ks.size
    out = np.zeros_like(ks)
    flags = np.zeros(ks.size, dtype=np.int32)
    _check(lib.msgpu_tracelog_multi(len(meds), _media(meds), ks.size, mi.ctypes.data_as(_ip), _cp(ks), _cp(out),
                                    mode, reeval_rel, flags.ctypes.data_as(_ip)), "msgpu_tracelog_multi")
    return out, flags


def dk_log_det_multi(meds, med_of, ks, mode=FP64):
    ks = np.ascontiguousarray(ks, dtype=np.complex128)
    mi = np.ascontiguousarray(med_of, dtype=np.int32)
    assert mi.size == ks.size
    out = np.zeros_like(ks)
    _check(lib.msgpu_dk_log_det_multi(len(meds), _media(meds), ks.size, mi.ctypes.data_as(_ip), _cp(ks), _cp(out),
                                      mode), "msgpu_dk_log_det_multi")
    return out


def resonances_multi(meds, starts, doms, nexpsup=4, maxit=50, toler=1e-12, mode=FP64):
    """starts[m], doms[m] = (xmin, xmax, ymin, ymax) per medium; returns ([roots of medium m], sweeps)."""
    nz = np.array([len(s) for s in starts], dtype=np.int32)
    z = np.ascontiguousarray(np.concatenate([np.asarray(s, dtype=np.complex128) for s in starts]))
    dm = np.ascontiguousarray(np.asarray(doms, dtype=np.float64).reshape(-1))
    assert dm.size == 4 * len(meds) and nz.size == len(meds)
    offs = np.concatenate([[0], np.cumsum(nz)[:-1]])
    sweeps = _check(lib.msgpu_resonances_multi(len(meds), _media(meds), _cp(dm), nz.ctypes.data_as(_ip), _cp(z),
                                               nexpsup, maxit, toler, mode), "msgpu_resonances_multi")
    return [z[o: o + c].copy() for o, c in zip(offs, nz)], sweeps


# ------------------------------------------------------------------ serving loop
REQ_LOGDET, REQ_SOLVE, REQ_RESONANCE = 0, 1, 2


class Server:
    """msgpu_server_*: requests admitted by device footprint and served with continuous
    batching.  submit_* return a handle; wait(handle) returns (outputs, result dict)."""

    def __init__(self, hbm_gb=64.0, max_active=64):
        init()
        self.h = lib.msgpu_server_create(hbm_gb, max_active)
        if not self.h:
            raise MsgpuError("msgpu_server_create failed: %s" % lib.msgpu_last_error().decode())
        self._live = {}

    def _submit(self, req, keep, outputs):
        rid = lib.msgpu_server_submit(self.h, ctypes.byref(req))
        _check(rid, "msgpu_server_submit")
        self._live[rid] = (req, keep, outputs)
        return rid

    def submit_logdet(self, med, ks, mode=FP64, reeval_rel=1e-6):
        ks = np.ascontiguousarray(ks, dtype=np.complex128)
        out = np.zeros_like(ks)
        r = Request(kind=REQ_LOGDET, med=ctypes.pointer(med.c), mode=mode, nk=ks.size, ks=_cp(ks),
                    reeval_rel=reeval_rel, out=_cp(out))
        return self._submit(r, (ks,), {"tracelog": out})

    def submit_solve(self, med, k, source=PLANE, mode=FP64, tol=0.0, grid=None):
        a = np.zeros(med.n, dtype=np.complex128)
        r = Request(kind=REQ_SOLVE, med=ctypes.pointer(med.c), mode=mode, k=_c(k), source=source, tol=tol, a=_cp(a))
        outs = {"a": a}
        if grid is not None:
            nx, ny, x0, x1, y0, y1 = grid
            psi = np.zeros(nx * ny, dtype=np.complex128)
            r.nx, r.ny, r.xmin, r.xmax, r.ymin, r.ymax, r.psi = nx, ny, x0, x1, y0, y1, _cp(psi)
            outs["psi"] = psi
        return self._submit(r, (), outs)

    def submit_resonance(self, med, starts, dom, nexpsup=4, maxit=50, toler=1e-12, mode=FP64):
        z = np.ascontiguousarray(starts, dtype=np.complex128)
        roots = np.zeros_like(z)
        nr = np.zeros(1, dtype=np.int32)
        r = Request(kind=REQ_RESONANCE, med=ctypes.pointer(med.c), mode=mode, nz=z.size, z=_cp(z),
                    nexpsup=nexpsup, maxit=maxit, toler=toler, roots=_cp(roots), nroots=nr.ctypes.data_as(_ip))
        r.dom[:] = list(dom)
        return self._submit(r, (z,), {"roots": roots, "nroots": nr})

    def wait(self, rid):
        res = Result()
        st = lib.msgpu_server_wait(self.h, rid, ctypes.byref(res))
        _, _, outs = self._live.pop(rid)
        _check(st, "server request %d" % rid)
        if "roots" in outs:
            outs = {"roots": outs["roots"][: int(outs["nroots"][0])].copy()}
        return outs, {"count": res.count, "queue_ms": res.queue_ms, "total_ms": res.total_ms}

    def stats(self):
        v = np.zeros(9)
        lib.msgpu_server_stats(self.h, _cp(v), 9)
        return {"steps": int(v[0]), "batched_matrices": int(v[1]), "max_requests_per_step": int(v[2]),
                "peak_admitted_bytes": v[3], "budget_bytes": v[4], "factor_hits": int(v[5]),
                "factor_misses": int(v[6]), "kept_factor_bytes": v[7], "resident_media": int(v[8])}

    def close(self):
        if self.h:
            lib.msgpu_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
