"""Pins the CPU oracle (oracle/msoracle.c) against the reference's own golden
values (Mathematica tables in src/test/test_green_bessel_data.h and
src/test/test_ms_matrix.c), against fixtures written by the compiled reference
(tests/golden/ref_fixtures.npz), and against closed forms and independent
libraries (scipy AMOS Bessel, numpy LAPACK, mpmath)."""
import numpy as np
import pytest
import scipy.special as sp

import pyoracle as O
from conftest import wrapped

HS = O.HARDSPHERE


def rel(a, b):
    return np.abs(np.asarray(a) / np.asarray(b) - 1)


# --------------------------------------------------------------------- Green / Bessel
def test_bessel_k_table(tables):
    t = tables["bessel_k_data"]
    errs = np.array([rel(O.bessel_k(int(np.floor(2 * r[0])), r[1] + 1j * r[2]), r[3] + 1j * r[4]) for r in t])
    small = np.abs(t[:, 0]) <= 6
    assert errs[small].max() < 1e-12          # orders used on the path: 0, 1, 1/2, 3/2
    assert errs.max() < 1e-5                  # nu=20 by upward recurrence (the reference misses these too)


@pytest.mark.parametrize("name,fn", [("green_data", O.green), ("green_imag_data", O.green_imag),
                                     ("dk_green_data", O.dk_green)])
def test_green_tables(tables, name, fn):
    t = tables[name]
    errs = np.array([rel(fn(int(r[0]), r[1] + 1j * r[2], 1.0), r[3] + 1j * r[4]) for r in t])
    assert errs.max() < 1e-12


def test_green_closed_forms():
    rng = np.random.default_rng(1)
    for _ in range(50):
        k = rng.uniform(0.1, 40) + 1j * rng.uniform(-2, 0.5)
        r = rng.uniform(0.01, 15)
        assert abs(O.green(3, k, r) / (-np.exp(1j * k * r) / (4 * np.pi * r)) - 1) < 1e-14
        assert abs(O.green(1, k, r) / (np.exp(1j * k * r) / (2j * k)) - 1) < 1e-14
        # d=2: -K0(-ikr)/(2 pi) = -(i/4) H0^(1)(kr)  (independent AMOS evaluation)
        assert abs(O.green(2, k, r) / (-0.25j * sp.hankel1(0, k * r)) - 1) < 1e-12
        # d=4: -(-ik/(2 pi r)) K1(-ikr)/(2 pi)
        z = -1j * k * r
        assert abs(O.green(4, k, r) / (-(-1j * k / (2 * np.pi * r)) * sp.kv(1, z) / (2 * np.pi)) - 1) < 1e-12


def test_green_fixtures(ref):
    for d in (1, 2, 3, 4):
        for kind, fn in (("green", O.green), ("green_imag", O.green_imag), ("dk_green", O.dk_green)):
            got = [fn(d, k, r) for k, r in zip(ref["green_k"], ref["green_r"])]
            assert rel(got, ref["%s_d%d" % (kind, d)]).max() < 1e-12, (kind, d)


# --------------------------------------------------------------------- scattering models
def test_invf_closed_forms():
    for k in (0.5 - 0.1j, 3.0 + 0j, 7.5 - 1j, 12 - 0.2j):
        for al in (0.05, 0.1, 0.4):
            want3 = k / (4 * np.pi) * (1 / np.tan(k * al) + 1j)
            assert abs(O.invf(HS, (al,), 3, k) / want3 - 1) < 1e-12
            want2 = 0.25j * sp.hankel1(0, k * al) / sp.jv(0, k * al)
            assert abs(O.invf(HS, (al,), 2, k) / want2 - 1) < 1e-12
    assert O.invf(O.MAXIMUM, (), 3, 4.0) == pytest.approx(1j * 4.0 / (4 * np.pi))


def test_invf_fixtures(ref):
    ks = ref["invf_k"]
    for mi in range(6):
        spec = ref["invf_model_%d" % mi]
        model, p = int(spec[0]), spec[1:]
        for d in (1, 2, 3, 4):
            got = [O.invf(model, p, d, k) for k in ks]
            assert rel(got, ref["invf_%d_d%d" % (mi, d)]).max() < 1e-12, (mi, d)


# --------------------------------------------------------------------- RNG and media
def test_rng_streams(ref):
    for s, stream in zip(ref["rng_seeds"], ref["rng_streams"]):
        assert np.array_equal(O.rng_stream(int(s), 32), stream)


def test_media_bit_exact(ref):
    for i, (d, n, shape, ratio, seed) in enumerate(ref["media_spec"]):
        pos = O.fill_medium(int(d), int(n), int(seed), int(shape), float(ratio))
        assert np.array_equal(pos.view(np.uint64), ref["media_pos_%d" % i].view(np.uint64)), i


def test_ball_inside_radius():
    for d in (2, 3, 4):
        pos = O.fill_medium(d, 300, 5)
        assert (np.linalg.norm(pos, axis=1) <= O.radius_of_ball(d, 300) * (1 + 1e-15)).all()


# --------------------------------------------------------------------- matrices
def test_matrix_fixtures(ref):
    for d in (1, 2, 3, 4):
        pos = ref["mat_pos_d%d" % d]
        for kind in ("ms", "dk", "normalized"):
            m = O.build_matrix(pos, 7.0 - 0.3j, HS, (0.1,), kind)
            want = ref["mat_%s_d%d" % ({"normalized": "norm"}.get(kind, kind), d)]
            assert np.array_equal(m, m.T)
            off = ~np.eye(m.shape[0], dtype=bool)
            assert rel(m[off], want[off]).max() < 1e-12, (d, kind)
            # the dk diagonal is a central difference with h=1e-4 of invf: ulp-level
            # differences of invf are amplified by |invf|/h
            tol = 1e-12 if kind != "dk" else 1e-15 * abs(O.invf(HS, (0.1,), d, 7.0 - 0.3j)) / 1e-4 / abs(want[0, 0])
            assert rel(np.diag(m), np.diag(want)).max() < max(tol, 1e-12), (d, kind)


# --------------------------------------------------------------------- LU / tracelog
def _tracelog_close(got, want, n, tol=1e-11):
    got, want = np.asarray(got), np.asarray(want)
    assert (np.abs(got.real - want.real) <= tol * np.maximum(1, np.abs(want.real))).all()
    # arg det = N Im(tracelog) is defined modulo 2 pi
    assert (np.abs(wrapped(n * (got.imag - want.imag))) <= 1e-8 * n).all()


def test_tracelog_mathematica():
    """Re tracelog vs the Mathematica values of test_ms_matrix.c:52-57 (d=2, N=50, hardsphere 0.4, seed 1)."""
    pos = O.fill_medium(2, 50, 1)
    ks = [1 - 1j, 3 - 2j, 6 - 6j, 10 - 0.5j, 10 - 10j]
    expc = [-0.7384505297011151 - 0.01552793781658203j, 3.198589101932406 - 0.04336958744130083j,
            22.04929061014481851 - 0.02092209133585608j, -0.2877986626571256 - 0.04027358665471191j,
            43.28298771155470174 + 0.01123031052443743j]
    for k, e in zip(ks, expc):
        t = O.tracelog(O.build_matrix(pos, k, HS, (0.4,)))
        assert abs(t.real / e.real - 1) < 5e-10
        assert abs(wrapped(50 * (t.imag - e.imag))) < 1e-6


@pytest.mark.parametrize("tag,p,n", [("test", 0.4, 50), ("cfg1", 0.1, 200), ("cfg3", 0.05, 500),
                                     ("n1000", 0.05, 1000), ("d2", 0.1, 400)])
def test_tracelog_fixtures(ref, tag, p, n):
    pos = O.fill_medium(2, 50, 1) if tag == "test" else ref["tl_%s_pos" % tag]
    ks = ref["tl_%s_k" % tag]
    if n >= 1000:
        ks = ks[:2]
    got = [O.tracelog(O.build_matrix(pos, k, HS, (p,))) for k in ks]
    _tracelog_close(got, ref["tl_%s_val" % tag][:len(ks)], n)


def test_lu_vs_numpy():
    rng = np.random.default_rng(3)
    for n in (1, 2, 7, 33, 100):
        a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        lu, ipiv, info = O.getrf(a)
        assert info == 0
        sign, logabs = np.linalg.slogdet(a)
        t = O.tracelog(a) * n
        assert abs(t.real - logabs) < 1e-11 * max(1, abs(logabs))
        assert abs(wrapped(t.imag - np.angle(sign))) < 1e-10
        b = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
        x = O.getrs(lu, ipiv, b)
        assert np.abs(a @ x - b).max() < 1e-10 * np.abs(b).max() * n


def test_lu_pivot_semantics():
    """izamax rule: max |Re|+|Im|, first index on ties; exact zero pivot -> info."""
    a = np.array([[1, 2], [1 + 1j, 2j]], dtype=complex)      # |1+1j|_1 = 2 > 1 -> row 1
    _, ipiv, _ = O.getrf(a)
    assert ipiv[0] == 2
    a = np.array([[2j, 1], [2, 3]], dtype=complex)           # tie 2 == 2 -> first (row 0)
    _, ipiv, _ = O.getrf(a)
    assert ipiv[0] == 1
    a = np.array([[0, 1], [0, 2]], dtype=complex)
    _, _, info = O.getrf(a)
    assert info == 1
    assert np.isneginf(O.tracelog(a).real)


def test_tracelog_n2_closed_form():
    pos = np.array([[0.0, 0.0, 0.0], [0.3, -0.4, 1.2]])
    r = np.linalg.norm(pos[1])
    for k in (2.0 - 0.1j, 6.5 - 0.9j):
        f = O.invf(HS, (0.1,), 3, k)
        g = -O.green(3, k, r)
        det = f * f - g * g
        t = O.tracelog(O.build_matrix(pos, k, HS, (0.1,)))
        assert abs(2 * t.real - np.log(abs(det))) < 1e-13
        assert abs(wrapped(2 * t.imag - np.angle(det))) < 1e-13


# --------------------------------------------------------------------- dk_log_det
def test_dk_log_det_fixtures(ref):
    pos = O.fill_medium(2, 50, 1)
    got = [O.dk_log_det(pos, k, HS, (0.4,)) for k in ref["dk_test_k"]]
    assert rel(got, ref["dk_test_val"]).max() < 1e-8
    pos = ref["tl_cfg1_pos"]
    got = [O.dk_log_det(pos, k, HS, (0.1,)) for k in ref["dk_cfg1_k"]]
    assert rel(got, ref["dk_cfg1_val"]).max() < 1e-8
    got = [O.dk_log_det(ref["dk_d2_pos"], k, HS, (0.1,)) for k in ref["dk_d2_k"]]
    assert rel(got, ref["dk_d2_val"]).max() < 1e-8


def test_dk_log_det_is_derivative():
    """d/dk log det M = tr(M^-1 M'): compare with a central difference of N*tracelog."""
    pos = O.fill_medium(3, 40, 2)
    for k in (3.0 - 0.2j, 8.0 - 0.6j):
        h = 1e-5
        f = lambda z: 40 * O.tracelog(O.build_matrix(pos, z, HS, (0.1,)))
        fd = (f(k + h) - f(k - h)) / (2 * h)
        # the branch of Im tracelog can jump by 2 pi / N between k-h and k+h; use the real part
        got = O.dk_log_det(pos, k, HS, (0.1,))
        assert abs(got.real - fd.real) < 1e-5 * abs(got)


# --------------------------------------------------------------------- solve / wavefunction
def test_scattering_amplitudes_mathematica(ref):
    """test_ms_matrix.c:229-231: total amplitude b(theta)^H a for d=2 N=20 hardsphere 0.4, k=1."""
    pos = O.fill_medium(2, 20, 1)
    a, info = O.solve(pos, 1.0, HS, (0.4,))
    assert info == 0
    assert rel(a, ref["sol_test_a"]).max() < 1e-12
    expc = [3.9719373901881365 - 11.9533442247113j, 0.02483484083147164 - 8.152614522250493j,
            -3.569044782687082 + 3.1417554077424166j, -1.6098304636914875 + 4.718906749770716j,
            4.510670835267466 + 5.531912234005656j]
    for th, e in zip([0., 30, 90, 120, 180], expc):
        atot = np.vdot(O.plane_wave(pos, 1.0, th), a)
        assert abs(atot / e - 1) < 1e-12
    xs = [O.total_cross_section(pos, k, HS, (0.4,)) for k in (0.1, 1., 3., 10., 100.)]
    assert rel(xs, [20.52583442320853, 11.953344224711243, 9.862454676777112, 4.816530470222862,
                    0.0028798604611001526]).max() < 1e-12


@pytest.mark.parametrize("tag,model,p,k,grid,source", [
    ("cfg2", HS, (0.1,), 10.0, (17, 13, -8, 8, -6, 6), O.PLANE),
    ("d2", O.MAXIMUM, (), 10.0, (11, 9, -8, 8, -8, 8), O.PLANE),
    ("sph", HS, (0.1,), 4 - 0.1j, (9, 7, -5, 5, -4, 4), O.SPHERICAL)])
def test_solve_and_wavefunction_fixtures(ref, tag, model, p, k, grid, source):
    pos = ref["sol_%s_pos" % tag]
    a, info = O.solve(pos, k, model, p, source)
    assert rel(a, ref["sol_%s_a" % tag]).max() < 1e-10
    psi = O.wavefunction_grid(pos, a, k, *grid, source=source)
    want = ref["sol_%s_psi" % tag]
    fin = np.isfinite(want)           # a spherical source is singular at the origin
    assert np.array_equal(fin, np.isfinite(psi))
    assert np.abs(psi[fin] - want[fin]).max() < 1e-10 * np.abs(want[fin]).max()


def test_traceinv_and_inverse_iteration(ref):
    a3 = np.array([-1.997 + 0.45j, 1.875 - 1.47j, -1.641 + 0.204j, -1.74 - 1.613j, -1.197 - 1.071j,
                   1.883 - 1.765j, 1.958 + 0.195j, 1.278 + 1.105j, -1.08 + 1.842j]).reshape(3, 3, order="F")
    assert abs(O.traceinv(a3) - (-2.5885257744954697 - 0.8796395991818305j)) < 1e-13
    # inverse iteration: the Mathematica eigenpair of test_ms_matrix.c:137-138 (30 steps, 1e-12,
    # from the constant unit vector), and the compiled reference's own result
    conv, v, lam = O.inverse_iteration(a3, np.ones(3) / np.sqrt(3), 30, 1e-12)
    assert conv == 1
    assert abs(lam - (-0.38598154234503224 + 0.15816377074842095j)) < 1e-12
    assert np.abs(v - np.array([0.16644079736120876, 0.4293668879243407 + 0.46904982189636174j,
                                0.2043459505806509 + 0.7253802682650067j])).max() < 1e-10
    assert abs(lam - ref["a3_ivit_lambda"]) < 1e-14
    assert np.abs(v - ref["a3_ivit_vec"]).max() < 1e-14
    # a wrong eigenvalue would fail: the other two eigenvalues of a3 are far from lam
    ev = np.linalg.eigvals(a3)
    assert np.min(np.abs(ev - lam)) < 1e-12 and np.sort(np.abs(ev))[1] > 2 * abs(lam)


# --------------------------------------------------------------------- grids and roots
def test_kgrid_map_order():
    k = O.kgrid_map(4, 3, 1.0, 4.0, -1.0, 0.0)
    assert k[0] == 1 + 0j and k[3] == 4 + 0j and k[4] == 1 - 0.5j and k[11] == 4 - 1j


def test_kgrids_bit_exact_vs_reference(ref):
    """get_arg_cmap, get_arg_ccut and the log-cut mesh of the compiled reference, bit for bit."""
    for i, g in enumerate(ref["grid_map_spec"]):
        got = O.kgrid_map(int(g[0]), int(g[1]), *g[2:])
        assert np.array_equal(got, ref["grid_map_%d" % i]), i
    for i, (ns, ends) in enumerate(zip(ref["grid_cut_ns"], ref["grid_cut_ends"])):
        got = O.kgrid_cut(int(ns), ends[0], ends[1])
        assert got.size == 3 * (ns + 2)
        assert np.array_equal(got, ref["grid_cut_%d" % i]), i
    for i, g in enumerate(ref["grid_log_spec"]):
        got = O.kgrid_logcut(int(g[0]), *g[1:])
        assert np.array_equal(got, ref["grid_log_%d" % i]), i


def test_kgrid_cut_geometry():
    """The triplets straddle the cut perpendicularly: the middle point of triplet il is on the
    segment (il - 1 steps from zmin), the side points are +-i dz away (dz the complex step)."""
    zmin, zmax, ns = 1 - 0.5j, 5 - 0.1j, 6
    k = O.kgrid_cut(ns, zmin, zmax).reshape(ns + 2, 3)
    dz = (zmax - zmin) / (ns - 1)
    assert np.allclose(k[:, 1], zmin + dz * (np.arange(ns + 2) - 1), rtol=0, atol=1e-14)
    assert np.allclose(k[:, 2] - k[:, 1], 1j * dz, rtol=0, atol=1e-14)
    assert np.allclose(k[:, 0] - k[:, 1], -1j * dz, rtol=0, atol=1e-14)


def test_kgrid_logcut_geometry():
    k = O.kgrid_logcut(9, 7.5, -1e-3, -10.0)
    assert np.all(k.real == 7.5)
    assert abs(k[0].imag + 1e-3) < 1e-18 and abs(k[-1].imag + 10.0) < 1e-12
    assert np.allclose(np.diff(np.log(-k.imag)), np.log(1e4) / 8, rtol=1e-12)


def assert_same_root_sets(got, want, rtol):
    """Both directions of assert_cvector_subset (src/test/assertion.c:136-162)."""
    got, want = np.asarray(got), np.asarray(want)
    assert got.size == want.size and want.size > 0, (got, want)
    err = max(max(np.min(np.abs(got - w)) / abs(w) for w in want), max(np.min(np.abs(want - g)) / abs(g) for g in got))
    print("root sets: %d roots, max relative distance %.2e" % (want.size, err))
    assert err < rtol, err


def test_aberth_fixtures(ref):
    pos = O.fill_medium(2, 20, 1)
    roots = O.aberth(pos, ref["ab_d2_starts"], tuple(ref["ab_d2_dom"]), HS, (0.4,), 4, 60, 1e-12)
    assert_same_root_sets(roots, ref["ab_d2_roots"], 1e-10)
    roots = O.aberth(ref["ab_d3_pos"], ref["ab_d3_starts"], tuple(ref["ab_d3_dom"]), HS, (0.1,), 4, 60, 1e-12)
    assert_same_root_sets(roots, ref["ab_d3_roots"], 1e-10)


def eig_case(ref, tag):
    spec = ref["eig_%s_spec" % tag]
    d, n, seed = int(spec[0]), int(spec[1]), int(spec[2])
    return dict(pos=ref["eig_%s_pos" % tag], alpha=spec[3], dom=tuple(spec[4:8]), nt=int(spec[8]),
                maxit=int(spec[9]), toler=spec[10], method=int(spec[11]), kroot=ref["eig_%s_kroot" % tag],
                mu=ref["eig_%s_mu" % tag], eigv=ref["eig_%s_eigv" % tag])


def halton_starts(nt, dom):
    """The [eigenstate_list] starts: Halton points sorted by -Im k, stable (ms_core.c:413-415)."""
    z = O.halton(nt, *dom)
    return z[np.argsort(-z.imag, kind="stable")]


@pytest.mark.parametrize("tag", ["mev_d2", "iti_d2", "mev_d3", "iti_d3", "mev_d2n60", "iti_d2n60"])
def test_muller_root_sets_vs_reference(ref, tag):
    c = eig_case(ref, tag)
    roots = O.roots_muller(c["pos"], halton_starts(c["nt"], c["dom"]), c["dom"], c["method"], HS, (c["alpha"],),
                           c["maxit"], c["toler"])
    assert_same_root_sets(roots, c["kroot"], 1e-10)


@pytest.mark.parametrize("tag", ["det_d2", "mev_d3", "iti_d2n60"])
def test_eigenvectors_vs_reference(ref, tag):
    """compute_eigenstates (ms_core.c:296-313): at each root, 50 inverse iterations from the
    constant unit vector; mu and the phase-normalised eigenvector as save_eigenstate writes them."""
    c = eig_case(ref, tag)
    n = c["pos"].shape[0]
    for k, mu, v in zip(c["kroot"], c["mu"], c["eigv"]):
        m = O.build_matrix(c["pos"], k, HS, (c["alpha"],))
        conv, vo, lam = O.inverse_iteration(m, np.ones(n) / np.sqrt(n), 50, 1e-10)
        assert abs(lam - mu) < 1e-12, (k, lam, mu)
        assert np.abs(vo - v).max() < 1e-10, (k, np.abs(vo - v).max())


def test_aberth_n2_closed_form():
    """N=2: det M = invf^2 - G(r)^2, roots solve invf(k) = +-G(k,r); check with mpmath."""
    import mpmath as mp
    r = 1.3
    pos = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
    dom = (2.0, 9.0, -2.0, 0.0)
    roots = O.aberth(pos, O.halton(8, *dom), dom, HS, (0.4,), 2, 80, 1e-13)
    assert len(roots) >= 1
    al = mp.mpf("0.4")
    for z in roots:
        def f(k):
            invf = k / (4 * mp.pi) * (mp.cot(k * al) + 1j)
            g = mp.exp(1j * k * r) / (4 * mp.pi * r)
            return invf * invf - g * g
        root = mp.findroot(f, mp.mpc(z.real, z.imag))
        assert abs(complex(root) - z) < 1e-10 * abs(z)
