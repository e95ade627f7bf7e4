"""Parity of the HIP path (libmsgpu.so through its C ABI) against the CPU oracle
and the golden fixtures written by the compiled reference.

Tolerances (FP64 path): off-diagonal matrix elements 1e-12 relative; the
finite-difference dk diagonal eps*|invf|/h; tracelog Re 1e-10*max(1,|ref|),
arg det = N Im(tracelog) compared modulo 2 pi to 1e-10 rad; dk_log_det 1e-9
relative; amplitudes and psi 1e-10 relative to max|ref|.
"""
import numpy as np
import pytest

import pyoracle as O
from conftest import wrapped

pytestmark = pytest.mark.gpu

HS = O.HARDSPHERE


@pytest.fixture(scope="module")
def G():
    import msmodel_amd as G
    G.init()
    return G


def rel(a, b):
    return np.abs(np.asarray(a) / np.asarray(b) - 1)


def medium(G, pos, model=HS, p=(0.1,)):
    pos = np.asarray(pos)
    n, d = pos.shape
    m = G.Medium(d, n, model, p)
    m.set_pos(pos)
    return m


def tl_close(got, want, n, tol=1e-10, argtol=1e-10, floor=None):
    """tracelog = log det / N: Re to tol * max(1, |Re|); arg det = N Im modulo 2 pi to argtol
    absolute (radians).  floor (optional, per cell): (Re, arg) disagreement of two
    independent FP64 CPU factorizations of the same matrix; a cell may then differ by up
    to 10 times that instead (matrices with condition numbers near 1e21-1e31, where FP64
    LU itself cannot fix arg det to 1e-10).  The achieved errors are printed."""
    got, want = np.asarray(got), np.asarray(want)
    dre = np.abs(got.real - want.real) / np.maximum(1, np.abs(want.real))
    darg = np.abs(wrapped(n * (got.imag - want.imag)))
    tre, targ = np.full(got.shape, tol), np.full(got.shape, argtol)
    if floor is not None:
        fl = np.asarray(floor)
        tre, targ = np.maximum(tre, 10 * fl[:, 0]), np.maximum(targ, 10 * fl[:, 1])
    print("N=%d, %d cells: max rel |d Re tracelog| %.2e, max |d arg det| %.2e rad (tolerance %.1e)"
          % (n, got.size, dre.max(), darg.max(), targ.max()))
    assert (dre <= tre).all(), (dre, tre)
    assert (darg <= targ).all(), (darg, targ)


def lapack_floor(m, want, n):
    """(Re, arg det) distance between the reference value and numpy's LAPACK slogdet of the
    same matrix, and between the reference value and the oracle's LU."""
    sgn, ld = np.linalg.slogdet(m)
    o = O.tracelog(m)
    return (max(abs(ld / n - want.real) / max(1, abs(want.real)), abs(o.real - want.real) / max(1, abs(want.real))),
            max(abs(wrapped(np.angle(sgn) - n * want.imag)), abs(wrapped(n * (o.imag - want.imag)))))


# --------------------------------------------------------------------- media
def test_fill_medium_bit_exact(G, ref):
    for i, (d, n, shape, ratio, seed) in enumerate(ref["media_spec"]):
        m = G.Medium(int(d), int(n), HS, (0.1,), int(shape), float(ratio)).fill(int(seed))
        assert np.array_equal(m.pos.view(np.uint64), ref["media_pos_%d" % i].view(np.uint64)), i


# --------------------------------------------------------------------- assembly
@pytest.mark.parametrize("d", [1, 2, 3, 4])
@pytest.mark.parametrize("kind", ["ms", "dk", "normalized"])
def test_assembly_vs_reference(G, ref, d, kind):
    pos = ref["mat_pos_d%d" % d]
    m = medium(G, pos)
    got = G.build_matrix(m, 7.0 - 0.3j, kind)
    want = ref["mat_%s_d%d" % ({"normalized": "norm"}.get(kind, kind), d)]
    assert np.array_equal(got, got.T)
    off = ~np.eye(pos.shape[0], dtype=bool)
    assert rel(got[off], want[off]).max() < 1e-12
    tol = 1e-12 if kind != "dk" else max(1e-12, 1e-15 * abs(O.invf(HS, (0.1,), d, 7.0 - 0.3j)) / 1e-4 /
                                           abs(want[0, 0]))
    assert rel(np.diag(got), np.diag(want)).max() < tol


@pytest.mark.parametrize("d", [1, 2, 3, 4])
def test_assembly_vs_oracle_many_k(G, d):
    pos = O.fill_medium(d, 70, 3)
    m = medium(G, pos, HS, (0.2,))
    for k in (0.3 - 0.05j, 2.5 - 1.0j, 11.0 + 0j, 23.7 - 0.4j, 38.0 - 1.9j):
        got = G.build_matrix(m, k)
        want = O.build_matrix(pos, k, HS, (0.2,))
        assert rel(got, want).max() < 1e-12, (d, k)


def test_invf_models(G, ref):
    ks = ref["invf_k"]
    for mi in range(6):
        spec = ref["invf_model_%d" % mi]
        model, p = int(spec[0]), spec[1:]
        for d in (1, 2, 3, 4):
            m = G.Medium(d, 2, model, p[:{0: 1, 1: 1, 2: 2, 3: 0}[model]] if model != 3 else (0.0,))
            got = [G.invf(m, k) for k in ks]
            assert rel(got, ref["invf_%d_d%d" % (mi, d)]).max() < 1e-12, (mi, d)


# --------------------------------------------------------------------- LU / tracelog
@pytest.mark.parametrize("n", [1, 2, 5, 63, 64, 65, 130, 200])
def test_zgetrf_matches_oracle(G, n):
    rng = np.random.default_rng(n)
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    lu, ipiv, info = G.zgetrf(a)
    olu, oipiv, oinfo = O.getrf(a)
    assert info == oinfo == 0
    assert np.array_equal(ipiv, oipiv)
    assert np.abs(lu - olu).max() < 1e-11 * max(1, np.abs(olu).max())


def test_zgesv_and_zgetrs(G):
    rng = np.random.default_rng(5)
    for n in (7, 100, 257):
        a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        b = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
        x, lu, ipiv, info = G.zgesv(a, b)
        assert info == 0
        assert np.abs(a @ x - b).max() < 1e-10 * n
        x2, info = G.zgetrs(lu, ipiv, b)
        assert np.abs(x2 - x).max() < 1e-12 * np.abs(x).max() * n


def test_tracelog_lu_dropin(G):
    pos = O.fill_medium(2, 50, 1)
    m = medium(G, pos, HS, (0.4,))
    for k in (1 - 1j, 3 - 2j, 10 - 0.5j, 10 - 10j):
        mm = O.build_matrix(pos, k, HS, (0.4,))
        t, lu = G.tracelog_lu(mm)
        want = O.tracelog(mm)
        tl_close([t], [want], 50, floor=[lapack_floor(mm, want, 50)])


def test_tracelog_lu_concurrent_calls(G):
    """Drop-in tracelog_lu called from many host threads at once (the calls are coalesced
    into shared batched factorizations, sizes mixed): every call returns what it returns
    alone, trace-log and LAPACK-format factor bit for bit; an empty matrix gives 0/0."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(21)
    sizes = [60, 130, 60, 200, 130, 60, 70, 200, 60, 130, 60, 60] * 2
    mats = [rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) for n in sizes]
    alone = [G.tracelog_lu(m) for m in mats]
    with ThreadPoolExecutor(max_workers=12) as ex:
        together = list(ex.map(G.tracelog_lu, mats))
    for (t0, lu0), (t1, lu1), m in zip(alone, together, mats):
        assert t0 == t1
        assert np.array_equal(lu0, lu1)
    for i in (0, 3):
        tl_close([alone[i][0]], [O.tracelog(mats[i])], sizes[i])
    t, _ = G.tracelog_lu(np.zeros((0, 0), np.complex128))
    assert np.isnan(t.real)


@pytest.mark.parametrize("tag,p,n", [("test", 0.4, 50), ("cfg1", 0.1, 200), ("cfg3", 0.05, 500),
                                     ("n1000", 0.05, 1000), ("d2", 0.1, 400)])
def test_tracelog_batch_fp64_vs_reference(G, ref, tag, p, n):
    pos = O.fill_medium(2, 50, 1) if tag == "test" else ref["tl_%s_pos" % tag]
    m = medium(G, pos, HS, (p,))
    ks = ref["tl_%s_k" % tag]
    out, flags = G.tracelog_batch(m, ks, G.FP64)
    want = ref["tl_%s_val" % tag]
    # the reference test's k = 6-6i and 10-10i (d=2 N=50) have condition numbers near 1e21
    # and 1e31: there the tolerance follows the CPU factorizations' own disagreement
    floor = [lapack_floor(O.build_matrix(pos, k, HS, (p,)), w, n) for k, w in zip(ks, want)] if tag == "test" else None
    tl_close(out, want, n, floor=floor)


def test_tracelog_batch_many_k_vs_oracle(G, ref):
    pos = ref["tl_cfg1_pos"]
    m = medium(G, pos, HS, (0.1,))
    ks = O.kgrid_map(9, 5, 1.0, 20.0, -1.0, 0.0)
    out, _ = G.tracelog_batch(m, ks, G.FP64)
    want = [O.tracelog(O.build_matrix(pos, k, HS, (0.1,))) for k in ks]
    tl_close(out, want, 200)


def test_tracelog_fp32_error_is_bounded(G, ref):
    pos = ref["tl_cfg3_pos"]
    m = medium(G, pos, HS, (0.05,))
    ks = ref["tl_cfg3_k"]
    o64, _ = G.tracelog_batch(m, ks, G.FP64)
    o32, _ = G.tracelog_batch(m, ks, G.FP32)
    err = np.abs(o32.real - o64.real)
    print("fp32 tracelog |dRe| per cell:", err)
    assert err.max() < 1e-4


# --------------------------------------------------------------------- dk log det
def test_dk_log_det_vs_reference(G, ref):
    pos = O.fill_medium(2, 50, 1)
    m = medium(G, pos, HS, (0.4,))
    got = G.dk_log_det_batch(m, ref["dk_test_k"], G.FP64)
    assert rel(got, ref["dk_test_val"]).max() < 1e-8
    m = medium(G, ref["tl_cfg1_pos"], HS, (0.1,))
    got = G.dk_log_det_batch(m, ref["dk_cfg1_k"], G.FP64)
    assert rel(got, ref["dk_cfg1_val"]).max() < 1e-8
    m = medium(G, ref["dk_d2_pos"], HS, (0.1,))
    got = G.dk_log_det_batch(m, ref["dk_d2_k"], G.FP64)
    assert rel(got, ref["dk_d2_val"]).max() < 1e-8
    assert abs(G.dk_log_det_ms(m, ref["dk_d2_k"][0]) / ref["dk_d2_val"][0] - 1) < 1e-8


# --------------------------------------------------------------------- solve / psi
@pytest.mark.parametrize("tag,model,p,k,grid,source", [
    ("cfg2", HS, (0.1,), 10.0, (17, 13, -8, 8, -6, 6), O.PLANE),
    ("d2", O.MAXIMUM, (0.0,), 10.0, (11, 9, -8, 8, -8, 8), O.PLANE),
    ("sph", HS, (0.1,), 4 - 0.1j, (9, 7, -5, 5, -4, 4), O.SPHERICAL)])
def test_solve_and_wavefunction_vs_reference(G, ref, tag, model, p, k, grid, source):
    pos = ref["sol_%s_pos" % tag]
    m = medium(G, pos, model, p)
    a, steps = G.solve(m, k, source, G.FP64)
    assert rel(a, ref["sol_%s_a" % tag]).max() < 1e-10
    psi = G.wavefunction_grid(m, k, a, *grid, source=source)
    want = ref["sol_%s_psi" % tag]
    fin = np.isfinite(want)
    assert np.abs(psi[fin] - want[fin]).max() < 1e-10 * np.abs(want[fin]).max()
    x = np.zeros(pos.shape[1])
    x[0] = 0.37
    x[-1] = -0.21
    one = G.eval_total_wave(m, k, a, x, source)
    want1 = O.total_wave(pos, a, k, x, source)
    assert abs(one - want1) < 1e-10 * abs(want1)


@pytest.mark.parametrize("d,k,grid", [
    (3, 10.0, (300, 3000, -3, 3, -4, 4)),         # real k: four chunks of rows, the last one ragged
    (3, 7 - 0.3j, (257, 300, -2, 2, -3, 1)),      # complex k
    (2, 6 - 0.2j, (270001, 2, -5, 5, -1, 1)),     # rows wider than a chunk: one row per chunk
    (1, 3 + 0.0j, (5, 3, -2, 2, 0, 0))])          # a grid smaller than one chunk
def test_wavefunction_grid_chunks_vs_oracle(G, d, k, grid):
    """psi grids computed in chunks of rows (downloaded while later chunks compute)
    against the oracle's eval_total_wave over get_pos_wfun's grid (incident_wave_util.c:87-96,
    wave_function_util.c:76-81), d = 1..3, real and complex k, ragged chunks."""
    pos = O.fill_medium(d, 30, 4)
    m = medium(G, pos, HS, (0.1,))
    a, _ = G.solve(m, k, O.PLANE, G.FP64)
    psi = G.wavefunction_grid(m, k, a, *grid)
    want = O.wavefunction_grid(pos, a, k, *grid)
    assert psi.shape == want.shape
    err = np.abs(psi - want).max() / np.abs(want).max()
    print("grid %s: max relative error %.2e" % (grid[:2], err))
    assert err < 1e-10


def test_mixed_solve_refines_to_fp64(G, ref):
    pos = ref["sol_cfg2_pos"]
    m = medium(G, pos, HS, (0.1,))
    a64, _ = G.solve(m, 10.0, G.PLANE, G.FP64)
    a32, steps = G.solve(m, 10.0, G.PLANE, G.MIXED, 1e-12)
    assert steps >= 1
    assert np.abs(a32 - a64).max() < 1e-10 * np.abs(a64).max()


def test_mixed_solve_at_a_resonance_falls_back_to_fp64(G, ref):
    """At a root of det M (cond M ~ 1e12, far beyond 1/eps32) complex64 refinement stalls:
    the solve must not return its unconverged iterate but the complex128 solution, whose
    residual is at the level of the complex128 backward error."""
    pos = O.fill_medium(2, 20, 1)
    m = medium(G, pos, HS, (0.4,))
    k = ref["ab_d2_roots"][0]
    a64, _ = G.solve(m, k, G.PLANE, G.FP64)
    a32, steps = G.solve(m, k, G.PLANE, G.MIXED, 1e-12)
    assert steps >= 1
    assert np.array_equal(a32, a64)
    M = O.build_matrix(pos, k, HS, (0.4,))
    phi = np.exp(1j * k * pos[:, 0])
    res = np.linalg.norm(M @ a32 - phi)
    print("|a| %.2e, residual %.2e, backward error %.2e" % (np.linalg.norm(a32), res,
          res / (np.linalg.norm(M, 2) * np.linalg.norm(a32))))
    assert np.linalg.norm(a32) > 1e6 * np.linalg.norm(phi) / np.linalg.norm(M, 2)      # truly near-singular
    assert res <= 1e-13 * 20 * np.linalg.norm(M, 2) * np.linalg.norm(a32)


def test_scattering_amplitudes_mathematica(G):
    pos = O.fill_medium(2, 20, 1)
    m = medium(G, pos, HS, (0.4,))
    xs = [G.total_cross_section(m, k) for k in (0.1, 1., 3., 10., 100.)]
    assert rel(xs, [20.52583442320853, 11.953344224711243, 9.862454676777112, 4.816530470222862,
                    0.0028798604611001526]).max() < 1e-12


def test_traceinv_inverse_iteration_mathematica(G):
    a3 = np.array([-1.997 + 0.45j, 1.875 - 1.47j, -1.641 + 0.204j, -1.74 - 1.613j, -1.197 - 1.071j,
                   1.883 - 1.765j, 1.958 + 0.195j, 1.278 + 1.105j, -1.08 + 1.842j]).reshape(3, 3, order="F")
    assert abs(G.traceinv(a3) - (-2.5885257744954697 - 0.8796395991818305j)) < 1e-13
    v0 = np.ones(3, dtype=complex) / np.sqrt(3)
    conv, v, lam = G.inverse_iteration(a3, v0, 30, 1e-12)
    assert conv == 1
    assert abs(lam - (-0.38598154234503224 + 0.15816377074842095j)) < 1e-12
    expc = np.array([0.16644079736120876, 0.4293668879243407 + 0.46904982189636174j,
                     0.2043459505806509 + 0.7253802682650067j])
    assert np.abs(v - expc).max() < 1e-10


# --------------------------------------------------------------------- resonances
def test_resonances_vs_reference_roots(G, ref):
    pos = O.fill_medium(2, 20, 1)
    m = medium(G, pos, HS, (0.4,))
    dom = tuple(ref["ab_d2_dom"])
    roots, sweeps = G.resonances(m, ref["ab_d2_starts"], dom, 4, 60, 1e-12, G.FP64)
    want = ref["ab_d2_roots"]
    # every root found on the GPU is a zero of det M: a Newton step of the oracle stays put
    for z in roots:
        step = 1.0 / O.dk_log_det(pos, z, HS, (0.4,))
        assert abs(step) < 1e-10 * abs(z), (z, step)
    # and the root set is the reference's, in both directions (assert_cvector_subset)
    assert len(roots) == len(want)
    for w in want:
        assert np.min(np.abs(roots - w)) < 1e-10 * abs(w), w
    for z in roots:
        assert np.min(np.abs(want - z)) < 1e-10 * abs(z), z
