"""MSGPU_FP64_LDLT: the trace-log map with the pivot-free LDL^T (scatterers in order of
increasing distance from their centroid) on the cells whose error estimate
rho = sum_r sum_{k<r} |L_rk|^2 |D_k| / |D_r| stays below 1e7, the pivoted LU elsewhere.

M(k) is complex symmetric (ms_matrix_util.c:36-41), so log det M = sum_i log D_ii of
M = L D L^T when that factorisation exists; the reference computes log det from zgetrf_
(ms_matrix_util.c:185-202).  Each cell is checked against the pivoted complex128 LU map
(which the parity tests pin to the reference), at 1e-8 in N tracelog = log det (real
part; the imaginary part modulo 2 pi), and at small N against the oracle; cells the
mode computes with the LU must equal the LU map bit for bit.  The device's max |L| and
rho are checked against an unblocked LDL^T in numpy of the oracle's matrix in the same
order.
"""
import numpy as np
import pytest

import pyoracle as O
from conftest import wrapped

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import msmodel_amd as G
    G.init()
    return G


def kmap(nx, ny, x0, x1, y0, y1):
    kr = np.linspace(x0, x1, nx)
    ki = np.linspace(y0, y1, ny)
    return (kr[None, :] + 1j * ki[::-1, None]).ravel()


def check(ldl, flags, lu, n, tol=1e-8):
    dre = np.abs(n * (ldl.real - lu.real))
    dim = np.abs(wrapped(n * (ldl.imag - lu.imag)))
    print("%d cells, %d through the LU; max |d log det| re %.2e, arg %.2e" % (ldl.size, flags.sum(), dre.max(),
                                                                                dim.max()))
    assert np.isfinite(ldl).all()
    assert dre.max() <= tol and dim.max() <= tol
    assert np.array_equal(ldl[flags == 1], lu[flags == 1])


def test_map_n1000_against_lu(G):
    """The bench's map (d=3, N=1000 unit-density ball, seed 1, alpha 0.05) on 16 x 16 cells
    over Re k in [5, 30] x Im k in [-2, 0]: the upper part of the map runs as LDL^T."""
    med = G.Medium(3, 1000, G.HARDSPHERE, (0.05,)).fill(1)
    ks = kmap(16, 16, 5, 30, -2, 0)
    lu, _ = G.tracelog_batch(med, ks, G.FP64)
    ldl, flags = G.tracelog_batch(med, ks, G.FP64_LDLT)
    check(ldl, flags, lu, 1000)
    # centre-out, the estimate stays below the bound on most of the map (239 of these 256
    # cells in numpy); the cells kept at Re k near 5 and Im k near -2 are the hardest
    assert (flags == 0).sum() >= 0.85 * ks.size
    assert flags[ks.imag > -1.0].sum() == 0


@pytest.mark.parametrize("d,n,alpha", [(3, 50, 0.1), (3, 130, 0.1), (2, 64, 0.3), (2, 400, 0.3), (1, 40, 0.2),
                                       (4, 90, 0.1)])
def test_sizes_and_dimensions(G, d, n, alpha):
    """N below, at and above one 64-row block, not multiples of 64, d = 1..4; at small N
    also against the oracle's LAPACK-style LU."""
    med = G.Medium(d, n, G.HARDSPHERE, (alpha,)).fill(7)
    ks = kmap(6, 4, 1, 9, -0.8, 0)
    lu, _ = G.tracelog_batch(med, ks, G.FP64)
    ldl, flags = G.tracelog_batch(med, ks, G.FP64_LDLT)
    check(ldl, flags, lu, n)
    if n <= 64:
        pos = O.fill_medium(d, n, 7)
        for k, t in zip(ks[:6], ldl[:6]):
            want = O.tracelog(O.build_matrix(pos, k, O.HARDSPHERE, (alpha,)))
            assert abs(n * (t.real - want.real)) <= 1e-9
            assert abs(wrapped(n * (t.imag - want.imag))) <= 1e-9


@pytest.mark.parametrize("d,n,model,p", [(3, 65, 1, (0.2,)), (2, 120, 2, (5.0, -0.5)), (2, 90, 3, (0.0,)),
                                         (3, 2, 0, (0.1,)), (2, 3, 0, (0.1,)), (4, 65, 2, (5.0, -0.5))])
def test_scattering_models_and_tiny_sizes(G, d, n, model, p):
    """The soft-sphere, resonant and maximal-coupling inverse amplitudes (the diagonal of M,
    ms_matrix_util.c:36-41) and the smallest media, N = 2, 3 (medium_util.c:37-40 rejects
    N < 2): against the LU map, at N <= 65 also against the oracle's LAPACK-style LU."""
    med = G.Medium(d, n, model, p).fill(11)
    ks = kmap(5, 3, 2, 8, -0.6, 0)
    lu, _ = G.tracelog_batch(med, ks, G.FP64)
    ldl, flags = G.tracelog_batch(med, ks, G.FP64_LDLT)
    check(ldl, flags, lu, n)
    if n <= 65:
        pos = O.fill_medium(d, n, 11)
        for k, t in zip(ks, ldl):
            want = O.tracelog(O.build_matrix(pos, k, model, p))
            assert abs(n * (t.real - want.real)) <= 1e-9, (k, t, want)
            assert abs(wrapped(n * (t.imag - want.imag))) <= 1e-9, (k, t, want)


def test_multi_medium(G):
    """Cells of several media and sizes in one call, as the LU map does them."""
    meds = [G.Medium(3, n, G.HARDSPHERE, (0.1,)).fill(s) for n, s in ((300, 1), (200, 2), (64, 3))]
    rng = np.random.default_rng(4)
    med_of = rng.integers(0, 3, 60)
    ks = rng.uniform(2, 12, 60) + 1j * rng.uniform(-1.5, 0, 60)
    lu, _ = G.tracelog_multi(meds, med_of, ks, G.FP64)
    ldl, flags = G.tracelog_multi(meds, med_of, ks, G.FP64_LDLT)
    for m in range(3):
        sel = med_of == m
        check(ldl[sel], flags[sel], lu[sel], meds[m].n)


def centre_out(pos):
    r2 = ((pos - pos.mean(axis=0)) ** 2).sum(axis=1)
    return np.argsort(r2, kind="stable")


def ldlt_estimate(m):
    """max |L| and rho of the unblocked LDL^T of m (numpy, no pivoting)."""
    a = m.copy()
    n = a.shape[0]
    s = np.zeros(n)
    dm = np.zeros(n)
    g = 0.0
    for j in range(n):
        d = a[j, j]
        dm[j] = abs(d)
        l = a[j + 1:, j] / d
        if l.size:
            g = max(g, np.abs(l).max())
        s[j + 1:] += np.abs(l) ** 2 * abs(d)
        a[j + 1:, j + 1:] -= np.outer(l, a[j, j + 1:])
    return g, float((s / dm).sum())


ESTIMATE_CASES = [(2, 500, 0.1, 3, [3 - 0.3j, 8 - 2j, 15 - 1.5j]),   # rho 6e5, 4e16, 2e12 in numpy
                  (3, 300, 0.05, 5, [10 - 0.4j, 18 - 1.2j, 6 - 2.5j])]


@pytest.mark.parametrize("d,n,alpha,seed,ks", ESTIMATE_CASES)
def test_error_estimate_and_rejection(G, d, n, alpha, seed, ks):
    """The device's max |L| and rho against numpy's unblocked LDL^T in the same
    (centre-out) order; cells with rho above the bound go to the LU, bit for bit."""
    ks = np.array(ks)
    med = G.Medium(d, n, G.HARDSPHERE, (alpha,)).fill(seed)
    _, st = G.tracelog_stats(med, ks, G.FP64_LDLT)
    lu, _ = G.tracelog_batch(med, ks, G.FP64)
    ldl, flags = G.tracelog_batch(med, ks, G.FP64_LDLT)
    pos = O.fill_medium(d, n, seed)
    p = centre_out(pos)
    for i, k in enumerate(ks):
        m = O.build_matrix(pos, k, O.HARDSPHERE, (alpha,))
        g, rho = ldlt_estimate(m[np.ix_(p, p)])
        print("k=%s  max|L| %.3e / %.3e  rho %.3e / %.3e  flag %d" % (k, st[i, 0], g, st[i, 1], rho, flags[i]))
        if rho < 1e6:                          # well below the bound: same values, kept
            assert st[i, 0] == pytest.approx(g, rel=1e-6)
            assert st[i, 1] == pytest.approx(rho, rel=1e-6)
            assert flags[i] == 0
        elif rho > 1e9:                        # far above: both say so, recomputed with the LU
            assert not (st[i, 1] <= 1e8)
            assert flags[i] == 1
    assert np.array_equal(ldl[flags == 1], lu[flags == 1])
    check(ldl, flags, lu, n)


def test_result_independent_of_input_order(G):
    """The LDL^T runs on the centre-out order, so a permutation of the input positions
    gives the same bits."""
    pos = O.fill_medium(3, 700, 9)
    perm = np.random.default_rng(1).permutation(700)
    a = G.Medium(3, 700, G.HARDSPHERE, (0.05,))
    a.set_pos(pos)
    b = G.Medium(3, 700, G.HARDSPHERE, (0.05,))
    b.set_pos(pos[perm])
    ks = kmap(8, 4, 4, 20, -1.5, 0)
    ra, fa = G.tracelog_batch(a, ks, G.FP64_LDLT)
    rb, fb = G.tracelog_batch(b, ks, G.FP64_LDLT)
    assert np.array_equal(fa, fb)
    keep = fa == 0
    assert keep.sum() > 0
    assert np.array_equal(ra[keep], rb[keep])


def _persistent_case(G):
    """A d=3 map at N = 300 (Np = 320, five 64-column panels) and an uneven multi-medium
    batch (d = 2 and 3, N from 90 to 450, all padded sizes below 512)."""
    med = G.Medium(3, 300, G.HARDSPHERE, (0.05,)).fill(5)
    ks = kmap(24, 12, 4.0, 24.0, -1.2, 0.0)
    meds = [G.Medium(3 if i % 2 else 2, n, G.HARDSPHERE, (0.1,)).fill(100 + i)
            for i, n in enumerate((150, 230, 310, 450, 90))]
    rng = np.random.default_rng(7)
    midx = rng.integers(0, len(meds), 160).astype(np.int32)
    mks = rng.uniform(2, 20, 160) - 1j * rng.uniform(0, 0.6, 160)
    a, fa = G.tracelog_batch(med, ks, G.FP64_LDLT)
    b, fb = G.tracelog_multi(meds, midx, mks, G.FP64_LDLT)
    return a, fa, b, fb


def test_persistent_kernel_equals_per_phase_launches(G, tmp_path):
    """MSGPU_LDLT_PERSIST_NP: one persistent launch, each workgroup taking items from a
    counter and running assembly, LDL^T and reduction alone, gives the per-phase launches'
    results bit for bit (in a process of its own: the variable is read once)."""
    import os
    import subprocess
    import sys
    from conftest import ROOT
    code = (
        "import sys, numpy as np\n"
        "sys.path.insert(0, %r); sys.path.insert(0, %r); sys.path.insert(0, %r)\n"
        "import test_gpu_ldlt as T, msmodel_amd as G\n"
        "G.init()\n"
        "a, fa, b, fb = T._persistent_case(G)\n"
        "np.savez(%r, a=a, fa=fa, b=b, fb=fb)\n"
        % (ROOT, os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle"), str(tmp_path / "p.npz")))
    env = dict(os.environ, MSGPU_LDLT_PERSIST_NP="512")
    print("running a python process with MSGPU_LDLT_PERSIST_NP=512", flush=True)
    subprocess.run([sys.executable, "-c", code], env=env, check=True, timeout=240)
    p = np.load(str(tmp_path / "p.npz"))
    a, fa, b, fb = _persistent_case(G)
    assert np.array_equal(p["a"], a) and np.array_equal(p["fa"], fa)
    assert np.array_equal(p["b"], b) and np.array_equal(p["fb"], fb)
    lu, _ = G.tracelog_batch(G.Medium(3, 300, G.HARDSPHERE, (0.05,)).fill(5), kmap(24, 12, 4.0, 24.0, -1.2, 0.0),
                             G.FP64)
    check(a, fa, lu, 300)
