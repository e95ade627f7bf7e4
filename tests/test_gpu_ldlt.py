"""MSGPU_FP64_LDLT: the trace-log map with the pivot-free LDL^T on the cells whose growth
is predicted small, the pivoted LU elsewhere.

M(k) is complex symmetric (ms_matrix_util.c:36-41), so log det M = sum_i log D_ii of
M = L D L^T when that factorisation exists; the reference computes log det from zgetrf_
(ms_matrix_util.c:185-202).  Each cell is checked against the pivoted complex128 LU map
(which the parity tests pin to the reference), at 1e-8 in N tracelog = log det (real
part; the imaginary part modulo 2 pi), and at small N against the oracle; cells the
mode computes with the LU must equal the LU map bit for bit.
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
    assert (flags == 0).sum() >= ks.size // 3          # a good part of the map avoids the LU
    assert flags[ks.imag > -0.5].sum() == 0            # near the real axis the growth stays small


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


def test_growth_rejection_falls_back(G):
    """Cells far below the real axis (growth far beyond the bound) all go to the LU."""
    med = G.Medium(3, 400, G.HARDSPHERE, (0.05,)).fill(2)
    ks = np.array([6 - 3j, 12 - 2.5j, 20 - 3.5j])
    lu, _ = G.tracelog_batch(med, ks, G.FP64)
    ldl, flags = G.tracelog_batch(med, ks, G.FP64_LDLT)
    assert flags.tolist() == [1, 1, 1]
    assert np.array_equal(ldl, lu)


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
