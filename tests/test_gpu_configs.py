"""Parity at the sizes of the benchmark configurations.

The oracle finishes in seconds on samples of these workloads, so each test runs
the full GPU workload and checks a sample of its outputs against the oracle, plus
size-independent properties on the whole output:
  map N=200 (128x64 cells)          sample of cells vs oracle tracelog, 1e-10
  solve N=1000 + 1024^2 psi          residual |M a - phi| / |phi| with the oracle's M,
                                     FP32-LU + refinement vs FP64 1e-10, psi samples
                                     vs the oracle sum 1e-10
  map N=500, complex64 + FP64 re-evaluation
                                     re-evaluated cells equal the FP64 map, the others
                                     within the complex64 error; sample vs oracle
  uneven resonance search (d=2)      every root is a zero of det M: one oracle Newton
                                     step |1/(f'/f)| < 1e-10 |k|
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


def kgrid(nx, ny, xmin, xmax, ymin, ymax):
    dx, dy = (xmax - xmin) / (nx - 1), (ymax - ymin) / (ny - 1)
    l = np.arange(nx * ny)
    return (xmin + (l % nx) * dx) + 1j * (ymax - (l // nx) * dy)


def scaled(d, n, seed, radius):
    pos = O.fill_medium(d, n, seed)
    return pos * (radius / O.radius_of_ball(d, n))


def test_map_n200(G):
    med = G.Medium(3, 200, G.HARDSPHERE, (0.1,)).fill(1, radius=1.0)
    pos = scaled(3, 200, 1, 1.0)
    assert np.array_equal(med.pos, pos)
    ks = kgrid(128, 64, 1.0, 20.0, -1.0, 0.0)
    out, _ = G.tracelog_batch(med, ks, G.FP64)
    assert np.isfinite(out).all()
    for l in np.random.default_rng(1).choice(ks.size, 48, replace=False):
        want = O.tracelog(O.build_matrix(pos, ks[l], HS, (0.1,)))
        assert abs(out[l].real - want.real) <= 1e-10 * max(1.0, abs(want.real)), l
        assert abs(wrapped(200 * (out[l].imag - want.imag))) <= 1e-10, l


def test_solve_and_psi_n1000(G):
    med = G.Medium(3, 1000, G.HARDSPHERE, (0.1,)).fill(2, radius=5.0)
    pos = scaled(3, 1000, 2, 5.0)
    a64, _ = G.solve(med, 10.0, G.PLANE, G.FP64)
    a32, steps = G.solve(med, 10.0, G.PLANE, G.MIXED, 1e-12)
    m = O.build_matrix(pos, 10.0, HS, (0.1,))
    phi = np.exp(1j * 10.0 * pos[:, 0])
    for a in (a64, a32):
        assert np.linalg.norm(m @ a - phi) <= 1e-11 * np.linalg.norm(phi)
    assert np.abs(a32 - a64).max() <= 1e-10 * np.abs(a64).max()
    assert steps >= 1
    psi = G.wavefunction_grid(med, 10.0, a64, 1024, 1024, -6.0, 6.0, -6.0, 6.0)
    assert np.isfinite(psi).all()
    rng = np.random.default_rng(2)
    dx = 12.0 / 1023
    for _ in range(40):
        ix, iy = rng.integers(0, 1024, 2)
        x = np.array([-6.0 + ix * dx, 6.0 - iy * dx, 0.0])
        want = O.total_wave(pos, a64, 10.0, x)
        assert abs(psi[iy, ix] - want) <= 1e-10 * max(1.0, abs(want)), (ix, iy)


def test_map_n500_mixed(G):
    med = G.Medium(3, 500, G.HARDSPHERE, (0.05,)).fill(3, radius=1.0)
    pos = scaled(3, 500, 3, 1.0)
    ks = kgrid(64, 64, 5.0, 30.0, -2.0, 0.0)
    f64, _ = G.tracelog_batch(med, ks, G.FP64)
    mixed, flags = G.tracelog_batch(med, ks, G.MIXED, 1e-6)
    f32, _ = G.tracelog_batch(med, ks, G.FP32)
    re = flags == 1
    assert re.sum() > 0
    assert np.array_equal(mixed[re], f64[re])                 # re-evaluated cells are the FP64 ones
    assert np.array_equal(mixed[~re], f32[~re])               # the others keep the complex64 value
    err = np.abs(mixed.real - f64.real)
    assert err.max() <= 1e-3                                  # complex64 error on this medium
    for l in np.random.default_rng(3).choice(ks.size, 12, replace=False):
        want = O.tracelog(O.build_matrix(pos, ks[l], HS, (0.05,)))
        assert abs(f64[l].real - want.real) <= 1e-10 * max(1.0, abs(want.real)), l


def uneven_media(G, nmed=64, per=4):
    """The uneven-batch resonance search: d=2, nmed media (seeds 100..) with N log-uniform in
    [200, 2000] in unit disks, alpha=0.1, per starts each uniform in Re k in [1,40] x Im k in
    [-1,0] (the configuration's 10000 starts are the bench's; the test takes 4 per medium)."""
    rng = np.random.default_rng(100)
    meds, poss, starts, doms = [], [], [], []
    for s in range(100, 100 + nmed):
        n = int(round(np.exp(rng.uniform(np.log(200), np.log(2000)))))
        meds.append(G.Medium(2, n, G.HARDSPHERE, (0.1,)).fill(s, radius=1.0))
        poss.append(meds[-1].pos)
        starts.append(rng.uniform(1.0, 40.0, per) + 1j * rng.uniform(-1.0, 0.0, per))
        doms.append((1.0, 40.0, -1.0, 0.0))
    return meds, poss, starts, doms


def test_uneven_resonance_search(G):
    """64 media of N log-uniform in [200, 2000], 4 starts each, in one call: the mixed search
    (complex64 sweeps, complex128 polish) finds the same root set as the complex128 search
    in every medium (1e-10 |k|), and every root is a zero of det M: a complex128 Newton step
    on log det stays below 1e-10 |k| (the oracle's step for the media up to N = 260)."""
    meds, poss, starts, doms = uneven_media(G)
    print("media N:", sorted(m.n for m in meds), flush=True)
    r64, sw64 = G.resonances_multi(meds, starts, doms, 4, 40, 1e-12, G.FP64)
    r32, sw32 = G.resonances_multi(meds, starts, doms, 4, 40, 1e-12, G.MIXED)
    assert sw64 > 0 and sw32 > 0
    nroots, worst, nsmall = 0, 0.0, 0
    for m, pos, a, b in zip(meds, poss, r64, r32):
        assert len(a) == len(b), (m.n, np.sort_complex(a), np.sort_complex(b))
        for z in a:
            assert np.min(np.abs(b - z)) <= 1e-10 * abs(z), (m.n, z)
        for z in b:
            assert np.min(np.abs(a - z)) <= 1e-10 * abs(z), (m.n, z)
        if len(a):
            steps = 1.0 / G.dk_log_det_batch(m, a, G.FP64)
            worst = max(worst, float(np.max(np.abs(steps) / np.abs(a))))
            assert (np.abs(steps) <= 1e-10 * np.abs(a)).all(), m.n
            if m.n <= 260:
                for z in a:
                    assert abs(1.0 / O.dk_log_det(pos, z, HS, (0.1,))) <= 1e-10 * abs(z), (m.n, z)
                    nsmall += 1
        nroots += len(a)
    print("%d roots in 64 media (%d checked by the oracle), sweeps FP64 %d / mixed %d, worst Newton step %.1e |k|"
          % (nroots, nsmall, sw64, sw32, worst))
    assert nroots >= 10            # 19 with these starts (4 per medium over Re k in [1, 40])

def test_mixed_gated_map(G):
    """The pivot-growth gate (MSGPU_MIXED_GATED) on the N=500 map: the flagged cells are the
    ones whose complex64 factor's bound 16 eps32 exp(max log|U_ii| - min log|U_ii|) exceeds
    tol; they are the complex128 values and the others the complex64 ones; every kept cell
    is within tol of complex128; at tol = 1e-8 every cell is recomputed."""
    med = G.Medium(3, 500, G.HARDSPHERE, (0.05,)).fill(3, radius=1.0)
    ks = kgrid(24, 24, 5.0, 30.0, -2.0, 0.0)
    f64, _ = G.tracelog_batch(med, ks, G.FP64)
    f32, st32 = G.tracelog_stats(med, ks, G.FP32)
    bound = 16 * 2.0 ** -24 * np.exp(st32[:, 0] - st32[:, 1])
    for tol in (1e-8, 3e-6, 1e-4):
        out, flags = G.tracelog_batch(med, ks, G.MIXED_GATED, tol)
        want_flags = ~(bound <= tol)
        assert np.array_equal(flags.astype(bool), want_flags), tol
        assert np.array_equal(out[want_flags], f64[want_flags])
        assert np.array_equal(out[~want_flags], f32[~want_flags])
        err = np.hypot(out.real - f64.real, wrapped(500 * (out.imag - f64.imag)) / 500)
        print("tol %.0e: %d of %d cells recomputed, max error of the kept cells %.2e"
              % (tol, want_flags.sum(), ks.size, err[~want_flags].max() if (~want_flags).any() else 0.0))
        assert (err <= tol).all(), (tol, err.max())
        if tol == 1e-8:
            assert want_flags.all()
    # the serving loop gives the same tile
    srv = G.Server()
    got = srv.wait(srv.submit_logdet(med, ks[:40], G.MIXED_GATED, 1e-4))[0]["tracelog"]
    srv.close()
    assert np.array_equal(got, G.tracelog_batch(med, ks[:40], G.MIXED_GATED, 1e-4)[0])
