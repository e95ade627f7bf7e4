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


def test_uneven_resonance_search(G):
    rng = np.random.default_rng(100)
    meds, poss, starts, doms = [], [], [], []
    for s in range(100, 104):
        n = int(round(np.exp(rng.uniform(np.log(200), np.log(400)))))
        meds.append(G.Medium(2, n, G.HARDSPHERE, (0.1,)).fill(s, radius=1.0))
        poss.append(scaled(2, n, s, 1.0))
        starts.append(rng.uniform(1.0, 40.0, 4) + 1j * rng.uniform(-1.0, 0.0, 4))
        doms.append((1.0, 40.0, -1.0, 0.0))
    roots, sweeps = G.resonances_multi(meds, starts, doms, 4, 40, 1e-12, G.MIXED)
    assert sweeps > 0
    found = 0
    for pos, r in zip(poss, roots):
        for z in r:
            step = 1.0 / O.dk_log_det(pos, z, HS, (0.1,))
            assert abs(step) <= 1e-10 * abs(z), (z, step)
            found += 1
    assert found >= 1
