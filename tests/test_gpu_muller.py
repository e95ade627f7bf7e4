"""Resonances by the damped Muller method on 1/tr M(k)^-1 and on the smallest eigenvalue
of M(k) (msgpu_roots_muller; find_eigenstates_serial_2/_3, ms_core.c:335-401): every
returned root is a zero of det M (one oracle Newton step on log det is below 1e-9 |k|),
lies in the domain, and the roots are distinct; the two root functions find common roots
to 1e-9.
"""
import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu
P = (0.3,)
DOM = (1.0, 12.0, -1.0, 0.0)


@pytest.fixture(scope="module")
def G():
    import msmodel_amd as G
    G.init()
    return G


@pytest.fixture(scope="module")
def setup(G):
    med = G.Medium(2, 30, G.HARDSPHERE, P).fill(3)
    pos = O.fill_medium(2, 30, 3)
    rng = np.random.default_rng(3)
    starts = rng.uniform(1.5, 11.5, 8) + 1j * rng.uniform(-0.8, -0.05, 8)
    return med, pos, starts


def check_roots(pos, roots):
    assert len(roots) >= 1
    for z in roots:
        assert DOM[0] < z.real < DOM[1] and DOM[2] < z.imag < DOM[3]
        step = 1.0 / O.dk_log_det(pos, z, O.HARDSPHERE, P)
        assert abs(step) <= 1e-9 * abs(z), (z, step)
    for i in range(len(roots)):
        for j in range(i):
            assert abs(roots[i] - roots[j]) >= 1e-14 * abs(roots[i])


@pytest.mark.parametrize("method", ["ROOT_INVTRACEINV", "ROOT_MINEIGVAL"])
def test_muller_roots_are_zeros_of_det(G, setup, method):
    med, pos, starts = setup
    roots, rounds = G.roots_muller(med, starts, DOM, 30, 1e-12, getattr(G, method))
    assert rounds > 0
    check_roots(pos, roots)


def test_both_root_functions_agree(G, setup):
    med, _, starts = setup
    r1, _ = G.roots_muller(med, starts, DOM, 30, 1e-12, G.ROOT_INVTRACEINV)
    r2, _ = G.roots_muller(med, starts, DOM, 30, 1e-12, G.ROOT_MINEIGVAL)
    common = [z for z in r1 if np.min(np.abs(r2 - z)) <= 1e-9 * abs(z)] if len(r2) else []
    assert len(common) >= 1
