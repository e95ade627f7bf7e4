"""Eigenstates at resonances (msgpu_eigenstates, the batched form of compute_eigenstates,
ms_core.c:296-313) against the per-root drop-in inverse_iteration
(ms_matrix_util.c:279-314) that the reference's loop calls, the eigenpair property with
the oracle's M(k), and the checks of check_eigenstate_list (eigenstate_list_util.c:130-163).
"""
import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu
HS = O.HARDSPHERE
P = (0.3,)


@pytest.fixture(scope="module")
def G():
    import msmodel_amd as G
    G.init()
    return G


@pytest.fixture(scope="module")
def found(G):
    med = G.Medium(2, 40, G.HARDSPHERE, P).fill(7)
    dom = (1.0, 12.0, -1.0, 0.0)
    rng = np.random.default_rng(7)
    starts = rng.uniform(1.5, 11.5, 10) + 1j * rng.uniform(-0.9, -0.05, 10)
    roots, _ = G.resonances(med, starts, dom, 4, 50, 1e-12, G.FP64)
    assert len(roots) >= 2
    return med, O.fill_medium(2, 40, 7), np.asarray(roots)


def test_eigenpairs_at_roots(G, found):
    med, pos, roots = found
    vecs, mu, conv, suspect = G.eigenstates(med, roots, 50, 1e-10)
    assert conv.all() and not suspect.any()
    for k, v, m in zip(roots, vecs, mu):
        M = O.build_matrix(pos, k, HS, P)
        assert abs(np.linalg.norm(v) - 1.0) < 1e-12
        assert abs(m) < 1e-6                                  # a resonance: M(k) is nearly singular
        assert np.linalg.norm(M @ v - m * v) <= 1e-9 * np.abs(M).max()


def test_equals_per_root_inverse_iteration(G, found):
    med, _, roots = found
    vecs, mu, conv, _ = G.eigenstates(med, roots, 50, 1e-10)
    v0 = np.full(med.n, 1.0 / np.sqrt(med.n), dtype=np.complex128)
    for k, v, m, c in zip(roots, vecs, mu, conv):
        c2, v2, lam2 = G.inverse_iteration(G.build_matrix(med, k), v0, 50, 1e-10)
        assert bool(c2) == bool(c)
        assert np.linalg.norm(v - v2) <= 1e-8
        assert abs(m - lam2) <= 1e-6 * abs(lam2) + 1e-15


def test_suspect_states(G, found):
    med, _, roots = found
    ks = np.array([roots[0], roots[0] * (1 + 1e-9), 5.5 - 0.5j])     # a duplicate and a non-root
    _, mu, _, suspect = G.eigenstates(med, ks, 50, 1e-10)
    assert not suspect[0]
    assert suspect[1]                                         # within 1e-6 |k| of an earlier root
    assert suspect[2] and abs(mu[2]) > 1e-6                   # M(k) is far from singular there
