"""Edge cases of the device path against the oracle: LAPACK pivot semantics (izamax order,
ties, exact zero pivots and their info), the smallest media, padding boundaries and
empty batches.
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


def test_pivot_rule_small(G):
    """Largest |Re| + |Im|; the first row on ties (zgetrf_, 2 x 2 as the oracle's test)."""
    for a in (np.array([[1, 2], [1 + 1j, 2j]], dtype=complex), np.array([[2j, 1], [2, 3]], dtype=complex)):
        _, ipiv, info = G.zgetrf(a)
        _, oipiv, _ = O.getrf(a)
        assert info == 0 and np.array_equal(ipiv, oipiv)


def test_pivot_tie_in_a_large_matrix(G):
    """Two rows with the same |Re| + |Im| maximum in column 0 of a 300 x 300 matrix: the
    first wins, and every later pivot matches the oracle."""
    rng = np.random.default_rng(4)
    a = (rng.normal(size=(300, 300)) + 1j * rng.normal(size=(300, 300))) * 0.1
    a[7, 0], a[250, 0] = 10.0, 6.0 + 4.0j
    lu, ipiv, info = G.zgetrf(a)
    olu, oipiv, oinfo = O.getrf(a)
    assert info == oinfo == 0
    assert ipiv[0] == 8
    assert np.array_equal(ipiv, oipiv)
    assert np.abs(lu - olu).max() < 1e-11 * np.abs(olu).max()


@pytest.mark.parametrize("n,zero_col", [(2, 0), (70, 3), (200, 130)])
def test_exact_zero_pivot_sets_info(G, n, zero_col):
    """A zero column gives an exact zero pivot: info = its 1-based index (zgetrf_), and the
    trace-log's real part is -inf (tracelog_lu prints the LAPACK message and goes on)."""
    rng = np.random.default_rng(n)
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    a[:, zero_col] = 0.0
    _, _, info = G.zgetrf(a)
    _, _, oinfo = O.getrf(a)
    assert info == oinfo == zero_col + 1
    t, _ = G.tracelog_lu(a)
    assert np.isneginf(t.real)


@pytest.mark.parametrize("n", [2, 3, 63, 64, 65, 127, 128, 129])
def test_tracelog_sizes_around_padding(G, n):
    """The smallest media (a medium has at least 2 atoms: init_medium, medium_util.c:37-60)
    and sizes on both sides of the 64-row padding and the 128-row blocks, complex128 and
    complex64, against the oracle."""
    pos = O.fill_medium(3, n, 40 + n)
    med = G.Medium(3, n, G.HARDSPHERE, (0.2,))
    med.set_pos(pos)
    ks = np.array([0.7 - 0.05j, 3.0 + 0.0j, 8.5 - 0.9j])
    for mode, tol in ((G.FP64, 1e-10), (G.FP32, 1e-3)):
        out, _ = G.tracelog_batch(med, ks, mode)
        for k, t in zip(ks, out):
            want = O.tracelog(O.build_matrix(pos, k, HS, (0.2,)))
            assert abs(t.real - want.real) <= tol * max(1.0, abs(want.real)), (n, k, mode)
            if mode == G.FP64:
                assert abs(wrapped(n * (t.imag - want.imag))) <= 1e-10, (n, k)


def test_empty_batches(G):
    med = G.Medium(3, 50, G.HARDSPHERE, (0.1,)).fill(1)
    out, flags = G.tracelog_batch(med, np.zeros(0, np.complex128), G.FP64)
    assert out.size == 0 and flags.size == 0
    assert G.dk_log_det_batch(med, np.zeros(0, np.complex128), G.FP64).size == 0


@pytest.mark.parametrize("n,nan_col", [(40, 5), (300, 130)])
def test_nan_column_pivots_like_lapack(G, n, nan_col):
    """A NaN column: izamax (|Re| + |Im| compared with '>') keeps the first candidate, so
    from that column on every pivot is the diagonal row (ipiv[j] = j + 1) and info stays
    0; the pivots before it are those of the finite part."""
    rng = np.random.default_rng(n)
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    a[:, nan_col] = np.nan
    _, ipiv, info = G.zgetrf(a)
    _, oipiv, _ = O.getrf(a)
    assert info == 0
    assert np.array_equal(ipiv, oipiv)
    assert np.array_equal(ipiv[nan_col:], np.arange(nan_col + 1, n + 1))
    t, _ = G.tracelog_lu(a)
    assert np.isnan(t.real)


def test_overflowing_entries_leave_the_batch_intact(G):
    """At Im k = -800 the d = 3 Green function exp(ikr)/r overflows (exp(800 r)), so that
    cell's matrix holds infinities and its factorisation NaNs; the call returns and the
    other cells of the same batch are unaffected."""
    pos = O.fill_medium(3, 200, 9)
    med = G.Medium(3, 200, G.HARDSPHERE, (0.1,))
    med.set_pos(pos)
    ks = np.array([3.0 - 0.1j, 5.0 - 800.0j, 7.0 - 0.3j])
    for mode, tol in ((G.FP64, 1e-10), (G.MIXED, 1e-3), (G.FP32, 1e-3)):
        out, _ = G.tracelog_batch(med, ks, mode)
        assert not np.isfinite(out[1])
        for i in (0, 2):
            want = O.tracelog(O.build_matrix(pos, ks[i], HS, (0.1,)))
            assert abs(out[i].real - want.real) <= tol * max(1.0, abs(want.real)), (mode, ks[i])
