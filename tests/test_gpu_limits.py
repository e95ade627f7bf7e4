"""The largest matrices of the device path (N = 8192: a panel holds at most 8192 rows)
and the size limit.

At N = 8192 the oracle's O(N^3) LU would take minutes, so the checks are
size-independent properties against O(N^2) oracle work:
  solve (complex128, and complex64 LU + FP64 refinement)
                      residual |M a - phi| / |phi| with the oracle's M <= 1e-10
  tr(M^-1 M')         equals the central difference of N tracelog(k) at k +- 1e-6
                      (two pipelines: LU alone, and LU + N-column solve), relative 1e-6;
                      arg det turns by ~4e4 rad per unit k here, so the step keeps the
                      turn between k - h and k + h far below pi (the difference is taken
                      mod 2 pi) and the truncation error small
  N = 8193            rejected with an error by the batched and the serving entry points
"""
import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu
N = 8192
ALPHA = (0.05,)


@pytest.fixture(scope="module")
def G():
    import msmodel_amd as G
    G.init()
    return G


@pytest.fixture(scope="module")
def big(G):
    med = G.Medium(3, N, G.HARDSPHERE, ALPHA).fill(5)
    pos = O.fill_medium(3, N, 5)
    assert np.array_equal(med.pos, pos)
    return med, pos


@pytest.mark.parametrize("mode", ["FP64", "MIXED"])
def test_solve_residual_at_max_size(G, big, mode):
    med, pos = big
    k = 6.0
    a, _ = G.solve(med, k, G.PLANE, getattr(G, mode), 1e-12)
    assert np.isfinite(a).all()
    m = O.build_matrix(pos, k, O.HARDSPHERE, ALPHA)
    phi = np.exp(1j * k * pos[:, 0])
    assert np.linalg.norm(m @ a - phi) <= 1e-10 * np.linalg.norm(phi)


def test_dk_equals_difference_of_tracelog_at_max_size(G, big):
    med, _ = big
    k, h = 5.0 - 0.2j, 1e-6
    tl, _ = G.tracelog_batch(med, np.array([k + h, k - h]), G.FP64)
    dk = G.dk_log_det_batch(med, np.array([k]), G.FP64)[0]
    d = N * (tl[0] - tl[1])
    d = complex(d.real, (d.imag + np.pi) % (2 * np.pi) - np.pi)    # arg det is defined mod 2 pi
    fd = d / (2 * h)
    assert np.isfinite(dk)
    assert abs(fd - dk) <= 1e-6 * abs(dk), (fd, dk)


def test_size_limit_is_an_error(G):
    med = G.Medium(3, N + 1, G.HARDSPHERE, ALPHA).fill(1)
    with pytest.raises(G.MsgpuError):
        G.tracelog_batch(med, np.array([5.0 - 0.1j]), G.FP64)
    srv = G.Server()
    with pytest.raises(G.MsgpuError):
        srv.submit_logdet(med, np.array([5.0 - 0.1j]))
    srv.close()
