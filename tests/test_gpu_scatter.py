"""Cross sections and density of states of many (medium, real k) items in one batch
(msgpu_scattering_multi) against the oracle: the total cross section
(total_cross_section, ms_matrix_util.c:344-366), the density of states
-Im tr(M^-1 M')/(2 pi k) (ms_core.c:503-505) and the differential cross section
(ms_core.c:526-587) with the oracle's amplitudes and plane waves.
"""
import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu
SPECS = [(3, 60, 1, (0.2,)), (2, 90, 2, (0.3,)), (1, 30, 3, (0.5,)), (4, 40, 4, (0.1,))]
KS = np.array([0.8, 2.5, 6.0])


@pytest.fixture(scope="module")
def G():
    import msmodel_amd as G
    G.init()
    return G


@pytest.fixture(scope="module")
def media(G):
    out = []
    for d, n, seed, p in SPECS:
        med = G.Medium(d, n, G.HARDSPHERE, p).fill(seed)
        out.append((med, O.fill_medium(d, n, seed), p))
    return out


def items(media):
    med_of = np.repeat(np.arange(len(media)), KS.size).astype(np.int32)
    ks = np.tile(KS, len(media))
    return med_of, ks


def test_total_cross_section(G, media):
    med_of, ks = items(media)
    got = G.scattering_multi([m[0] for m in media], med_of, ks, G.CRSEC)
    for i, (j, k) in enumerate(zip(med_of, ks)):
        _, pos, p = media[j]
        want = O.total_cross_section(pos, k, O.HARDSPHERE, p)
        assert got[i] > 0
        assert abs(got[i] - want) <= 1e-10 * abs(want), (j, k, got[i], want)


def test_density_of_states(G, media):
    med_of, ks = items(media)
    got = G.scattering_multi([m[0] for m in media], med_of, ks, G.DOS)
    for i, (j, k) in enumerate(zip(med_of, ks)):
        _, pos, p = media[j]
        want = -O.dk_log_det(pos, k, O.HARDSPHERE, p).imag / (2 * np.pi * k)
        assert abs(got[i] - want) <= 1e-8 * max(1.0, abs(want)), (j, k, got[i], want)


def test_differential_cross_section(G, media):
    med_of, ks = items(media)
    thetas = np.array([0.0, 10.0, 45.0, 90.0, 170.0])
    got = G.scattering_multi([m[0] for m in media], med_of, ks, G.DCS, thetas)
    for i, (j, k) in enumerate(zip(med_of, ks)):
        _, pos, p = media[j]
        d = pos.shape[1]
        a, info = O.solve(pos, k, O.HARDSPHERE, p, O.PLANE)
        assert info == 0
        prefac = (k / (2 * np.pi)) ** (d - 3) / (16 * np.pi ** 2)
        for t, th in enumerate(thetas):
            b = O.plane_wave(pos, k, th)
            want = prefac * abs(np.vdot(b, a)) ** 2
            assert abs(got[i, t] - want) <= 1e-10 * max(want, 1e-300) + 1e-300, (j, k, th, got[i, t], want)
