"""Batches over several media (msgpu_*_multi): media of different N, d and
scattering model in one call.  Each item's result must be the one the
single-medium entry point gives for it (the kernels treat every matrix on its
own, so the comparison is exact), and the single-medium results are pinned to
the oracle and the reference fixtures elsewhere; a few items are also checked
against the oracle directly here.
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


def medium(G, pos, model, p):
    pos = np.asarray(pos)
    n, d = pos.shape
    m = G.Medium(d, n, model, p)
    m.set_pos(pos)
    return m


SPECS = [  # d, N, seed, model, params
    (3, 70, 3, O.HARDSPHERE, (0.1,)),
    (2, 130, 4, O.HARDSPHERE, (0.2,)),
    (3, 200, 5, O.SOFTSPHERE, (0.3,)),
    (4, 60, 6, O.HARDSPHERE, (0.1,)),
    (1, 45, 7, O.RESONANT, (2.0, 0.5)),
    (3, 64, 8, O.MAXIMUM, ()),
]


@pytest.fixture(scope="module")
def media(G):
    out = []
    for d, n, seed, model, p in SPECS:
        pos = O.fill_medium(d, n, seed)
        out.append((medium(G, pos, model, p), pos, model, p))
    return out


def items(nmed, per, seed):
    rng = np.random.default_rng(seed)
    med_of = np.repeat(np.arange(nmed), per)
    rng.shuffle(med_of)
    ks = rng.uniform(2.0, 9.0, med_of.size) + 1j * rng.uniform(-0.8, 0.0, med_of.size)
    return med_of.astype(np.int32), ks


@pytest.mark.parametrize("mode", ["FP64", "FP32", "MIXED"])
def test_tracelog_multi_equals_single(G, media, mode):
    meds = [m[0] for m in media]
    med_of, ks = items(len(meds), 7, 1)
    md = getattr(G, mode)
    got, flags = G.tracelog_multi(meds, med_of, ks, md, 1e-3)
    for j, m in enumerate(meds):
        sel = med_of == j
        want, wflags = G.tracelog_batch(m, ks[sel], md, 1e-3)
        assert np.array_equal(got[sel], want), (j, np.abs(got[sel] - want).max())
        assert np.array_equal(flags[sel], wflags)


def test_tracelog_multi_vs_oracle(G, media):
    meds = [m[0] for m in media]
    med_of, ks = items(len(meds), 3, 2)
    got, _ = G.tracelog_multi(meds, med_of, ks, G.FP64)
    for i in range(ks.size):
        _, pos, model, p = media[med_of[i]]
        want = O.tracelog(O.build_matrix(pos, ks[i], model, p))
        n = pos.shape[0]
        assert abs(got[i].real - want.real) <= 1e-10 * max(1.0, abs(want.real)), i
        assert abs(wrapped(n * (got[i].imag - want.imag))) <= 1e-10, i


def test_dk_multi_equals_single_and_oracle(G, media):
    meds = [m[0] for m in media]
    med_of, ks = items(len(meds), 3, 3)
    got = G.dk_log_det_multi(meds, med_of, ks, G.FP64)
    for j, m in enumerate(meds):
        sel = med_of == j
        assert np.array_equal(got[sel], G.dk_log_det_batch(m, ks[sel], G.FP64)), j
    for i in range(0, ks.size, 2):
        _, pos, model, p = media[med_of[i]]
        want = O.dk_log_det(pos, ks[i], model, p)
        assert abs(got[i] - want) <= 1e-9 * abs(want), (i, got[i], want)


# padded sizes 896, 832, 768, 704 (no tuning-table entries: one storage order and blocking),
# a few items each: the planner factors the small groups together, padded to the largest
MERGE_SPECS = [(3, 890, 11), (2, 820, 12), (3, 760, 13), (2, 700, 14)]


@pytest.fixture(scope="module")
def merge_media(G):
    return [G.Medium(d, n, G.HARDSPHERE, (0.1,)).fill(seed) for d, n, seed in MERGE_SPECS]


@pytest.mark.parametrize("mode", ["FP64", "FP32", "MIXED"])
def test_small_groups_padded_together_equal_single(G, merge_media, mode):
    """Results of media padded to a larger size in a shared batch equal each medium's own
    call bit for bit (trace-log, and tr(M^-1 M') where its solves allow the padding)."""
    md = getattr(G, mode)
    med_of, ks = items(len(merge_media), 3, 21)
    got, flags = G.tracelog_multi(merge_media, med_of, ks, md, 1e-3)
    for j, m in enumerate(merge_media):
        sel = med_of == j
        want, wflags = G.tracelog_batch(m, ks[sel], md, 1e-3)
        assert np.array_equal(got[sel], want), (j, np.abs(got[sel] - want).max())
        assert np.array_equal(flags[sel], wflags)
    if mode == "MIXED":
        return
    dk = G.dk_log_det_multi(merge_media, med_of, ks, md)
    for j, m in enumerate(merge_media):
        sel = med_of == j
        want = G.dk_log_det_batch(m, ks[sel], md)
        assert np.array_equal(dk[sel], want), (j, np.abs(dk[sel] - want).max())


def test_multi_validates_indices(G, media):
    with pytest.raises(G.MsgpuError):
        G.tracelog_multi([media[0][0]], np.array([0, 1], np.int32), np.array([3.0 + 0j, 4.0 + 0j]))


def test_resonances_multi_equals_single(G, ref):
    hs = O.HARDSPHERE
    a = medium(G, O.fill_medium(2, 20, 1), hs, (0.4,))
    b = medium(G, ref["ab_d3_pos"], hs, (0.1,))
    c = medium(G, O.fill_medium(2, 70, 5), hs, (0.4,))
    dom_c = (2.0, 8.0, -0.8, 0.0)
    rng = np.random.default_rng(4)
    starts_c = dom_c[0] + (dom_c[1] - dom_c[0]) * rng.uniform(0.05, 0.95, 9) + \
        1j * (dom_c[2] + (dom_c[3] - dom_c[2]) * rng.uniform(0.05, 0.95, 9))
    meds = [a, b, c]
    starts = [ref["ab_d2_starts"], ref["ab_d3_starts"], starts_c]
    doms = [tuple(ref["ab_d2_dom"]), tuple(ref["ab_d3_dom"]), dom_c]
    for mode in (G.FP64, G.MIXED):
        roots, sweeps = G.resonances_multi(meds, starts, doms, 4, 60, 1e-12, mode)
        assert sweeps > 0
        for m, s, dm, r in zip(meds, starts, doms, roots):
            want, _ = G.resonances(m, s, dm, 4, 60, 1e-12, mode)
            assert len(r) == len(want)
            assert np.allclose(np.sort_complex(r), np.sort_complex(want), rtol=1e-12, atol=0)
        # media a and b give the reference's root sets, in both directions
        for r, want in ((roots[0], ref["ab_d2_roots"]), (roots[1], ref["ab_d3_roots"])):
            assert len(r) == len(want), (mode, r, want)
            for w in want:
                assert np.min(np.abs(r - w)) < 1e-10 * abs(w), (mode, w)
