"""Root sets and eigenstates of the HIP path against those of the compiled reference
(tests/golden/ref_fixtures.npz, written by the reference's own drivers):

  Aberth on det M    msgpu_resonances vs solve_aberth + dk_log_det_ms (root_finder.c:522-577,
                     ms_core.c:320-328), d=2 N=20 and d=3 N=30, FP64 and MIXED
  damped Muller      msgpu_roots_muller vs find_eigenstates_serial_2/_3 (ms_core.c:335-401,
                     root_finder.c:342-407) on three media: 1/tr M^-1 against the reference's
                     root sets; the smallest eigenvalue against the oracle's per-start form
                     (the batched entry point gives each start its own inverse-iteration
                     vector where the reference shares one across its sequential loop, and the
                     smallest-eigenvalue function depends on that vector's history), every
                     root also a zero of det M
  eigenstates        msgpu_eigenstates vs compute_eigenstates (ms_core.c:296-313): mu and the
                     phase-normalised eigenvector that save_eigenstate writes

Root sets are compared in both directions (assert_cvector_subset, src/test/assertion.c:136-162):
every reference root has a GPU root within 1e-10 |k| and every GPU root a reference root.
"""
import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu
HS = O.HARDSPHERE


@pytest.fixture(scope="module")
def G():
    import msmodel_amd as G
    G.init()
    return G


def medium(G, pos, p):
    pos = np.asarray(pos)
    n, d = pos.shape
    m = G.Medium(d, n, HS, p)
    m.set_pos(pos)
    return m


def same_root_sets(got, want, rtol=1e-10):
    got, want = np.asarray(got), np.asarray(want)
    assert got.size == want.size and want.size > 0, (np.sort_complex(got), np.sort_complex(want))
    err = max(max(np.min(np.abs(got - w)) / abs(w) for w in want), max(np.min(np.abs(want - g)) / abs(g) for g in got))
    print("%d roots, max relative distance to the reference set %.2e" % (want.size, err))
    assert err < rtol, err


def eig_case(ref, tag):
    spec = ref["eig_%s_spec" % tag]
    return dict(pos=ref["eig_%s_pos" % tag], alpha=spec[3], dom=tuple(spec[4:8]), nt=int(spec[8]),
                maxit=int(spec[9]), toler=spec[10], method=int(spec[11]), kroot=ref["eig_%s_kroot" % tag],
                mu=ref["eig_%s_mu" % tag], eigv=ref["eig_%s_eigv" % tag])


def halton_starts(nt, dom):
    """Starts of the [eigenstate_list] driver: Halton (2,3) points sorted by -Im k, stable
    (run_eigenstate_list_serial, ms_core.c:413-415)."""
    z = O.halton(nt, *dom)
    return z[np.argsort(-z.imag, kind="stable")]


@pytest.mark.parametrize("mode", ["FP64", "MIXED"])
@pytest.mark.parametrize("case", ["d2", "d3"])
def test_aberth_root_sets(G, ref, case, mode):
    pos = O.fill_medium(2, 20, 1) if case == "d2" else ref["ab_d3_pos"]
    m = medium(G, pos, (0.4,) if case == "d2" else (0.1,))
    roots, sweeps = G.resonances(m, ref["ab_%s_starts" % case], tuple(ref["ab_%s_dom" % case]), 4, 60, 1e-12,
                                 getattr(G, mode))
    assert sweeps > 0
    same_root_sets(roots, ref["ab_%s_roots" % case])


@pytest.mark.parametrize("tag", ["iti_d2", "iti_d3", "iti_d2n60"])
def test_muller_invtraceinv_root_sets(G, ref, tag):
    c = eig_case(ref, tag)
    m = medium(G, c["pos"], (c["alpha"],))
    roots, rounds = G.roots_muller(m, halton_starts(c["nt"], c["dom"]), c["dom"], c["maxit"], c["toler"],
                                   G.ROOT_INVTRACEINV)
    assert rounds > 0
    same_root_sets(roots, c["kroot"])


@pytest.mark.parametrize("tag", ["mev_d2", "mev_d3", "mev_d2n60"])
def test_muller_mineigval_root_sets(G, ref, tag):
    c = eig_case(ref, tag)
    m = medium(G, c["pos"], (c["alpha"],))
    starts = halton_starts(c["nt"], c["dom"])
    roots, rounds = G.roots_muller(m, starts, c["dom"], c["maxit"], c["toler"], G.ROOT_MINEIGVAL)
    assert rounds > 0
    want = O.roots_muller(c["pos"], starts, c["dom"], O.MINEIGVAL_PER_START, HS, (c["alpha"],), c["maxit"],
                          c["toler"])
    same_root_sets(roots, want)
    for z in roots:        # a zero of det M: one oracle Newton step on log det stays put
        assert abs(1.0 / O.dk_log_det(c["pos"], z, HS, (c["alpha"],))) < 1e-9 * abs(z), z
    common = sum(np.min(np.abs(c["kroot"] - z)) < 1e-10 * abs(z) for z in roots)
    print("shared-vector reference: %d roots, %d of them found here (%d roots)" % (c["kroot"].size, common,
                                                                                 roots.size))


@pytest.mark.parametrize("tag", ["det_d2", "mev_d3", "iti_d2n60"])
def test_eigenstates_vs_reference(G, ref, tag):
    c = eig_case(ref, tag)
    m = medium(G, c["pos"], (c["alpha"],))
    vecs, mu, conv, suspect = G.eigenstates(m, c["kroot"], 50, 1e-10)
    dmu = np.abs(mu - c["mu"]).max()
    dv = np.abs(vecs - c["eigv"]).max()
    print("%d states: max |d mu| %.2e, max |d v| %.2e" % (len(mu), dmu, dv))
    assert dmu < 1e-12
    assert dv < 1e-10
    # check_eigenstate_list flags states within 1e-6 |k| of an earlier root or with |mu| > 1e-6
    for i, k in enumerate(c["kroot"]):
        dup = any(abs(k - c["kroot"][j]) < 1e-6 * abs(k) for j in range(i))
        assert bool(suspect[i]) == (dup or abs(c["mu"][i]) > 1e-6), i


def test_kgrids_bit_exact(G, ref):
    """msgpu_kgrid_map/_cut/_logcut against the reference's get_arg_cmap, get_arg_ccut and
    init_logcut + get_arg_lcut (complex_map_util.c:27-31, complex_cut_util.c:32-36,
    log_cut_util.c:72-101), bit for bit."""
    for i, g in enumerate(ref["grid_map_spec"]):
        assert np.array_equal(G.kgrid_map(int(g[0]), int(g[1]), *g[2:]), ref["grid_map_%d" % i]), i
    for i, (ns, ends) in enumerate(zip(ref["grid_cut_ns"], ref["grid_cut_ends"])):
        assert np.array_equal(G.kgrid_cut(int(ns), ends[0], ends[1]), ref["grid_cut_%d" % i]), i
    for i, g in enumerate(ref["grid_log_spec"]):
        assert np.array_equal(G.kgrid_logcut(int(g[0]), *g[1:]), ref["grid_log_%d" % i]), i
    with pytest.raises(G.MsgpuError):
        G.kgrid_logcut(5, 3.0, 0.1, -1.0)
