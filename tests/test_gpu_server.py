"""The serving loop (msgpu_server_*): mixed requests of different kinds, N and d.

Each request is served alone (a server of its own, nothing else in flight) and
then in a mix where all requests are in flight together under a small admission
budget, so that they are admitted in stages and share batched steps.  Every
result must be the same alone and in the mix; the comparison is exact because
every matrix is computed on its own.  The alone results are also checked
against the direct entry points (msgpu_tracelog_batch, msgpu_solve,
msgpu_wavefunction_grid, msgpu_resonances), which the parity tests pin to the
oracle.
"""
import threading
import time

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import msmodel_amd as G
    G.init()
    return G


def make_requests(G):
    """(kind, medium, args) for a reproducible mix."""
    rng = np.random.default_rng(12)
    reqs = []
    specs = [("logdet", 3, 90, 1), ("solve", 2, 150, 2), ("resonance", 2, 24, 3), ("logdet", 2, 260, 4),
             ("logdet", 3, 130, 5), ("solve", 3, 200, 6), ("resonance", 2, 40, 7), ("logdet", 3, 64, 8),
             ("solve", 2, 70, 9), ("logdet", 4, 50, 10)]
    for kind, d, n, seed in specs:
        med = G.Medium(d, n, G.HARDSPHERE, (0.3 if d == 2 else 0.1,)).fill(seed)
        if kind == "logdet":
            ks = rng.uniform(2, 10, 24) + 1j * rng.uniform(-1, 0, 24)
            mode = [G.FP64, G.MIXED, G.FP32][seed % 3]
            reqs.append((kind, med, dict(ks=ks, mode=mode, reeval_rel=1e-2)))
        elif kind == "solve":
            mode = G.MIXED if seed % 2 else G.FP64
            reqs.append((kind, med, dict(k=4.0 + 0.5 * seed, mode=mode, tol=1e-12,
                                         grid=(12, 9, -3.0, 3.0, -2.0, 2.0))))
        else:
            dom = (1.0, 8.0, -1.0, 0.0)
            st = dom[0] + (dom[1] - dom[0]) * rng.uniform(0.05, 0.95, 6) + \
                1j * (dom[2] + (dom[3] - dom[2]) * rng.uniform(0.05, 0.95, 6))
            reqs.append((kind, med, dict(starts=st, dom=dom, mode=G.MIXED if seed % 2 else G.FP64)))
    return reqs


def submit(srv, kind, med, a):
    if kind == "logdet":
        return srv.submit_logdet(med, a["ks"], a["mode"], a["reeval_rel"])
    if kind == "solve":
        return srv.submit_solve(med, a["k"], G_PLANE, a["mode"], a["tol"], a["grid"])
    return srv.submit_resonance(med, a["starts"], a["dom"], 4, 50, 1e-12, a["mode"])


G_PLANE = 0


def footprint_bound(kind, med, a):
    n2 = float(med.n) ** 2
    if kind == "logdet":
        return a["ks"].size * n2 * 16
    if kind == "solve":
        return n2 * 24
    return 2 * a["starts"].size * n2 * 16


def same(x, y):
    assert x.keys() == y.keys()
    for key in x:
        assert np.array_equal(x[key], y[key]), key


@pytest.fixture(scope="module")
def alone(G):
    reqs = make_requests(G)
    outs = []
    for kind, med, a in reqs:
        srv = G.Server(hbm_gb=64.0)
        print("alone", kind, med.c.d, med.n, flush=True)
        outs.append(srv.wait(submit(srv, kind, med, a))[0])
        srv.close()
    return reqs, outs


def test_alone_equals_direct_calls(G, alone):
    reqs, outs = alone
    for (kind, med, a), out in zip(reqs, outs):
        if kind == "logdet":
            want, _ = G.tracelog_batch(med, a["ks"], a["mode"], a["reeval_rel"])
            assert np.array_equal(out["tracelog"], want)
        elif kind == "solve":
            aw, _ = G.solve(med, a["k"], G.PLANE, a["mode"], a["tol"])
            assert np.array_equal(out["a"], aw)
            psi = G.wavefunction_grid(med, a["k"], aw, *a["grid"])
            assert np.array_equal(out["psi"], psi.ravel())
        else:
            rw, _ = G.resonances(med, a["starts"], a["dom"], 4, 50, 1e-12, a["mode"])
            assert np.array_equal(out["roots"], rw)


def test_mix_equals_alone(G, alone):
    reqs, outs = alone
    # a budget that holds only a few of these requests at once: staged admission
    srv = G.Server(hbm_gb=0.02, max_active=6)
    ids = [submit(srv, kind, med, a) for kind, med, a in reqs]
    got = [srv.wait(i) for i in ids]
    st = srv.stats()
    srv.close()
    for (out, info), want in zip(got, outs):
        same(out, want)
        assert info["total_ms"] >= info["queue_ms"] >= 0
    assert st["max_requests_per_step"] >= 2, st
    # admitted bytes stay within the budget, except for one request larger than it alone
    assert st["peak_admitted_bytes"] <= max(st["budget_bytes"], max(footprint_bound(*r) for r in reqs)), st


def test_arrivals_while_busy(G, alone):
    """Requests submitted from another thread while earlier ones are on the GPU."""
    reqs, outs = alone
    srv = G.Server(hbm_gb=8.0)
    ids = [None] * len(reqs)

    def late():
        for j in range(1, len(reqs), 2):
            time.sleep(0.002)
            ids[j] = submit(srv, *reqs[j])

    th = threading.Thread(target=late)
    for j in range(0, len(reqs), 2):
        ids[j] = submit(srv, *reqs[j])
        if j == 0:
            th.start()
    th.join()
    for j, i in enumerate(ids):
        same(srv.wait(i)[0], outs[j])
    srv.close()


def test_logdet_vs_oracle(G):
    med = G.Medium(3, 80, G.HARDSPHERE, (0.1,)).fill(21)
    ks = np.array([3.0 - 0.2j, 6.5 - 0.05j, 9.0 + 0j])
    srv = G.Server()
    out = srv.wait(srv.submit_logdet(med, ks, G.FP64))[0]["tracelog"]
    srv.close()
    pos = O.fill_medium(3, 80, 21)
    for k, t in zip(ks, out):
        want = O.tracelog(O.build_matrix(pos, k, O.HARDSPHERE, (0.1,)))
        assert abs(t.real - want.real) <= 1e-10 * max(1, abs(want.real))


def test_malformed_request_rejected(G):
    srv = G.Server()
    med = G.Medium(3, 10, G.HARDSPHERE, (0.1,)).fill(1)
    with pytest.raises(G.MsgpuError):
        srv.submit_logdet(med, np.zeros(0, np.complex128))
    srv.close()


def test_large_requests_mix_equals_alone(G):
    """Requests of the serving config's upper N range: a d=3 N=4096 log-det tile and a d=2
    N=4500 mixed-precision solve with a psi grid, each above the admission budget (they run
    alone), among small requests; every result is the same alone and in the mix."""
    rng = np.random.default_rng(40)
    big_tile = G.Medium(3, 4096, G.HARDSPHERE, (0.1,)).fill(41)
    big_solve = G.Medium(2, 4500, G.HARDSPHERE, (0.1,)).fill(42)
    small = G.Medium(3, 300, G.HARDSPHERE, (0.1,)).fill(43)
    reqs = [("logdet", big_tile, dict(ks=rng.uniform(5, 30, 8) - 1j * rng.uniform(0, 2, 8), mode=G.FP64,
                                      reeval_rel=1e-6)),
            ("logdet", small, dict(ks=rng.uniform(2, 10, 16) - 1j * rng.uniform(0, 1, 16), mode=G.MIXED,
                                   reeval_rel=1e-2)),
            ("solve", big_solve, dict(k=6.0, mode=G.MIXED, tol=1e-12, grid=(32, 24, -30.0, 30.0, -25.0, 25.0))),
            ("solve", small, dict(k=5.0, mode=G.FP64, tol=0.0, grid=(16, 12, -5.0, 5.0, -4.0, 4.0)))]
    budget_gb = 0.25
    for kind, med, a in reqs[::2]:
        assert footprint_bound(kind, med, a) > budget_gb * 2 ** 30
    outs = []
    for kind, med, a in reqs:
        srv = G.Server(hbm_gb=64.0)
        print("alone", kind, med.n, flush=True)
        outs.append(srv.wait(submit(srv, kind, med, a))[0])
        srv.close()
    srv = G.Server(hbm_gb=budget_gb, max_active=8)
    ids = [submit(srv, kind, med, a) for kind, med, a in reqs]
    for i, want in zip(ids, outs):
        same(srv.wait(i)[0], want)
    srv.close()
    # the big solve reached the refinement tolerance (or its complex128 fallback)
    aw, steps = G.solve(big_solve, 6.0, G.PLANE, G.MIXED, 1e-12)
    assert np.array_equal(outs[2]["a"], aw) and steps >= 1
    assert np.isfinite(outs[0]["tracelog"]).all()


def test_large_request_not_starved(G):
    """A request larger than the budget behind a steady stream of small ones: it is passed
    over a bounded number of times, then holds the queue until the device has drained and
    runs alone."""
    small = G.Medium(3, 200, G.HARDSPHERE, (0.1,)).fill(51)
    big = G.Medium(3, 1000, G.HARDSPHERE, (0.1,)).fill(52)
    ks_small = np.linspace(3.0, 9.0, 24) - 0.3j          # 24 x 16 x 200^2 B = 15 MB each
    ks_big = np.linspace(5.0, 25.0, 8) - 0.5j             # 128 MB, above the 50 MB budget
    srv = G.Server(hbm_gb=50e6 / 2 ** 30, max_active=8)
    t_sub, ids = [], []
    for j in range(43):
        if j == 3:
            t_big = time.perf_counter()
            big_id = srv.submit_logdet(big, ks_big, G.FP64)
        t_sub.append(time.perf_counter())
        ids.append(srv.submit_logdet(small, ks_small, G.FP64))
        time.sleep(0.001)
    admit = [t + srv.wait(i)[1]["queue_ms"] * 1e-3 for t, i in zip(t_sub, ids)]
    out_big, info_big = srv.wait(big_id)
    srv.close()
    admit_big = t_big + info_big["queue_ms"] * 1e-3
    passed = sum(1 for j in range(3, 43) if admit[j] < admit_big)
    print("small requests submitted after the large one and admitted before it:", passed)
    assert passed <= 6, passed                           # at most 4 bypasses (+ clock slack)
    assert np.isfinite(out_big["tracelog"]).all()


def test_kept_factors_and_resident_media(G):
    """Complex128 solves keep their factored M(k) on the device: a later solve over the same
    medium, model and k (another source, another grid) uses it instead of factoring again,
    with amplitudes and psi identical to the direct calls; media are shared by content."""
    med = G.Medium(3, 300, G.HARDSPHERE, (0.1,)).fill(21)
    twin = G.Medium(3, 300, G.HARDSPHERE, (0.1,)).fill(21)      # same content, another object
    grid = (16, 12, -3.0, 3.0, -2.0, 2.0)
    cases = [(med, 5.0 - 0.1j, G.PLANE), (twin, 5.0 - 0.1j, G.SPHERICAL), (med, 6.0 - 0.2j, G.PLANE),
             (med, 5.0 - 0.1j, G.PLANE)]
    srv = G.Server(hbm_gb=4.0)
    outs = []
    for m, k, src in cases:              # one after another: each finds the factors kept before it
        rid = srv.submit_solve(m, k, src, G.FP64, 0.0, grid)
        outs.append(srv.wait(rid)[0])
    st = srv.stats()
    srv.close()
    for (m, k, src), out in zip(cases, outs):
        aw, _ = G.solve(m, k, src, G.FP64)
        assert np.array_equal(out["a"], aw)
        assert np.array_equal(out["psi"], G.wavefunction_grid(m, k, aw, *grid, source=src).ravel())
    npad = (300 + 63) // 64 * 64
    print(st)
    assert (st["factor_hits"], st["factor_misses"]) == (2, 2), st
    assert st["kept_factor_bytes"] == 2 * (16 * npad * npad + 4 * npad), st
    assert st["resident_media"] == 1, st


def test_kept_factors_are_evicted_for_room(G):
    """A budget whose half holds one factor: the second k evicts the first, so the first k
    factors again; a log-det tile that needs the room drops the kept factor."""
    med = G.Medium(3, 300, G.HARDSPHERE, (0.1,)).fill(22)
    npad = (300 + 63) // 64 * 64
    fbytes = 16 * npad * npad + 4 * npad
    srv = G.Server(hbm_gb=3.0 * fbytes / 2 ** 30)
    got = []
    for k in (4.0 - 0.1j, 4.5 - 0.1j, 4.0 - 0.1j):
        got.append(srv.wait(srv.submit_solve(med, k, G.PLANE, G.FP64))[0]["a"])
    st = srv.stats()
    assert (st["factor_hits"], st["factor_misses"]) == (0, 3), st
    assert st["kept_factor_bytes"] == fbytes, st
    ks = np.array([4.0 - 0.1j, 4.2 - 0.3j, 4.4 - 0.2j])
    tl = srv.wait(srv.submit_logdet(med, ks, G.FP64))[0]["tracelog"]
    st = srv.stats()
    srv.close()
    assert st["kept_factor_bytes"] == 0, st           # 3 x 16 N^2 did not fit beside it
    assert st["peak_admitted_bytes"] <= st["budget_bytes"], st
    want, _ = G.tracelog_batch(med, ks, G.FP64)
    assert np.array_equal(tl, want)
    for k, a in zip((4.0 - 0.1j, 4.5 - 0.1j, 4.0 - 0.1j), got):
        assert np.array_equal(a, G.solve(med, k, G.PLANE, G.FP64)[0])
