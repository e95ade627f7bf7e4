"""The autotuner and the tuning table.  The swap + TRSM tile shape, panel shape, chunk
size, stream count, graph replay and the 128-column blocking only move data, change
launch shapes or keep the accumulation order, so results are bit-identical whatever
the table holds; the row-major factorisation rounds differently and is held to the
parity tolerance."""
import os

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import msmodel_amd as G
    G.init()
    return G


def test_autotune_report_and_results(G, tmp_path):
    # start from an explicit entry for Np = 320 (column-major, 64-column blocking): a
    # size without one takes the nearest size's entry
    path = os.path.join(str(tmp_path), "t320.txt")
    with open(path, "w") as f:
        f.write("320 0 0 256 4 0 0 0 0\n")
    assert G.tune_load(path) == 1
    med = G.Medium(3, 300, G.HARDSPHERE, (0.05,)).fill(3)
    ks = np.linspace(5, 30, 40) - 1j * np.linspace(0, 2, 40)
    before, _ = G.tracelog_batch(med, ks, G.FP64)
    dk_before = G.dk_log_det_batch(med, ks[:6], G.FP64)
    rep = G.autotune(300, G.FP64, 64)
    assert 0 <= rep["swap_variant"] < 6 and rep["chunk"] in (128, 256, 512) and rep["streams"] in (1, 2, 4)
    assert rep["row_major"] in (0, 1) and rep["panel_variant"] in (0, 1, 2, 3, 4) and rep["block128"] in (0, 1)
    assert rep["best_ms"] > 0
    after, _ = G.tracelog_batch(med, ks, G.FP64)
    if rep["row_major"] == 0 and rep["block128"] == 0:     # launch shapes only: exact
        assert np.array_equal(before, after)
    else:
        assert np.abs(after.real - before.real).max() <= 1e-10 * max(1, np.abs(before.real).max())
    dk_after = G.dk_log_det_batch(med, ks[:6], G.FP64)
    if rep["block128"] == 0:       # tr(M^-1 M') always factors row-major
        assert np.array_equal(dk_before, dk_after)
    else:
        assert np.abs(dk_after - dk_before).max() <= 1e-8 * np.abs(dk_before).max()


@pytest.mark.parametrize("prec_mode", ["FP64", "FP32"])
def test_every_swap_variant_same_result(G, tmp_path, prec_mode):
    mode = getattr(G, prec_mode)
    prec = 0 if mode == G.FP64 else 1
    med = G.Medium(3, 330, G.HARDSPHERE, (0.05,)).fill(5)    # Np = 384: six panels
    ks = np.linspace(6, 25, 24) - 0.5j
    ref = None
    for v in range(6):
        for chunk, streams, rm, pv, b2 in ((256, 4, 0, 0, 0), (16, 1, 0, 1, 1), (256, 2, 1, 2, 0), (16, 1, 1, v % 5, 1),
                                           (128, 2, 0, 3, 1), (64, 1, 0, 4, 0), (32, 2, 1, 5 + v % 3, 1),
                                           (32, 1, 0, 5 + (v + 1) % 3, 0)):
            path = os.path.join(str(tmp_path), "tune_%d_%d_%d_%d.txt" % (v, chunk, rm, b2))
            with open(path, "w") as f:
                f.write("384 %d %d %d %d %d %d 0 %d\n" % (prec, v, chunk, streams, rm, pv, b2))
            assert G.tune_load(path) == 1
            out, _ = G.tracelog_batch(med, ks, mode)
            if ref is None:
                ref = out
            if rm == 0 and b2 == 0 and pv < 5:     # launch shapes only: exact
                assert np.array_equal(out, ref), (v, chunk, streams, b2)
            else:   # row-major storage, the 128-column blocking (3M updates) and the left-looking
                    # panels (another summation order) round differently
                tol = 1e-10 if prec == 0 else 1e-4
                assert np.abs(out.real - ref.real).max() <= tol * max(1, np.abs(ref.real).max())
                dim = (330 * (out.imag - ref.imag) + np.pi) % (2 * np.pi) - np.pi
                assert np.abs(dim).max() <= (1e-8 if prec == 0 else 1e-2)
    path = os.path.join(str(tmp_path), "tune_restore.txt")
    with open(path, "w") as f:
        f.write("384 %d 0 256 4\n" % prec)
    G.tune_load(path)


@pytest.mark.parametrize("prec_mode", ["FP64", "FP32"])
def test_graph_replays_match_eager(G, tmp_path, prec_mode):
    """Chunks replayed from a captured HIP graph give the eager results bit for bit, on
    several replays with different inputs and with eager device work in between."""
    mode = getattr(G, prec_mode)
    prec = 0 if mode == G.FP64 else 1
    med = G.Medium(3, 200, G.HARDSPHERE, (0.1,)).fill(1)       # Np = 256
    other = G.Medium(2, 90, G.HARDSPHERE, (0.2,)).fill(4)
    sets = [np.linspace(1 + r, 20 - r, 40) - 1j * np.linspace(0, 1, 40)[::-1] * (0.5 + 0.1 * r) for r in range(4)]

    def table(graph):
        path = os.path.join(str(tmp_path), "g%d.txt" % graph)
        with open(path, "w") as f:
            f.write("256 %d 0 16 2 0 0 %d\n" % (prec, graph))
        assert G.tune_load(path) == 1

    table(0)
    eager = [G.tracelog_batch(med, ks, mode)[0] for ks in sets]
    table(1)
    for rep in range(2):
        for ks, want in zip(sets, eager):
            got, _ = G.tracelog_batch(med, ks, mode)
            assert np.array_equal(got, want), rep
            G.dk_log_det_batch(other, ks[:3], G.FP64)          # eager work between replays
    table(0)


@pytest.mark.parametrize("prec_mode", ["FP64", "FP32"])
@pytest.mark.parametrize("n,np_,variants", [(200, 256, (0, 1, 2, 3, 4, 5, 6, 7)), (500, 512, (0, 1, 2, 3, 4, 5, 6, 7)),
                                             (1000, 1024, (0, 1, 2, 3, 4, 5, 6, 7)), (1100, 1152, (0, 1, 2, 3, 5, 6)),
                                             (2000, 2048, (0, 3, 5, 6))])
def test_panel_shapes_beyond_one_register_row(G, tmp_path, prec_mode, n, np_, variants):
    """Panel shapes whose threads hold one or several rows of the panel (Np above the
    thread count: the rank-S update stages one register row at a time), right-looking (0-4)
    and left-looking (5-7), 256 to 1024 threads: every shape gives the same trace-log,
    row-major with 128-column blocks, and one cell matches the oracle."""
    mode = getattr(G, prec_mode)
    prec = 0 if mode == G.FP64 else 1
    pos = O.fill_medium(3, n, 12)
    med = G.Medium(3, n, G.HARDSPHERE, (0.05,))
    med.set_pos(pos)
    ks = np.array([6.0 - 0.3j, 14.0 - 1.1j, 22.0 - 0.05j])
    outs = []
    for pv in variants:
        path = os.path.join(str(tmp_path), "panel_%d.txt" % pv)
        with open(path, "w") as f:
            f.write("%d %d 4 64 1 1 %d 0 1\n" % (np_, prec, pv))
        assert G.tune_load(path) == 1
        outs.append(G.tracelog_batch(med, ks, mode)[0])
    tol = 1e-12 if prec == 0 else 1e-5
    for pv, out in zip(variants, outs):
        assert np.abs(out.real - outs[0].real).max() <= tol * np.abs(outs[0].real).max(), pv
        dim = (n * (out.imag - outs[0].imag) + np.pi) % (2 * np.pi) - np.pi
        assert np.abs(dim).max() <= (1e-9 if prec == 0 else 1e-2), pv
    want = O.tracelog(O.build_matrix(pos, ks[1], O.HARDSPHERE, (0.05,)))
    assert abs(outs[-1][1].real - want.real) <= (1e-10 if prec == 0 else 1e-3) * abs(want.real)
    G.tune_load(G.tuning_table_path())
