"""The drop-in eval_total_wave (incident_wave_util.c:87-96) as msmain's pixel loop calls it
(ms_core.c:629-640): one point per call, from many host threads at once.  By default the
calls are served by the resident psi kernel (requests in a ring of pinned host slots,
one workgroup per point); MSGPU_WAVE_RESIDENT=0 selects the launch-per-batch path.

Checked here:
  - every point against the oracle's eval_total_wave at 1e-10 relative, d = 1..4, plane
    and spherical sources, complex and real k, with calls on two contents interleaved
    from different threads;
  - more requests than the ring has slots (slot reuse) and a pause longer than the
    kernel's idle timeout between two rounds (the launch ends and is relaunched);
  - the resident and the launch-per-batch paths give the same bits.
"""
import os
import subprocess
import sys
import threading
import time

import numpy as np
import pytest

import pyoracle as O
from conftest import ROOT

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import msmodel_amd as G
    G.init()
    return G


def case(d, n, seed, k, source):
    pos = O.fill_medium(d, n, seed)
    rng = np.random.default_rng(seed)
    a = rng.normal(size=n) + 1j * rng.normal(size=n)
    pts = rng.uniform(-3, 3, size=(257, d))
    return pos, a, pts, k, source


CASES = [case(3, 1000, 2, 10.0 + 0j, O.PLANE), case(3, 300, 5, 5 - 0.2j, O.SPHERICAL),
         case(2, 400, 3, 6 - 0.1j, O.PLANE), case(1, 50, 4, 3 + 0j, O.PLANE), case(4, 120, 6, 4 - 0.3j, O.SPHERICAL)]


def medium(G, pos):
    n, d = pos.shape
    m = G.Medium(d, n, G.HARDSPHERE, (0.1,))
    m.set_pos(pos)
    return m


def run_threads(G, jobs, nthreads=8):
    """jobs: list of (medium, k, a, x, source); returns psi per job, computed by nthreads
    threads that each take every nthreads-th job (ctypes releases the GIL in the call)."""
    out = [None] * len(jobs)

    def work(t):
        for j in range(t, len(jobs), nthreads):
            m, k, a, x, src = jobs[j]
            out[j] = G.eval_total_wave(m, k, a, x, src)

    ts = [threading.Thread(target=work, args=(t,)) for t in range(nthreads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    return np.array(out)


def test_points_vs_oracle_interleaved_contents(G):
    meds = [medium(G, c[0]) for c in CASES]
    jobs, want = [], []
    for r in range(257):                   # the cases interleaved call by call
        for c, m in zip(CASES, meds):
            pos, a, pts, k, src = c
            jobs.append((m, k, a, pts[r], src))
            want.append(O.total_wave(pos, a, k, pts[r], src))
    print("%d calls from 8 threads over %d contents" % (len(jobs), len(CASES)), flush=True)
    got = run_threads(G, jobs)
    want = np.array(want)
    err = np.abs(got - want) / np.abs(want)
    print("max relative error %.2e" % err.max())
    assert err.max() < 1e-10


def test_ring_reuse_and_relaunch_after_idle(G):
    pos, a, pts, k, src = CASES[0]
    m = medium(G, pos)
    rng = np.random.default_rng(11)
    xs = rng.uniform(-4, 4, size=(9000, 3))                  # more requests than the 4096 slots
    jobs = [(m, k, a, x, src) for x in xs]
    t0 = time.perf_counter()
    first = run_threads(G, jobs[:6000], 16)
    dt = time.perf_counter() - t0
    print("6000 calls from 16 threads in %.3f s" % dt, flush=True)
    time.sleep(0.1)                                         # the launch idles out (20 ms) and ends
    second = run_threads(G, jobs[6000:], 16)
    got = np.concatenate([first, second])
    sub = np.arange(0, 9000, 97)
    want = np.array([O.total_wave(pos, a, k, xs[i], src) for i in sub])
    err = np.abs(got[sub] - want) / np.abs(want)
    print("max relative error %.2e over %d sampled points" % (err.max(), sub.size))
    assert err.max() < 1e-10
    # the same point asked twice gives the same bits
    again = run_threads(G, jobs[:50], 4)
    assert np.array_equal(again, first[:50])


def test_resident_equals_batched_launches(G, tmp_path):
    """The launch-per-batch path (MSGPU_WAVE_RESIDENT=0, in a process of its own) gives the
    same bits as the resident kernel."""
    pos, a, pts, k, src = CASES[1]
    np.savez(str(tmp_path / "in.npz"), pos=pos, a=a, pts=pts)
    code = (
        "import sys, numpy as np\n"
        "sys.path.insert(0, %r); sys.path.insert(0, %r)\n"
        "import test_gpu_wave_points as T, msmodel_amd as G\n"
        "G.init()\n"
        "z = np.load(%r)\n"
        "m = T.medium(G, z['pos'])\n"
        "jobs = [(m, %r, z['a'], x, %d) for x in z['pts']]\n"
        "np.save(%r, T.run_threads(G, jobs))\n"
        % (ROOT, os.path.join(ROOT, "tests"), str(tmp_path / "in.npz"), complex(k), src, str(tmp_path / "b.npy")))
    env = dict(os.environ, MSGPU_WAVE_RESIDENT="0",
               PYTHONPATH=os.pathsep.join([ROOT, os.path.join(ROOT, "oracle")] +
                                          os.environ.get("PYTHONPATH", "").split(os.pathsep)))
    print("running a python process with MSGPU_WAVE_RESIDENT=0", flush=True)
    subprocess.run([sys.executable, "-c", code], env=env, check=True, timeout=240)
    batched = np.load(str(tmp_path / "b.npy"))
    m = medium(G, pos)
    resident = run_threads(G, [(m, k, a, x, src) for x in pts])
    assert np.array_equal(resident, batched)
