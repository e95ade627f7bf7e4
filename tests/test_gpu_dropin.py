"""The drop-in: MSModel's own msmain, relinked so that its matrix, LAPACK and
wavefunction calls resolve to libmsgpu.so (oracle/_ref/msmain_gpu, made by
`make -f oracle/ref.mk gpu`), against the unmodified msmain (MKL) on the same
configuration files.  Both programs run the reference's task drivers
(run_tracelog_omp over maps, cuts and log cuts, run_wave_function_omp, and the
[eigenstate_list] driver with each of its three root finders:
find_eigenstates_serial_1/_2/_3 then compute_eigenstates) with two OpenMP
threads, so the library is also called from concurrent host threads.

Compared outputs (the .dat file msmain writes, complex_data_util.c:21-28,
wave_function_util.c:228-245, eigenstate_list_util.c:296-317, eigenstate_util.c:202-213):
  [kplane] [kcut] [logkcut] data   tracelog: Re to 1e-10 relative, N*Im modulo 2 pi to 1e-10
  [wavefunction]     Re psi and |psi|^2 on the grid: 1e-9 of the grid maximum
  [eigenstate_list]  the root sets both ways (1e-10 |k|); per state mu to 1e-12 and the
                     phase-normalised eigenvector to 1e-10

The two executables are built in the container that holds the MSModel sources and
travel with the tree; the test fails (it does not skip) when they are missing.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT, wrapped

REF = os.path.join(ROOT, "oracle", "_ref")
CPU_EXE = os.path.join(REF, "msmain")
GPU_EXE = os.path.join(REF, "msmain_gpu")

pytestmark = pytest.mark.gpu

CONF_D2 = """[medium]
dimension=2
natom=30
model=hardsphere 0.1
shape=ball
ratio=1

[kplane]
nseed=1
xrange=2:8
yrange=-1:0
xsample=8
ysample=4
color=temperature div 5
title=dropin

[wavefunction]
nseed=1
iseed=1
k=6.+0i
source=plane
nx=20
ny=16
xrange=-4:4
yrange=-4:4
color1=temperature div 5
color2=temperature div 5

[kcut]
nseed=1
kmin=2-0.8i
kmax=7-0.1i
sample=9
title=dropincut

[logkcut]
nseed=1
rek=4.5
imkmin=-0.001
imkmax=-1.5
sample=11
title=dropinlog

[eigenstate_list]
seed=1
ntarget=12
xrange=2:6
yrange=-1:0
method=determinant
maxit=60
toler=1e-10
verb=0

[eigenstate_list]
seed=1
ntarget=12
xrange=2:6
yrange=-1:0
method=mineigval
maxit=30
toler=1e-12
verb=0

[eigenstate_list]
seed=1
ntarget=12
xrange=2:6
yrange=-1:0
method=invtraceinv
maxit=30
toler=1e-12
verb=0
"""

CONF_D3 = """[medium]
dimension=3
natom=150
model=hardsphere 0.1
shape=ball
ratio=1

[kplane]
nseed=2
xrange=1:12
yrange=-1:0
xsample=6
ysample=3
color=temperature div 5
title=dropin3

[wavefunction]
nseed=1
iseed=3
k=5.-0.2i
source=spherical
nx=12
ny=10
xrange=-3:3
yrange=-2.5:2.5
color1=temperature div 5
color2=temperature div 5

[kcut]
nseed=1
kmin=3-0.6i
kmax=9-0.05i
sample=7
title=dropincut3

[logkcut]
nseed=1
rek=6.5
imkmin=-0.002
imkmax=-1.2
sample=7
title=dropinlog3

[eigenstate_list]
seed=2
ntarget=8
xrange=2:6
yrange=-1:0
method=determinant
maxit=60
toler=1e-10
verb=0
"""


def run(exe, conf, tmp, extra_env=None):
    assert os.path.exists(exe), "%s is missing: build it with `make -f oracle/ref.mk all gpu`" % exe
    os.makedirs(tmp, exist_ok=True)
    path = os.path.join(tmp, "t.conf")
    with open(path, "w") as f:
        f.write(conf)
    env = dict(os.environ, MKL_THREADING_LAYER="GNU", MKL_NUM_THREADS="1", OMP_NUM_THREADS="2")
    env.update(extra_env or {})
    print("running", os.path.basename(exe), "on", path, extra_env or "", flush=True)
    subprocess.run([exe, "-p", "2", path], cwd=tmp, env=env, check=True, stdout=subprocess.DEVNULL,
                   stderr=subprocess.PIPE, timeout=240)
    dats = [os.path.join(dp, fn) for dp, _, fs in os.walk(os.path.join(tmp, "out")) for fn in fs if fn.endswith(".dat")]
    assert len(dats) == 1, dats
    return parse_dat(dats[0])


def parse_dat(path):
    """{section name: [one dict per occurrence, in file order]}"""
    sections, cur = {}, None
    for line in open(path):
        line = line.strip()
        if line.startswith("[") and line.endswith("]"):
            cur = {}
            sections.setdefault(line[1:-1], []).append(cur)
        elif "=" in line and cur is not None:
            key, val = line.split("=", 1)
            cur[key] = val
    return sections


def cvec(text):
    return np.array([complex(t.replace("i", "j")) for t in text.split()])


def rvec(text):
    return np.array([float(t) for t in text.split()])


@pytest.fixture(scope="module", params=["d2", "d3"])
def outputs(request, tmp_path_factory):
    conf = CONF_D2 if request.param == "d2" else CONF_D3
    base = tmp_path_factory.mktemp("dropin_" + request.param)
    return run(CPU_EXE, conf, str(base / "cpu")), run(GPU_EXE, conf, str(base / "gpu"))


@pytest.mark.parametrize("section", ["kplane", "kcut", "logkcut"])
def test_tracelog_tasks(outputs, section):
    cpu, gpu = outputs
    assert section in cpu and section in gpu, section
    n = int(cpu["medium"][0]["natom"])
    for sc, sg in zip(cpu[section], gpu[section]):
        a, b = cvec(sc["data"]), cvec(sg["data"])
        assert a.shape == b.shape and a.size > 0
        ere = np.abs(a.real - b.real).max() / max(1.0, np.abs(a.real).max())
        earg = np.abs(wrapped(n * (a.imag - b.imag))).max()
        print("[%s] %d cells: max rel |d Re| %.2e, max |d arg det| %.2e" % (section, a.size, ere, earg))
        assert ere <= 1e-10
        assert earg <= 1e-10


def test_tracelog_keep_input_same_map(tmp_path):
    """MSGPU_TRACELOG_KEEP_INPUT=1 (tracelog_lu leaves the matrix as given instead of writing
    zgetrf_'s factor back) changes no trace-log that msmain writes."""
    plain = run(GPU_EXE, CONF_D3, str(tmp_path / "plain"))
    keep = run(GPU_EXE, CONF_D3, str(tmp_path / "keep"), {"MSGPU_TRACELOG_KEEP_INPUT": "1"})
    for sp, sk in zip(plain["kplane"], keep["kplane"]):
        assert sp["data"] == sk["data"]


def test_tracelog_keep_input_leaves_matrix(tmp_path):
    """In a process with MSGPU_TRACELOG_KEEP_INPUT=1, tracelog_lu returns the same value and
    leaves its argument unchanged."""
    code = (
        "import numpy as np, msmodel_amd as G\n"
        "rng = np.random.default_rng(3)\n"
        "m = rng.normal(size=(90, 90)) + 1j * rng.normal(size=(90, 90))\n"
        "t, a = G.tracelog_lu(m)\n"
        "assert np.array_equal(a, m)\n"
        "np.save(%r, np.array([t]))\n" % str(tmp_path / "t.npy"))
    env = dict(os.environ, MSGPU_TRACELOG_KEEP_INPUT="1",
               PYTHONPATH=os.pathsep.join([ROOT] + os.environ.get("PYTHONPATH", "").split(os.pathsep)))
    print("running a python process with MSGPU_TRACELOG_KEEP_INPUT=1", flush=True)
    subprocess.run([sys.executable, "-c", code], env=env, check=True, timeout=240)
    import msmodel_amd as G
    rng = np.random.default_rng(3)
    m = rng.normal(size=(90, 90)) + 1j * rng.normal(size=(90, 90))
    t, a = G.tracelog_lu(m)
    assert not np.array_equal(a, m)            # the default writes the factor back
    assert np.load(str(tmp_path / "t.npy"))[0] == t


def test_wavefunction(outputs):
    cpu, gpu = outputs
    for key in ("function", "density"):
        a, b = rvec(cpu["wavefunction"][0][key]), rvec(gpu["wavefunction"][0][key])
        assert a.shape == b.shape and a.size > 0
        assert np.abs(a - b).max() <= 1e-9 * np.abs(a).max(), key


def test_eigenstate_lists(outputs):
    cpu, gpu = outputs
    assert "eigenstate_list" in cpu and "eigenstate_list" in gpu
    assert len(cpu["eigenstate_list"]) == len(gpu["eigenstate_list"])
    for sc, sg in zip(cpu["eigenstate_list"], gpu["eigenstate_list"]):
        assert sc["method"] == sg["method"]

        def states(sec):
            ns = int(sec["nstate"])
            k = np.array([complex(sec["st%d_kroot" % i].replace("i", "j")) for i in range(ns)])
            mu = np.array([complex(sec["st%d_mu" % i].replace("i", "j")) for i in range(ns)])
            v = [cvec(sec["st%d_eigv" % i]) for i in range(ns)]
            return k, mu, v

        (ka, mua, va), (kb, mub, vb) = states(sc), states(sg)
        assert ka.size > 0 and ka.size == kb.size, (sc["method"], ka, kb)
        for z in ka:
            assert np.abs(kb - z).min() <= 1e-10 * abs(z), (sc["method"], z)
        for z in kb:
            assert np.abs(ka - z).min() <= 1e-10 * abs(z), (sc["method"], z)
        # the states in the same order (sorted by -Im k): eigenvalue and eigenvector
        dmu = max(abs(x - y) for x, y in zip(mua, mub))
        dv = max(np.abs(x - y).max() for x, y in zip(va, vb))
        print("[eigenstate_list] %s: %d states, max |d k| %.2e, max |d mu| %.2e, max |d v| %.2e"
              % (sc["method"], ka.size, np.abs(ka - kb).max(), dmu, dv))
        assert np.abs(ka - kb).max() <= 1e-10 * np.abs(ka).max()
        assert dmu <= 1e-12
        assert dv <= 1e-10
