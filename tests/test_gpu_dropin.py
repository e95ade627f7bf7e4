"""The drop-in: MSModel's own msmain, relinked so that its matrix, LAPACK and
wavefunction calls resolve to libmsgpu.so (oracle/_ref/msmain_gpu, made by
`make -f oracle/ref.mk gpu`), against the unmodified msmain (MKL) on the same
configuration files.  Both programs run the reference's task drivers
(run_tracelog_omp, run_wave_function_omp, find_eigenstates_serial_1) with two
OpenMP threads, so the library is also called from concurrent host threads.

Compared outputs (the .dat file msmain writes, complex_data_util.c:21-28,
wave_function_util.c:228-245, eigenstate_list_util.c:296-317):
  [kplane] data      tracelog map: Re to 1e-10 relative, N*Im modulo 2 pi to 1e-8
  [wavefunction]     Re psi and |psi|^2 on the grid: 1e-9 of the grid maximum
  [eigenstate_list]  the root sets: every root within 1e-9 of one root of the other
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, wrapped

REF = os.path.join(ROOT, "oracle", "_ref")
CPU_EXE = os.path.join(REF, "msmain")
GPU_EXE = os.path.join(REF, "msmain_gpu")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (os.path.exists(CPU_EXE) and os.path.exists(GPU_EXE)),
                                 reason="oracle/_ref/msmain and msmain_gpu are built from the MSModel sources")]

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

[eigenstate_list]
seed=1
ntarget=12
xrange=2:6
yrange=-1:0
method=determinant
maxit=60
toler=1e-10
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
"""


def run(exe, conf, tmp):
    os.makedirs(tmp, exist_ok=True)
    path = os.path.join(tmp, "t.conf")
    with open(path, "w") as f:
        f.write(conf)
    env = dict(os.environ, MKL_THREADING_LAYER="GNU", MKL_NUM_THREADS="1", OMP_NUM_THREADS="2")
    print("running", os.path.basename(exe), "on", path, flush=True)
    subprocess.run([exe, "-p", "2", path], cwd=tmp, env=env, check=True, stdout=subprocess.DEVNULL,
                   stderr=subprocess.PIPE, timeout=240)
    dats = [os.path.join(dp, fn) for dp, _, fs in os.walk(os.path.join(tmp, "out")) for fn in fs if fn.endswith(".dat")]
    assert len(dats) == 1, dats
    return parse_dat(dats[0])


def parse_dat(path):
    sections, cur = {}, None
    for line in open(path):
        line = line.strip()
        if line.startswith("[") and line.endswith("]"):
            cur = sections.setdefault(line[1:-1], {})
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


def test_kplane_map(outputs):
    cpu, gpu = outputs
    n = int(cpu["medium"]["natom"])
    a, b = cvec(cpu["kplane"]["data"]), cvec(gpu["kplane"]["data"])
    assert a.shape == b.shape and a.size > 0
    assert np.abs(a.real - b.real).max() <= 1e-10 * max(1.0, np.abs(a.real).max())
    assert np.abs(wrapped(n * (a.imag - b.imag))).max() <= 1e-8


def test_wavefunction(outputs):
    cpu, gpu = outputs
    for key in ("function", "density"):
        a, b = rvec(cpu["wavefunction"][key]), rvec(gpu["wavefunction"][key])
        assert a.shape == b.shape and a.size > 0
        assert np.abs(a - b).max() <= 1e-9 * np.abs(a).max(), key


def test_eigenstate_roots(outputs):
    cpu, gpu = outputs
    if "eigenstate_list" not in cpu:
        pytest.skip("no eigenstate section in this configuration")

    def roots(sec):
        return np.array([complex(sec["st%d_kroot" % i].replace("i", "j")) for i in range(int(sec["nstate"]))])

    a, b = roots(cpu["eigenstate_list"]), roots(gpu["eigenstate_list"])
    assert a.size > 0 and a.size == b.size
    for z in a:
        assert np.abs(b - z).min() <= 1e-9 * max(1.0, abs(z)), z
    for z in b:
        assert np.abs(a - z).min() <= 1e-9 * max(1.0, abs(z)), z
