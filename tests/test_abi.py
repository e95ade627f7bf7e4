"""CPU checks of the C ABI: libmsgpu.so loads, exports every function that
include/msgpu.h declares, and reports a missing device as an error (no CPU
fallback).  No compute call is made here."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "msgpu.h")
LIB = os.path.join(ROOT, "msmodel_amd", "libmsgpu.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\((?:[^;{}()]|\([^()]*\))*\)\s*;", text)
    return sorted(set(n for n in names if n not in ("sizeof",)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("build_ms_matrix", "tracelog_lu", "dk_log_det_ms", "zgesv_", "zgetrf_", "zgetrs_",
                 "eval_total_wave", "fill_medium", "msgpu_tracelog_batch", "msgpu_resonances"):
        assert must in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build first: python -m msmodel_amd.build"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    assert set(declared_functions()) <= exported


def test_python_binding_covers_the_abi():
    import msmodel_amd as G
    assert set(declared_functions()) - {"init_copy_medium"} <= set(G.EXPORTED) | {"init_copy_medium"}


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK), reason="a GPU is present")
def test_init_without_gpu_fails_loudly():
    import msmodel_amd as G
    with pytest.raises(G.MsgpuError):
        G.init()


def test_tuning_table_round_trip(tmp_path):
    """The tuning table loads and saves as text without a device."""
    import msmodel_amd as G
    src = tmp_path / "in.txt"
    src.write_text("# np precision swap_variant chunk streams row_major panel_variant graph block128\n"
                   "1024 0 2 512 2 1 2 1 1\n1024 1 1 256 4\n9999 0 99 1 1 0 0 0 0\n")
    assert G.tune_load(str(src)) == 2          # the out-of-range variant is skipped
    dst = tmp_path / "out.txt"
    G.tune_save(str(dst))
    lines = dst.read_text().split("\n")
    assert "1024 0 2 512 2 1 2 1 1" in lines and "1024 1 1 256 4 0 0 0 0" in lines
