"""The allocation-free entry point msgpu_tracelog_batch_noalloc against the allocating
msgpu_tracelog_batch, through the C ABI: bit-for-bit equal results on two calls in a row
(the second with a smaller input), in FP64 and FP32; sizes beyond its scratch and the
MIXED mode are rejected with an error before any device work; and the calls leave the
device's free memory where it was (they allocate nothing).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import msmodel_amd as G
    G.init()
    return G


def device_free_bytes():
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        return None
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    if hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) != 0:
        return None
    return free.value


def kmap(nx, ny, x0, x1, y0, y1):
    l = np.arange(nx * ny)
    return (x0 + (l % nx) * (x1 - x0) / (nx - 1)) + 1j * (y1 - (l // nx) * (y1 - y0) / (ny - 1))


def test_noalloc_equals_batch_two_calls(G):
    big = G.Medium(3, 1000, G.HARDSPHERE, (0.05,)).fill(1)
    small = G.Medium(2, 300, G.HARDSPHERE, (0.1,)).fill(7)
    ks_big = kmap(8, 5, 5.0, 30.0, -2.0, 0.0)            # 40 cells: three chunks of the scratch
    ks_small = np.linspace(2.0, 9.0, 7) - 0.5j           # 7 cells
    want_big, _ = G.tracelog_batch(big, ks_big, G.FP64)
    want_small, _ = G.tracelog_batch(small, ks_small, G.FP64)
    free0 = device_free_bytes()
    print("case: N=1000 d=3 40 cells FP64, then N=300 d=2 7 cells FP64", flush=True)
    got_big = G.tracelog_batch_noalloc(big, ks_big, G.FP64)
    got_small = G.tracelog_batch_noalloc(small, ks_small, G.FP64)
    free1 = device_free_bytes()
    assert np.array_equal(got_big, want_big)
    assert np.array_equal(got_small, want_small)
    if free0 is not None and free1 is not None:
        print("device free bytes before %d after %d" % (free0, free1))
        assert free0 - free1 < 32 << 20


def test_noalloc_fp32_and_limits(G):
    med = G.Medium(3, 500, G.HARDSPHERE, (0.05,)).fill(3)
    ks = kmap(5, 4, 5.0, 30.0, -2.0, 0.0)
    print("case: N=500 d=3 20 cells FP32", flush=True)
    want, _ = G.tracelog_batch(med, ks, G.FP32)
    assert np.array_equal(G.tracelog_batch_noalloc(med, ks, G.FP32), want)
    assert G.tracelog_batch_noalloc(med, np.zeros(0, np.complex128), G.FP64).size == 0
    with pytest.raises(G.MsgpuError, match="MSGPU_FP64 or MSGPU_FP32"):
        G.tracelog_batch_noalloc(med, ks, G.MIXED)
    huge = G.Medium(3, 4100, G.HARDSPHERE, (0.05,))
    huge.set_pos(np.zeros((4100, 3)))                    # rejected before any device work
    with pytest.raises(G.MsgpuError, match="allocation-free scratch holds"):
        G.tracelog_batch_noalloc(huge, ks[:2], G.FP64)
    # the largest complex128 size it takes, once
    top = G.Medium(3, 4032, G.HARDSPHERE, (0.05,)).fill(5)
    print("case: N=4032 d=3 2 cells FP64", flush=True)
    want, _ = G.tracelog_batch(top, ks[:2], G.FP64)
    assert np.array_equal(G.tracelog_batch_noalloc(top, ks[:2], G.FP64), want)
