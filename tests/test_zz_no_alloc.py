"""osy_fluid_traces_noalloc (the path lines with library-owned staging, for
callers that cannot allocate) against osy_fluid_traces, both through the C
ABI: the same int32 entries, bit for bit, on two calls in a row, the second
asking for fewer entries, and on a second, smaller grid; requests beyond the
staging array's 262144 entries per buffer are refused before any device
work."""
import ctypes

import numpy as np
import pytest

from opensayal_amd import Fluid, OpenSayalError, _abi, workloads as wl
from opensayal_amd._abi import check, lib

pytestmark = pytest.mark.gpu

LIMIT = 1 << 18


def traces_via(fn_name, fl, distance, length, d_t):
    shape = (fl.height // distance, fl.width // distance, length)
    tx = np.full(shape, 0x5A5A5A5A, np.int32)
    ty = np.full(shape, 0x5A5A5A5A, np.int32)
    fn = getattr(lib(), fn_name)
    check(fn(fl._h, distance, length, d_t, tx.ctypes.data_as(ctypes.c_void_p),
             ty.ctypes.data_as(ctypes.c_void_p), tx.nbytes))
    return tx, ty


def stepped(cfg, steps, seed):
    fl = wl.start(wl.SimSpec(cfg, seed, steps))
    fl.step(steps)
    return fl


def assert_same(fl, distance, length, d_t):
    want = traces_via("osy_fluid_traces", fl, distance, length, d_t)
    got = traces_via("osy_fluid_traces_noalloc", fl, distance, length, d_t)
    for a, b in zip(got, want):
        assert a.shape == b.shape
        assert np.array_equal(a, b), f"{np.count_nonzero(a != b)} entries differ"
    return got


def test_noalloc_traces_match_allocating_entry_point():
    big = stepped(wl.plain_small(), 40, seed=21)
    # first call: 25 x 14 seeds x 200 points = 70,000 entries per buffer
    tx1, _ = assert_same(big, 10, 200, 0.05)
    assert (tx1 == -1).any() and (tx1 >= 0).any()  # solid and fluid seeds
    # second call on the same handle asks for fewer entries (12 x 7 x 9)
    assert_same(big, 20, 9, 0.05)
    # a second, smaller grid in 16-bit storage after both
    small = stepped(wl.wind_tunnel(96, 54, pipe_height=13, cx=40, cy=27, radius=5.0,
                                   storage="f16"), 25, seed=22)
    assert_same(small, 7, 5, 0.3)
    big.close()
    small.close()


def test_noalloc_traces_limit_and_arguments():
    fl = Fluid(wl.plain_small())
    # 256 x 144 seeds x 8 points = 294,912 entries: over the limit
    shape = (144, 256, 8)
    assert np.prod(shape) > LIMIT
    tx = np.zeros(shape, np.int32)
    ty = np.zeros(shape, np.int32)
    st = lib().osy_fluid_traces_noalloc(fl._h, 1, 8, 0.05, tx.ctypes.data_as(ctypes.c_void_p),
                                        ty.ctypes.data_as(ctypes.c_void_p), tx.nbytes)
    assert st == _abi.OSY_E_INVALID
    assert b"262144" in lib().osy_last_error()
    # exactly at the limit is served: 256 x 128 cells, distance 1, length 8
    fl2 = stepped(wl.wind_tunnel(256, 128, pipe_height=32, cx=100, cy=64, radius=8.0), 3, 23)
    got = assert_same(fl2, 1, 8, 0.05)
    assert got[0].size == LIMIT
    # a wrong buffer size is refused as by osy_fluid_traces
    st = lib().osy_fluid_traces_noalloc(fl._h, 20, 4, 0.05, tx.ctypes.data_as(ctypes.c_void_p),
                                        ty.ctypes.data_as(ctypes.c_void_p), 12)
    assert st == _abi.OSY_E_INVALID
    with pytest.raises(OpenSayalError):
        check(lib().osy_fluid_traces_noalloc(fl._h, 0, 4, 0.05, None, None, 0))
    fl.close()
    fl2.close()
