"""The smoke advection's inverse-distance weights take a refined hardware
reciprocal in place of IEEE divisions (fluid_device.hpp: recip_nr,
inv_distance, idw_ratio), and the projection rounds two faces to 16-bit
storage with one packed conversion (rnd2).  These checks run every binary32
input through both forms on the GPU: the reciprocal path of
1.0 / (distance + 1e-6) (src/fluid.cu:690-693) matches bit for bit, the
refined reciprocal keeps the 2^-50 bound the weights' quotient relies on,
the quotient matches IEEE division on 2^32 pairs, and the packed roundings
match the scalar ones."""
import ctypes

import pytest

from opensayal_amd import _abi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode, what", [(0, "1/(s + 1e-6) in double, to float"),
                                        (1, "reciprocal within 2^-50"),
                                        (2, "weight quotient vs IEEE division"),
                                        (3, "paired FP16 rounding vs scalar"),
                                        (4, "paired BF16 rounding vs scalar"),
                                        (5, "pressure quotient, d_t = 0.05"),
                                        (6, "pressure quotient, d_t = 1/60"),
                                        (7, "corrected square root vs sqrtf")])
def test_arith_shortcut_exhaustive(mode, what):
    n = ctypes.c_uint64()
    first = ctypes.c_uint32()
    _abi.check(_abi.lib().osy_arith_check(mode, ctypes.byref(n), ctypes.byref(first)))
    assert n.value == 0, f"{what}: {n.value} inputs differ, first 0x{first.value:08x}"
    assert first.value == 0xFFFFFFFF
