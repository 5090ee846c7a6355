"""The oracle's 16-bit storage mode: rounding to binary16 / bfloat16 (round
to nearest, ties to even, subnormals kept) checked against torch's and
numpy's conversions over a large strided sample of all 2^32 binary32 values
and exhaustively around the subnormal and overflow boundaries; and the
state after a step is representable in the storage type."""
import warnings

import numpy as np
import pytest

from oracle.oracle import OracleFluid, round16
from tests.helpers import make_config, noise

torch = pytest.importorskip("torch")


def all_floats(stride, offset=0):
    k = np.arange(offset, 1 << 32, stride, dtype=np.uint64).astype(np.uint32)
    return k.view(np.float32)


def boundary_floats():
    parts = []
    for lo, hi in ((0x38000000, 0x39000000), (0x477f0000, 0x47800000),
                   (0x00000000, 0x00100000), (0x7f7f0000, 0x7f800001)):
        k = np.arange(lo, hi, dtype=np.uint64).astype(np.uint32)
        parts += [k, k | 0x80000000]
    return np.concatenate(parts).view(np.float32)


def ref16(x, storage):
    t = torch.from_numpy(x)
    dt = torch.float16 if storage == 1 else torch.bfloat16
    return t.to(dt).to(torch.float32).numpy()


def same_bits(a, b):
    nan = np.isnan(a) & np.isnan(b)
    return np.array_equal(a.view(np.uint32)[~nan], b.view(np.uint32)[~nan]) and \
        np.array_equal(np.isnan(a), np.isnan(b))


@pytest.mark.parametrize("storage", [1, 2])
def test_rounding_matches_torch(storage):
    for x in (all_floats(61), all_floats(61, 30), boundary_floats()):
        assert same_bits(round16(x, storage), ref16(x, storage))


def test_f16_rounding_matches_numpy():
    x = np.concatenate([all_floats(997), boundary_floats()[::7]])
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        ref = x.astype(np.float16).astype(np.float32)
    assert same_bits(round16(x, 1), ref)


def test_storage_zero_is_identity():
    x = all_floats(4099)
    assert same_bits(round16(x, 0), x)


@pytest.mark.parametrize("storage", [1, 2])
def test_state_is_representable_after_steps(storage):
    c = make_config(61, 37, projection_n=9, wind_speed=2.0, g=-5.0, enable_decay=1,
                    decay_rate=0.3, storage=storage)
    o = OracleFluid(c)
    o.set_state(u=noise((37, 61), 1, 0.3), v=noise((37, 61), 2, 0.3))
    o.update(steps=3)
    for a in (o.u, o.v, o.smoke):
        assert same_bits(round16(a, storage), a)
    # and differs from the single-precision run
    f = OracleFluid(make_config(61, 37, projection_n=9, wind_speed=2.0, g=-5.0, enable_decay=1,
                                decay_rate=0.3))
    f.set_state(u=noise((37, 61), 1, 0.3), v=noise((37, 61), 2, 0.3))
    f.update(steps=3)
    assert not np.array_equal(f.u, o.u)
