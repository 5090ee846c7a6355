"""The XCD-aware tile order (xcd_tile in fluid_device.hpp) restated: a
bijection of [0, n) for every n, giving each XCD (workgroup index mod 8) a
contiguous run of tiles."""
import pytest


def xcd_tile(b, n, xcds=8):
    x, k = b % xcds, b // xcds
    q, r = divmod(n, xcds)
    return x * q + min(x, r) + k


@pytest.mark.parametrize("n", list(range(1, 70)) + [450, 451, 457, 1000, 4097])
def test_bijection_and_contiguous_runs(n):
    tiles = [xcd_tile(b, n) for b in range(n)]
    assert sorted(tiles) == list(range(n))
    for x in range(8):
        run = [xcd_tile(b, n) for b in range(x, n, 8)]
        assert run == list(range(run[0], run[0] + len(run))) if run else True
