"""The domain of the smoke sampler's staged square root (sqrt_staged).

The staged sampler takes the distances from a sample position to the
centres of the four cells around it.  Cell centres are (i + 1/2) * cs, so
their magnitude is at least 1/2 * cs; a nonzero difference between a
position and a centre near it is then at least 2^-25 (the spacing of
binary32 numbers from 1/4 up), and the squared distance dx*dx + dy*dy,
computed in binary32 as the sampler does, is either 0 or at least 2^-50.
sqrt_staged is exact on 0 and on every input from 2^-96 up (exhaustive
device check, osy_arith_check mode 7).  This test walks positions at and
a few ulps around centres, for small and large cell indices and several
cell sizes, and checks the bound.
"""
import numpy as np
import pytest

f32 = np.float32


def centre(i, cs):
    # centre<CS1>: (float)i + 0.5 for cs == 1, else ((double)i + 0.5) * cs in
    # double rounded to float
    if cs == 1:
        return f32(f32(i) + f32(0.5))
    return f32((float(i) + 0.5) * cs)


def around(v, k=4):
    """v and the k binary32 neighbours on each side of it."""
    out = [v]
    up = dn = v
    for _ in range(k):
        up = np.nextafter(up, f32(np.inf), dtype=f32)
        dn = np.nextafter(dn, f32(-np.inf), dtype=f32)
        out += [up, dn]
    return np.array(out, f32)


@pytest.mark.parametrize("cs", [1, 2, 3, 8])
def test_staged_squared_distance_is_zero_or_normal(cs):
    bound = f32(2.0 ** -50)
    worst = f32(np.inf)
    for i in [-1, 0, 1, 2, 7, 63, 64, 511, 1023, 1919, 2047, 4095, 16383]:
        c = centre(i, cs)
        for j in [-1, 0, 1, 100, 1079, 4095]:
            cy = centre(j, cs)
            # positions at and near the centres, and at and near the centres
            # of the neighbouring cells the sampler also weighs
            xs = np.concatenate([around(c), around(centre(i + 1, cs)), around(centre(i - 1, cs))])
            ys = np.concatenate([around(cy), around(centre(j + 1, cs))])
            dx = (xs[:, None] - c).astype(f32)
            dy = (ys[None, :] - cy).astype(f32)
            t = ((dx * dx).astype(f32) + (dy * dy).astype(f32)).astype(f32)
            nz = t[t != 0]
            assert np.all(nz >= bound), (cs, i, j, nz.min())
            if nz.size:
                worst = min(worst, nz.min())
    # with unit cells the bound is reached within a factor of 16 (centres
    # at 1/2, neighbours 2^-25 apart): the argument is tight
    if cs == 1:
        assert worst < bound * 16
