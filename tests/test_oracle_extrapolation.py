"""Pins the oracle's extrapolation (src/fluid.cu:720-747) against a literal
simulation of the reference kernel.

The reference launches (ceil(W/64), H) blocks of 64 x 1 threads; every
thread runs apply_extrapolation_at on its cell.  The simulation runs each
block's statements in lockstep (all active lanes read, then write, statement
by statement; the `if` arm before the `else` arm) and runs whole blocks one
after another in a given order.  Under every block order the result is the
same, and the oracle must equal it on every value.  (Only if two blocks
interleave statement by statement can row 1 copy v(1,1) into v(0,1) after
row 0 has zeroed v(0,1); the oracle keeps the serialised value 0 there, and
likewise at v(W-1,1).)"""
import itertools

import numpy as np
import pytest

from oracle.oracle import OracleFluid
from tests.helpers import cell, make_config, noise


def simulate(u0, v0, W, H, order):
    u = {(i, j): cell(u0, i, j) for i in range(W) for j in range(H)}
    v = {(i, j): cell(v0, i, j) for i in range(W) for j in range(H)}
    for (bx, j) in order:
        lanes = list(range(bx * 64, min(bx * 64 + 64, W)))

        def stmt(dst, tgt, src_of, active):
            vals = [(tgt(i), src_of(i)) for i in active]  # all lanes read ...
            for t, val in vals:                            # ... then write
                dst[t] = val

        if j == 0:
            stmt(u, lambda i: (i, 0), lambda i: u[(i, 1)], lanes)
            stmt(v, lambda i: (i, 1), lambda i: np.float32(0), lanes)
        elif j == H - 1:
            stmt(u, lambda i: (i, H - 1), lambda i: u[(i, H - 2)], lanes)
        first = [i for i in lanes if i == 0]
        last = [i for i in lanes if i != 0 and i == W - 1]
        stmt(v, lambda i: (0, j), lambda i: v[(1, j)], first)
        stmt(u, lambda i: (1, j), lambda i: np.float32(0), first)
        stmt(v, lambda i: (W - 1, j), lambda i: v[(W - 2, j)], last)
    return u, v


def to_arrays(d, W, H):
    a = np.zeros((H, W), np.float32)
    for (i, j), val in d.items():
        a[H - 1 - j, i] = val
    return a


@pytest.mark.parametrize("W,H", [(7, 5), (70, 4)])
def test_extrapolation_matches_every_block_order(W, H):
    c = make_config(W, H)
    o = OracleFluid(c)
    u0, v0 = noise((H, W), 11, 1.0), noise((H, W), 12, 1.0)
    o.set_state(u=u0, v=v0)
    o.extrapolation()
    nbx = (W + 63) // 64
    blocks = [(bx, j) for j in range(H) for bx in range(nbx)]
    if len(blocks) <= 6:
        orders = [list(p) for p in itertools.permutations(blocks)]
    else:
        rng = np.random.Generator(np.random.PCG64(3))
        orders = [blocks, blocks[::-1]]
        orders += [[blocks[k] for k in rng.permutation(len(blocks))] for _ in range(60)]
    for order in orders:
        su, sv = simulate(u0, v0, W, H, order)
        assert np.array_equal(o.u, to_arrays(su, W, H))
        assert np.array_equal(o.v, to_arrays(sv, W, H))


def test_extrapolation_net_effect():
    W, H = 12, 9
    o = OracleFluid(make_config(W, H))
    o.set_state(u=noise((H, W), 13, 1.0), v=noise((H, W), 14, 1.0))
    o.extrapolation()
    assert all(cell(o.u, 1, j) == 0 for j in range(H))
    assert all(cell(o.v, i, 1) == 0 for i in range(W))
    assert cell(o.v, 0, 1) == 0 and cell(o.v, W - 1, 1) == 0
