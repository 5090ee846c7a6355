"""Pins the oracle's red-black projection (src/fluid.cu:221-295) with closed
forms for single-cell updates, SOR convergence, and order invariances."""
import numpy as np
import pytest

from oracle.oracle import OracleFluid
from tests.helpers import cell, make_config, noise, set_cell


def div_at(o, i, j):
    return ((cell(o.u, i + 1, j) - cell(o.u, i, j)) + cell(o.v, i, j + 1)) - cell(o.v, i, j)


def box(W=8, H=8, **kw):
    """Closed box: every border solid, no obstacle."""
    return make_config(W, H, obstacle_enable=0, enable_drain=0, **kw)


def test_zero_sweeps_is_identity():
    c = box(20, 14)
    o = OracleFluid(c)
    u0, v0 = noise((14, 20), 1), noise((14, 20), 2)
    o.set_state(u=u0, v=v0)
    o.projection(sweeps=0)
    assert np.array_equal(o.u, u0) and np.array_equal(o.v, v0)


def test_single_cell_update_with_o_1_zeroes_its_divergence():
    """Cell (3,3) is even, interior, s = 4.  With o = 1, d = div/4 is added
    to u(3,3) and v(3,3) and taken from u(4,3) and v(3,4); the cell's
    divergence becomes exactly 0 for dyadic data."""
    c = box(projection_o=1.0)
    o = OracleFluid(c)
    set_cell(o.u, 4, 3, 1.0)
    o.colour_phase(0)
    assert cell(o.u, 3, 3) == 0.25 and cell(o.u, 4, 3) == 0.75
    assert cell(o.v, 3, 3) == 0.25 and cell(o.v, 3, 4) == -0.25
    assert div_at(o, 3, 3) == 0.0
    # the odd phase then acts on the neighbours, which now have divergence
    before = {(i, j): div_at(o, i, j) for (i, j) in ((2, 3), (4, 3), (3, 2), (3, 4))}
    assert all(d != 0 for d in before.values())
    o.colour_phase(1)
    for (i, j) in before:
        assert abs(div_at(o, i, j)) < 1e-7


def test_overrelaxation_scales_residual():
    """With over-relaxation o and s = 4 the updated cell keeps (1 - o) of
    its divergence."""
    for om in (0.5, 1.5, 1.9):
        c = box(projection_o=om)
        o = OracleFluid(c)
        set_cell(o.u, 4, 3, 1.0)
        o.colour_phase(0)
        assert div_at(o, 3, 3) == pytest.approx(1.0 - np.float32(om), abs=1e-6)


def test_solid_neighbour_gates_its_face_and_s_is_3():
    """Cell (1,4) touches the left wall: s = 3, u(1,4) (the wall face) does
    not move, the other three faces move by d = div / 3."""
    c = box(10, 10, projection_o=1.0)
    o = OracleFluid(c)
    assert cell(o.total_s, 1, 4) == 3
    set_cell(o.u, 2, 4, 0.75)  # (1,4): 1+4 odd -> odd colour
    set_cell(o.u, 1, 4, 0.5)
    o.colour_phase(1)
    d = np.float32(np.float32(0.75) - np.float32(0.5)) / np.float32(3)
    assert cell(o.u, 1, 4) == np.float32(0.5)
    assert cell(o.u, 2, 4) == np.float32(0.75) - d
    assert cell(o.v, 1, 4) == d and cell(o.v, 1, 5) == -d


def test_pressure_accumulates_d_rho_cs_over_dt():
    c = box(projection_o=1.0, enable_pressure=1, density=2.0, d_t=0.05)
    o = OracleFluid(c)
    set_cell(o.u, 4, 3, 1.0)
    o.colour_phase(0)
    assert cell(o.p, 3, 3) == pytest.approx(0.25 * 2.0 / 0.05, rel=1e-6)
    assert np.count_nonzero(o.p) == 1


@pytest.mark.parametrize("omega,factor", [(1.0, 5e-3), (1.5, 1e-4), (1.9, 1e-7)])
def test_sor_converges_in_fp64(omega, factor):
    """In a closed box SOR drives the divergence towards 0 for o in (0, 2);
    a sign or index error in the update makes it grow instead.  Faster
    with stronger over-relaxation, as SOR theory predicts for this grid."""
    c = box(24, 18, projection_o=omega)
    o = OracleFluid(c, np.float64)
    # noise on fluid-fluid faces only: wall faces never move, so they must
    # carry no net flux for a divergence-free field to exist
    fl = o.solid == 0
    fu = np.zeros_like(fl)
    fu[:, 1:] = fl[:, 1:] & fl[:, :-1]
    fv = np.zeros_like(fl)
    fv[:-1, :] = fl[:-1, :] & fl[1:, :]
    o.set_state(u=noise((18, 24), 4, 1.0, np.float64) * fu,
                v=noise((18, 24), 5, 1.0, np.float64) * fv)
    res = [o.divergence_norm()[0]]
    for _ in range(3):
        o.projection(sweeps=60)
        res.append(o.divergence_norm()[0])
    assert all(b < a for a, b in zip(res, res[1:]))
    assert res[-1] < factor * res[0]


def py_colour_phase(o, colour, order_seed):
    """apply_projection_at (src/fluid.cu:229-262) over one colour, cells in a
    random order, float32 arithmetic."""
    W, H = o.cfg.width, o.cfg.height
    f32 = np.float32
    cells = [(i, j) for j in range(1, H - 1) for i in range(1, W - 1) if (i + j) % 2 == colour]
    rng = np.random.Generator(np.random.PCG64(order_seed))
    rng.shuffle(cells)
    S = lambda i, j: cell(o.solid, i, j)
    for (i, j) in cells:
        if S(i, j):
            continue
        uu, vv = cell(o.u, i, j), cell(o.v, i, j)
        top_v, right_u = cell(o.v, i, j + 1), cell(o.u, i + 1, j)
        div = f32(f32(f32(right_u - uu) + top_v) - vv)
        vd = f32(f32(o.cfg.projection_o) * f32(div / f32(cell(o.total_s, i, j))))
        if not S(i - 1, j):
            set_cell(o.u, i, j, f32(cell(o.u, i, j) + vd))
        if not S(i + 1, j):
            set_cell(o.u, i + 1, j, f32(cell(o.u, i + 1, j) - vd))
        if not S(i, j - 1):
            set_cell(o.v, i, j, f32(cell(o.v, i, j) + vd))
        if not S(i, j + 1):
            set_cell(o.v, i, j + 1, f32(cell(o.v, i, j + 1) - vd))


def test_colour_phase_is_order_independent():
    """Same-colour cells share no face: any order within a phase gives the
    same bits as the oracle's row order."""
    c = make_config(23, 17, obstacle_center_x=11, obstacle_center_y=8, obstacle_radius=3.5)
    a, b = OracleFluid(c), OracleFluid(c)
    u0, v0 = noise((17, 23), 6, 0.5), noise((17, 23), 7, 0.5)
    a.set_state(u=u0, v=v0)
    b.set_state(u=u0, v=v0)
    for colour, seed in ((0, 1), (1, 2), (0, 3)):
        a.colour_phase(colour)
        py_colour_phase(b, colour, seed)
        assert np.array_equal(a.u, b.u) and np.array_equal(a.v, b.v)


def test_divergence_norm_closed_form():
    """One nonzero face u(i,j) = a gives divergence -a in cell (i,j) and +a
    in cell (i-1,j): L2 = sqrt(2)|a|, max |div| = |a|; a v face likewise."""
    c = box(12, 10)
    o = OracleFluid(c, np.float64)
    set_cell(o.u, 5, 4, 0.75)
    l2, mx = o.divergence_norm()
    assert l2 == pytest.approx(np.sqrt(2) * 0.75, rel=1e-15) and mx == 0.75
    set_cell(o.v, 2, 6, -2.0)
    l2, mx = o.divergence_norm()
    assert l2 == pytest.approx(np.sqrt(2 * 0.75 ** 2 + 2 * 4.0), rel=1e-15) and mx == 2.0


def test_pressure_range_is_min_and_max_over_all_cells():
    """thrust::reduce min/max over all W*H pressures (solid cells hold 0)."""
    c = box(14, 9, enable_pressure=1, projection_n=3, projection_o=1.5)
    o = OracleFluid(c)
    o.set_state(u=noise((9, 14), 41, 1.0), v=noise((9, 14), 42, 1.0))
    o.update()
    lo, hi = o.pressure_range
    assert lo == o.p.min() and hi == o.p.max()
    assert lo < 0 < hi


def test_thread_count_invariance():
    c = make_config(97, 61, projection_n=30)
    outs = []
    for nt in (1, 3, 8):
        o = OracleFluid(c, nthreads=nt)
        o.set_state(u=noise((61, 97), 8), v=noise((61, 97), 9))
        o.projection()
        outs.append((o.u.copy(), o.v.copy()))
    for u, v in outs[1:]:
        assert np.array_equal(u, outs[0][0]) and np.array_equal(v, outs[0][1])
