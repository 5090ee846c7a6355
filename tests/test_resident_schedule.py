"""CPU emulation of the resident projection's stage schedule (k_project_resident).

Every tile keeps its region (the owned tile plus a 2k-cell halo) for the
whole projection and runs the sweeps in stages of k.  Between stages it
writes only its ring (owned cells within 2k of the tile's edge) to the
buffer of the next stage and reloads only its halo from the rings its
neighbours wrote.  The emulation sweeps each region on its own: faces
past the region read 0, as the kernel's lane shifts and its last row give,
and gates come from the grid's flags.  Whatever the exchange buffers hold
outside the rings is poisoned with NaN before every stage, so a halo cell
that no ring covers shows up.  The result must equal the oracle's
projection bit for bit.  With a ring one cell too narrow it must not.
"""
import numpy as np
import pytest

from oracle.oracle import OracleFluid
from tests.helpers import make_config, noise
from tests.test_tiled_schedule import device_flags

CF_UPD, CF_FU, CF_FV = 2, 4, 8
f32 = np.float32


def sweep_region(cfg, F, U, V, S, gy0, gx0, sweeps):
    """`sweeps` red-black sweeps over one region in place (even colour
    first); F holds the grid's flags for the region plus one column right
    and one row above (0 outside the grid)."""
    rh, rw = U.shape
    o, dens, cs, dt = f32(cfg.projection_o), f32(cfg.density), f32(int(cfg.cell_size)), f32(cfg.d_t)
    rr, qq = np.mgrid[0:rh, 0:rw]
    for ph in range(2 * sweeps):
        colour = ph & 1
        f = F[:rh, :rw]
        act = ((f & CF_UPD) != 0) & (((gy0 + rr + gx0 + qq) & 1) == colour)
        r, q = rr[act], qq[act]
        fc = f[r, q].astype(np.int64)
        uu, vv = U[r, q], V[r, q]
        ur = np.where(q + 1 < rw, U[r, np.minimum(q + 1, rw - 1)], f32(0))
        vt = np.where(r + 1 < rh, V[np.minimum(r + 1, rh - 1), q], f32(0))
        div = (((ur - uu).astype(f32) + vt).astype(f32) - vv).astype(f32)
        s = ((fc >> 4) & 7).astype(f32)
        with np.errstate(divide="ignore", invalid="ignore"):
            vd = (o * (div / s).astype(f32)).astype(f32)
        if cfg.enable_pressure:
            with np.errstate(invalid="ignore", over="ignore"):
                S[r, q] = (S[r, q] + (((vd * dens).astype(f32) * cs).astype(f32) / dt).astype(f32)).astype(f32)
        gl = (fc & CF_FU) != 0
        gr = (F[r, q + 1] & CF_FU) != 0
        gb = (fc & CF_FV) != 0
        gt = (F[r + 1, q] & CF_FV) != 0
        U[r, q] = np.where(gl, (uu + vd).astype(f32), uu)
        inr = gr & (q + 1 < rw)
        U[r[inr], q[inr] + 1] = (ur[inr] - vd[inr]).astype(f32)
        V[r, q] = np.where(gb, (vv + vd).astype(f32), vv)
        int_ = gt & (r + 1 < rh)
        V[r[int_] + 1, q[int_]] = (vt[int_] - vd[int_]).astype(f32)


def emulate_resident(cfg, flags, u, v, p, tw, th, K, ring=None):
    """The resident projection of cfg.projection_n sweeps; `ring` is the
    width of the exchanged ring (2K in the kernel)."""
    H, W = u.shape
    ring = 2 * K if ring is None else ring
    n = cfg.projection_n
    nst = -(-n // K)
    ntx, nty = -(-W // tw), -(-H // th)
    rh, rw = th + 4 * K, tw + 4 * K
    bufs = [(u.copy(), v.copy()), (np.full_like(u, np.nan), np.full_like(v, np.nan))]
    tiles = []
    for ty in range(nty):
        for tx in range(ntx):
            gx0, gy0 = tx * tw - 2 * K, ty * th - 2 * K
            jj, ii = np.mgrid[gy0:gy0 + rh + 1, gx0:gx0 + rw + 1]
            inside = (ii >= 0) & (ii < W) & (jj >= 0) & (jj < H)
            F = np.where(inside, flags[np.clip(jj, 0, H - 1), np.clip(ii, 0, W - 1)], 0).astype(np.int64)
            ins = inside[:rh, :rw]
            jc, ic = np.clip(jj[:rh, :rw], 0, H - 1), np.clip(ii[:rh, :rw], 0, W - 1)
            U = np.where(ins, u[jc, ic], f32(0)).astype(f32)
            V = np.where(ins, v[jc, ic], f32(0)).astype(f32)
            S = np.where(ins, p[jc, ic], f32(0)).astype(f32)
            own = np.zeros((rh, rw), bool)
            own[2 * K:2 * K + th, 2 * K:2 * K + tw] = True
            own &= ins
            orow, ocol = np.mgrid[0:rh, 0:rw]
            orow, ocol = orow - 2 * K, ocol - 2 * K
            in_ring = own & ((orow < ring) | (orow >= th - ring) | (ocol < ring) | (ocol >= tw - ring))
            tiles.append(dict(gx0=gx0, gy0=gy0, F=F, U=U, V=V, S=S, ins=ins, jc=jc, ic=ic,
                              own=own, ring=in_ring))
    u_out, v_out, p_out = np.full_like(u, np.nan), np.full_like(v, np.nan), p.copy()
    for st in range(nst):
        sweeps = K if st + 1 < nst else n - (nst - 1) * K
        if st > 0:  # reload the halo from the rings of the previous stage
            bu, bv = bufs[st & 1]
            for t in tiles:
                halo = ~t["own"]
                hu = np.where(t["ins"], bu[t["jc"], t["ic"]], f32(0)).astype(f32)
                hv = np.where(t["ins"], bv[t["jc"], t["ic"]], f32(0)).astype(f32)
                t["U"][halo], t["V"][halo] = hu[halo], hv[halo]
        for t in tiles:
            sweep_region(cfg, t["F"], t["U"], t["V"], t["S"], t["gy0"], t["gx0"], sweeps)
        if st + 1 < nst:  # rings into the next stage's buffer, the rest poisoned
            bu, bv = bufs[(st + 1) & 1]
            bu.fill(np.nan)
            bv.fill(np.nan)
            for t in tiles:
                m = t["ring"]
                bu[t["jc"][m], t["ic"][m]] = t["U"][m]
                bv[t["jc"][m], t["ic"][m]] = t["V"][m]
        else:
            for t in tiles:
                m = t["own"]
                u_out[t["jc"][m], t["ic"][m]] = t["U"][m]
                v_out[t["jc"][m], t["ic"][m]] = t["V"][m]
                p_out[t["jc"][m], t["ic"][m]] = t["S"][m]
    return u_out, v_out, p_out


CASES = [
    (150, 100, 9, 2, 4, 0),    # 2 x 5 tiles of 120 x 24, stages of 2 and a last of 1
    (250, 130, 13, 5, 8, 1),   # 108 x 44 tiles (the 1920x1080 shape), pressure on
    (300, 170, 10, 1, 4, 0),   # ten one-sweep stages
    (60, 25, 7, 3, 4, 0),      # a grid smaller than one tile
]


def run_case(W, H, n, K, R, pressure, ring=None):
    cfg = make_config(W, H, projection_n=n, enable_pressure=pressure,
                      obstacle_center_x=W // 2 + 1, obstacle_center_y=H // 2 - 2,
                      obstacle_radius=min(6.5, H / 4))
    o = OracleFluid(cfg)
    o.set_state(u=noise((H, W), 41, 0.5), v=noise((H, W), 42, 0.5),
                p=noise((H, W), 43, 0.5) if pressure else None)
    flags = device_flags(o)
    tw, th = 128 - 4 * K, 8 * R - 4 * K
    u, v, p = emulate_resident(cfg, flags, o.u[::-1].copy(), o.v[::-1].copy(),
                               o.p[::-1].copy(), tw, th, K, ring)
    o.projection()
    same = (np.array_equal(u[::-1].view(np.uint32), o.u.view(np.uint32)) and
            np.array_equal(v[::-1].view(np.uint32), o.v.view(np.uint32)))
    if pressure:
        same = same and np.array_equal(p[::-1].view(np.uint32), o.p.view(np.uint32))
    return same


@pytest.mark.parametrize("W,H,n,K,R,pressure", CASES)
def test_resident_schedule_matches_oracle(W, H, n, K, R, pressure):
    assert run_case(W, H, n, K, R, pressure)


@pytest.mark.parametrize("W,H,n,K,R,pressure", CASES[:2])
def test_narrower_ring_is_caught(W, H, n, K, R, pressure):
    """A planted mistake: rings of 2k - 1 cells leave halo cells that no
    ring covers, and the emulation no longer matches the oracle."""
    assert not run_case(W, H, n, K, R, pressure, ring=2 * K - 1)
