"""CPU emulation of the register-resident projection kernel (k_project_regs).

Eight wavefronts of 64 lanes; wavefront w holds region rows [w*R, w*R+R),
lane l cells 2l and 2l+1.  The emulation keeps the kernel's per-lane
registers as arrays over (wavefront, row, lane) and follows its steps: the
DPP lane shifts (lane l reads lane l+1, or lane l-1, with 0 past the ends),
the per-phase publication of each wavefront's row 0 through LDS, the update
sent back to the wavefront above after the phase, the per-row parity of the
active cell, the box of rows and columns that still reach the tile, and the
write-back of the owned tile.  Its result must equal the oracle's projection
bit for bit.
"""
import numpy as np
import pytest

from oracle.oracle import OracleFluid
from tests.helpers import make_config, noise
from tests.test_tiled_schedule import device_flags

G_L, G_R, G_B, G_T, G_UPD, G_SLOW = 1, 2, 4, 8, 16, 128
CF_UPD, CF_FU, CF_FV = 2, 4, 8
NWV = 8
f32 = np.float32


def gate_byte(f, fr, ft):
    f, fr, ft = int(f), int(fr), int(ft)
    if not f & CF_UPD:
        return 0
    g = G_UPD
    g |= G_L if f & CF_FU else 0
    g |= G_R if fr & CF_FU else 0
    g |= G_B if f & CF_FV else 0
    g |= G_T if ft & CF_FV else 0
    s = (f >> 4) & 7
    if s in (0, 3):
        g |= G_SLOW
    else:
        g |= (s >> 1) << 5
    return g


def from_right(x):
    out = np.zeros_like(x)
    out[..., :-1] = x[..., 1:]
    return out


def from_left(x):
    out = np.zeros_like(x)
    out[..., 1:] = x[..., :-1]
    return out


def quotient(div, g, o):
    l2 = (g >> 5) & 3
    y = (np.uint32(0x3F800000) - (l2.astype(np.uint32) << np.uint32(23))).view(np.float32)
    q = (div * y).astype(f32)
    slow = (g & G_SLOW) != 0
    if slow.any():
        s = np.array([bin(int(x) & 15).count("1") for x in g.ravel()], f32).reshape(g.shape)
        with np.errstate(divide="ignore", invalid="ignore"):
            q = np.where(slow, (div / s).astype(f32), q)
    return (f32(o) * q).astype(f32)


def gate_mask(g, bit):
    return ((g >> bit) & 1).astype(bool)


def emulate_launch(cfg, flags, u_in, v_in, p, tw, th, K, sweeps, R, mask=True):
    H, W = u_in.shape
    assert tw == 128 - 4 * K and th + 4 * K == NWV * R
    ntx, nty = -(-W // tw), -(-H // th)
    o = f32(cfg.projection_o)
    pressure = bool(cfg.enable_pressure)
    u_out = np.full_like(u_in, np.nan)
    v_out = np.full_like(v_in, np.nan)
    p_out = p.copy()
    written = np.zeros((H, W), np.int32)
    lanes = np.arange(64)
    for bid in range(ntx * nty):
        tx, ty = bid % ntx, bid // ntx
        x0, y0 = tx * tw, ty * th
        gx0, gy0 = x0 - 2 * K, y0 - 2 * K
        gi = gx0 + 2 * lanes
        u0 = np.zeros((NWV, R, 64), f32)
        u1 = np.zeros_like(u0)
        v0 = np.zeros_like(u0)
        v1 = np.zeros_like(u0)
        s0 = np.zeros_like(u0)
        s1 = np.zeros_like(u0)
        gw = np.zeros((NWV, R, 64), np.int64)

        def flag(i, j):
            return int(flags[j, i]) if (0 <= i < W and 0 <= j < H) else 0

        def val(a, i, j):
            return a[j, i] if (0 <= i < W and 0 <= j < H) else f32(0)

        for w in range(NWV):
            for rho in range(R):
                gj = gy0 + w * R + rho
                for l in range(64):
                    i0 = gi[l]
                    u0[w, rho, l], u1[w, rho, l] = val(u_in, i0, gj), val(u_in, i0 + 1, gj)
                    v0[w, rho, l], v1[w, rho, l] = val(v_in, i0, gj), val(v_in, i0 + 1, gj)
                    s0[w, rho, l], s1[w, rho, l] = val(p, i0, gj), val(p, i0 + 1, gj)
                    g0 = gate_byte(flag(i0, gj), flag(i0 + 1, gj), flag(i0, gj + 1))
                    g1 = gate_byte(flag(i0 + 1, gj), flag(i0 + 2, gj), flag(i0 + 1, gj + 1))
                    gw[w, rho, l] = g0 | (g1 << 8)
        nphase = 2 * sweeps
        for s in range(nphase):
            c = s & 1
            e = nphase - 1 - s
            lo, chi, rhi = 2 * K - 1 - e, 2 * K + tw + e, 2 * K + th + e
            lc = 2 * lanes
            m0 = np.where((lc >= lo) & (lc < chi), 0xFF, 0)
            m1 = np.where((lc + 1 >= lo) & (lc + 1 < chi), 0xFF, 0)
            if not mask:  # every lane of a wavefront that has a row in the box updates
                m0 = m1 = np.full(64, 0xFF)
            vrow = np.zeros((NWV, 2, 64), f32)
            for w in range(1, NWV):
                vrow[w, 0], vrow[w, 1] = v0[w, 0], v1[w, 0]
            dsend = np.zeros((NWV, 64), f32)
            pbs = []
            for w in range(NWV):
                vt0 = vrow[w + 1, 0] if w < NWV - 1 else np.zeros(64, f32)
                vt1 = vrow[w + 1, 1] if w < NWV - 1 else np.zeros(64, f32)
                pb = (c + gy0 + w * R) & 1
                pbs.append(pb)
                dtop = np.zeros(64, f32)
                wave_in_box = lo - w * R < R and rhi - w * R > 0
                for rho in range(R):
                    r = w * R + rho
                    if mask and (r < lo or r >= rhi):
                        continue
                    if not mask and not wave_in_box:
                        continue
                    b = (pb + rho) & 1
                    if b == 0:
                        g = gw[w, rho] & m0
                        uu, ur, vv = u0[w, rho].copy(), u1[w, rho].copy(), v0[w, rho].copy()
                        vt = v0[w, rho + 1].copy() if rho + 1 < R else vt0
                    else:
                        g = (gw[w, rho] >> 8) & m1
                        uu, ur, vv = u1[w, rho].copy(), from_right(u0[w, rho]), v1[w, rho].copy()
                        vt = v1[w, rho + 1].copy() if rho + 1 < R else vt1
                    div = (((ur - uu).astype(f32) + vt).astype(f32) - vv).astype(f32)
                    d = quotient(div, g, o)
                    if pressure:
                        t = ((((d * f32(cfg.density)).astype(f32) * f32(int(cfg.cell_size))).astype(f32))
                             / f32(cfg.d_t)).astype(f32)
                        tgt = s0 if b == 0 else s1
                        tgt[w, rho] = np.where(g & G_UPD, (tgt[w, rho] + t).astype(f32), tgt[w, rho])
                    neg0, pos0 = f32(-0.0), f32(0.0)
                    dl = np.where(gate_mask(g, 0), d, neg0)
                    dr = np.where(gate_mask(g, 1), d, pos0)
                    db = np.where(gate_mask(g, 2), d, neg0)
                    dt = np.where(gate_mask(g, 3), d, pos0)
                    if b == 0:
                        u0[w, rho] = (uu + dl).astype(f32)
                        u1[w, rho] = (ur - dr).astype(f32)
                        v0[w, rho] = (vv + db).astype(f32)
                        if rho + 1 < R:
                            v0[w, rho + 1] = (vt - dt).astype(f32)
                        else:
                            dtop = dt
                    else:
                        u1[w, rho] = (uu + dl).astype(f32)
                        u0[w, rho] = (u0[w, rho] - from_left(dr)).astype(f32)
                        v1[w, rho] = (vv + db).astype(f32)
                        if rho + 1 < R:
                            v1[w, rho + 1] = (vt - dt).astype(f32)
                        else:
                            dtop = dt
                if w < NWV - 1:
                    dsend[w] = dtop
            for w in range(1, NWV):
                if pbs[w] == 0:
                    v1[w, 0] = (v1[w, 0] - dsend[w - 1]).astype(f32)
                else:
                    v0[w, 0] = (v0[w, 0] - dsend[w - 1]).astype(f32)
        for w in range(NWV):
            for rho in range(R):
                r = w * R + rho
                gj = gy0 + r
                if r < 2 * K or r >= 2 * K + th or gj >= H:
                    continue
                for l in range(64):
                    for bb, (ua, va, sa) in enumerate(((u0, v0, s0), (u1, v1, s1))):
                        lcc = 2 * l + bb
                        i = gx0 + lcc
                        if 2 * K <= lcc < 2 * K + tw and i < W:
                            u_out[gj, i] = ua[w, rho, l]
                            v_out[gj, i] = va[w, rho, l]
                            if pressure:
                                p_out[gj, i] = sa[w, rho, l]
                            written[gj, i] += 1
    assert np.all(written == 1)
    return u_out, v_out, p_out


def emulate_projection(cfg, o, K, R, mask=True):
    flags = device_flags(o)
    tw, th = 128 - 4 * K, NWV * R - 4 * K
    u, v, p = o.u[::-1].copy(), o.v[::-1].copy(), o.p[::-1].copy()
    done = 0
    while done < cfg.projection_n:
        sw = min(K, cfg.projection_n - done)
        u, v, p = emulate_launch(cfg, flags, u, v, p, tw, th, K, sw, R, mask)
        done += sw
    return u[::-1], v[::-1], p[::-1]


@pytest.mark.parametrize("mask", [True, False])
@pytest.mark.parametrize("W,H,n,K,R,pressure", [
    (150, 40, 5, 2, 4, 0),     # two tiles across, several down, n not a multiple of k
    (130, 70, 4, 4, 8, 1),     # pressure on, k = 4
    (60, 25, 3, 3, 4, 0),      # grid smaller than one tile row
    (250, 90, 11, 5, 8, 0),    # the default shape, tiles on both edges
    (230, 150, 13, 6, 10, 0),  # 10 rows per wavefront
])
def test_regs_emulation_matches_oracle(W, H, n, K, R, pressure, mask):
    cfg = make_config(W, H, projection_n=n, enable_pressure=pressure,
                      obstacle_center_x=W // 2 + 1, obstacle_center_y=H // 2 - 2,
                      obstacle_radius=min(6.5, H / 4))
    o = OracleFluid(cfg)
    o.set_state(u=noise((H, W), 31, 0.5), v=noise((H, W), 32, 0.5),
                p=noise((H, W), 33, 0.5) if pressure else None)
    u, v, p = emulate_projection(cfg, o, K, R, mask)
    o.projection()
    assert np.array_equal(u.view(np.uint32), o.u.view(np.uint32))
    assert np.array_equal(v.view(np.uint32), o.v.view(np.uint32))
    if pressure:
        assert np.array_equal(p.view(np.uint32), o.p.view(np.uint32))
