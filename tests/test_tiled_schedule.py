"""CPU emulation of the tiled projection kernel (k_project_tiled).

The emulation follows the kernel's index arithmetic step by step: region
origin, de-interleaved LDS layout, per-thread items, the per-phase box of
cells that still reach the owned tile, and the write-back.  Every LDS and
global index is asserted in bounds.  The result must equal the oracle's
projection bit for bit, for tile shapes smaller and larger than the grid,
sweep counts that are not multiples of k, and with pressure on.
"""
import numpy as np
import pytest

from oracle.oracle import OracleFluid
from tests.helpers import make_config, noise

G_UPD, G_L, G_R, G_B, G_T = 1, 2, 4, 8, 16
CF_FLUID, CF_UPD, CF_FU, CF_FV = 1, 2, 4, 8
NT = 256


def device_flags(o):
    """The flag bytes k_build_flags writes, from the oracle's masks, in the
    device layout (row j = bottom-up)."""
    H, W = o.solid.shape
    S = o.solid[::-1].astype(bool)  # S[j, i]
    ts = o.total_s[::-1]
    f = np.zeros((H, W), np.uint8)
    for j in range(H):
        for i in range(W):
            if S[j, i]:
                v = 0
            else:
                v = CF_FLUID
                if 1 <= i <= W - 2 and 1 <= j <= H - 2:
                    v |= CF_UPD
                if i > 0 and not S[j, i - 1]:
                    v |= CF_FU
                if j > 0 and not S[j - 1, i]:
                    v |= CF_FV
            f[j, i] = v | (ts[j, i] << 4)
    return f


def emulate_launch(cfg, flags, u_in, v_in, p, pw, rh, K, sweeps):
    H, W = u_in.shape
    tw, th = 2 * pw - 4 * K, rh - 4 * K
    ntx, nty = -(-W // tw), -(-H // th)
    u_out = np.full_like(u_in, np.nan)
    v_out = np.full_like(v_in, np.nan)
    p_out = p.copy()
    pwp, rw = pw + 1, 2 * pw
    rs, fs = 2 * pwp, rw + 1
    f32 = np.float32
    o_ = f32(cfg.projection_o)
    for bid in range(ntx * nty):
        tx, ty = bid % ntx, bid // ntx
        x0, y0 = tx * tw, ty * th
        gx0, gy0 = x0 - 2 * K, y0 - 2 * K
        assert gx0 % 2 == 0
        U = np.full(rh * rs, np.nan, np.float32)
        V = np.full(rh * rs, np.nan, np.float32)
        F = np.zeros((rh + 1) * fs, np.uint8)
        for q in range((rh + 1) * fs):
            r, c = divmod(q, fs)
            gi, gj = gx0 + c, gy0 + r
            if 0 <= gi < W and 0 <= gj < H:
                F[q] = flags[gj, gi]
        nitems = pw * rh
        for q in range(nitems):
            r, pp = divmod(q, pw)
            for b in range(2):
                gi, gj = gx0 + 2 * pp + b, gy0 + r
                inside = 0 <= gi < W and 0 <= gj < H
                idx = r * rs + b * pwp + pp
                assert idx < rh * rs
                U[idx] = u_in[gj, gi] if inside else 0.0
                V[idx] = v_in[gj, gi] if inside else 0.0
        items = []
        for q in range(nitems):
            r, pp = divmod(q, pw)
            g2, pacc = 0, [0.0, 0.0]
            for b in range(2):
                lc = 2 * pp + b
                f, fr, ft = F[r * fs + lc], F[r * fs + lc + 1], F[(r + 1) * fs + lc]
                g = 0
                g |= G_UPD if f & CF_UPD else 0
                g |= G_L if f & CF_FU else 0
                g |= G_R if fr & CF_FU else 0
                g |= G_B if f & CF_FV else 0
                g |= G_T if ft & CF_FV else 0
                g |= ((int(f) >> 4) & 7) << 5
                g2 |= g << (8 * b)
                gi, gj = gx0 + lc, gy0 + r
                owned = (2 * K <= lc < 2 * K + tw and 2 * K <= r < 2 * K + th
                         and gi < W and gj < H)
                pacc[b] = p[gj, gi] if owned else f32(0)
            items.append([r, pp, g2, pacc])
        nphase = 2 * sweeps
        for s in range(nphase):
            c = s & 1
            e = nphase - 1 - s
            lo, chi, rhi = 2 * K - 1 - e, 2 * K + tw + e, 2 * K + th + e
            for it in items:
                r, pp, gg, pacc = it
                b = (c + gy0 + r) & 1
                lc = 2 * pp + b
                g = (gg >> (8 * b)) & 0xFF
                if not (g & G_UPD) or r < lo or r >= rhi or lc < lo or lc >= chi:
                    continue
                own = r * rs + b * pwp + pp
                right = own - pwp + 1 if b else own + pwp
                top = own + rs
                assert max(own, right, top) < rh * rs and min(own, right, top) >= 0
                uu, ur, vv, vt = U[own], U[right], V[own], V[top]
                assert not np.isnan([uu, ur, vv, vt]).any()
                div = f32(f32(f32(ur - uu) + vt) - vv)
                d = f32(o_ * f32(div / f32(g >> 5)))
                if cfg.enable_pressure:
                    t = f32(f32(f32(d * f32(cfg.density)) * f32(int(cfg.cell_size))) / f32(cfg.d_t))
                    pacc[b] = f32(pacc[b] + t)
                if g & G_L:
                    U[own] = f32(uu + d)
                if g & G_R:
                    U[right] = f32(ur - d)
                if g & G_B:
                    V[own] = f32(vv + d)
                if g & G_T:
                    V[top] = f32(vt - d)
        for r, pp, gg, pacc in items:
            if r < 2 * K or r >= 2 * K + th:
                continue
            gj = gy0 + r
            if gj >= H:
                continue
            for b in range(2):
                lc = 2 * pp + b
                gi = gx0 + lc
                if lc < 2 * K or lc >= 2 * K + tw or gi >= W:
                    continue
                assert np.isnan(u_out[gj, gi])  # each cell written once
                u_out[gj, gi] = U[r * rs + b * pwp + pp]
                v_out[gj, gi] = V[r * rs + b * pwp + pp]
                if cfg.enable_pressure:
                    p_out[gj, gi] = pacc[b]
    assert not np.isnan(u_out).any() and not np.isnan(v_out).any()
    return u_out, v_out, p_out


def emulate_projection(cfg, o, pw, rh, K):
    flags = device_flags(o)
    u, v, p = o.u[::-1].copy(), o.v[::-1].copy(), o.p[::-1].copy()  # device orientation
    done = 0
    while done < cfg.projection_n:
        sw = min(K, cfg.projection_n - done)
        u, v, p = emulate_launch(cfg, flags, u, v, p, pw, rh, K, sw)
        done += sw
    return u[::-1], v[::-1], p[::-1]


@pytest.mark.parametrize("W,H,n,pw,rh,K,pressure", [
    (40, 30, 7, 12, 16, 2, 0),    # several tiles, n not a multiple of k
    (40, 30, 4, 12, 16, 1, 1),    # k = 1, pressure on
    (26, 19, 5, 40, 40, 3, 0),    # one tile larger than the grid
    (33, 21, 6, 10, 14, 2, 1),    # odd width, tiles clipped at both edges
])
def test_tiled_emulation_matches_oracle(W, H, n, pw, rh, K, pressure):
    cfg = make_config(W, H, projection_n=n, enable_pressure=pressure,
                      obstacle_center_x=W // 2 - 3, obstacle_center_y=H // 2 + 1,
                      obstacle_radius=4.5)
    o = OracleFluid(cfg)
    o.set_state(u=noise((H, W), 21, 0.5), v=noise((H, W), 22, 0.5),
                p=noise((H, W), 23, 0.5) if pressure else None)
    u, v, p = emulate_projection(cfg, o, pw, rh, K)
    o.projection()
    assert np.array_equal(u.view(np.uint32), o.u.view(np.uint32))
    assert np.array_equal(v.view(np.uint32), o.v.view(np.uint32))
    if pressure:
        assert np.array_equal(p.view(np.uint32), o.p.view(np.uint32))


def test_items_fit_registers_and_lds_for_default_shape():
    pw, rh, K = 64, 64, 4
    assert -(-pw * rh // NT) <= 24
    lds = rh * 2 * (pw + 1) * 4 * 2 + (rh + 1) * (2 * pw + 1)
    assert lds <= 80 * 1024  # two workgroups per CU
