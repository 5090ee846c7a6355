"""CPU emulation of the tiled projection kernel (k_project_tiled).

The emulation follows the kernel's index arithmetic: region origin,
wavefront/lane to (row, cell pair) mapping, de-interleaved LDS layout, the
per-phase box of cells that still reach the owned tile, and the write-back.
It asserts that every (row, pair) of the region is visited by exactly one
lane, that every LDS and global index is in bounds, and that every owned
cell is written once.  The result must equal the oracle's projection bit for
bit, for tiles smaller and larger than the grid, sweep counts that are not
multiples of k, one to eight wavefronts per row, and with pressure on.
"""
import numpy as np
import pytest

from oracle.oracle import OracleFluid
from tests.helpers import make_config, noise

G_UPD, G_L, G_R, G_B, G_T = 1, 2, 4, 8, 16
CF_FLUID, CF_UPD, CF_FU, CF_FV = 1, 2, 4, 8
WAVES = 8
LDS_LIMIT = 160 * 1024


def tuning_geometry(tw, th, k):
    rw, rh = tw + 4 * k, th + 4 * k
    pw = rw // 2
    n = -(-pw // 64)
    wpr = 1 if n <= 1 else 2 if n <= 2 else 4 if n <= 4 else 8
    return rw, rh, pw, wpr


def lds_bytes(tw, th, k, pressure):
    rw, rh, pw, _ = tuning_geometry(tw, th, k)
    cells = rh * 2 * (pw + 1)
    return cells * 4 * (3 if pressure else 2) + cells + (rh + 1) * (rw + 1)


def device_flags(o):
    """The flag bytes k_build_flags writes, from the oracle's masks, in the
    device orientation (row j bottom-up)."""
    H, W = o.solid.shape
    S = o.solid[::-1].astype(bool)
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


def lanes(pw, wpr):
    """(r0, rstep, q) of every lane of the workgroup that has a pair."""
    out = []
    for wave in range(WAVES):
        cb, r0, rstep = wave % wpr, wave // wpr, WAVES // wpr
        for lane in range(64):
            q = cb * 64 + lane
            if q < pw:
                out.append((r0, rstep, q))
    return out


def emulate_launch(cfg, flags, u_in, v_in, p, tw, th, K, sweeps):
    H, W = u_in.shape
    P = W  # the emulation keeps unpadded rows; the kernel's pitch only scales gk
    rw, rh, pw, wpr = tuning_geometry(tw, th, K)
    pwp, rs, fs = pw + 1, 2 * (pw + 1), rw + 1
    ncell = rh * rs
    ntx, nty = -(-W // tw), -(-H // th)
    f32 = np.float32
    o_ = f32(cfg.projection_o)
    pressure = bool(cfg.enable_pressure)
    u_out = np.full_like(u_in, np.nan)
    v_out = np.full_like(v_in, np.nan)
    p_out = p.copy()
    written = np.zeros((H, W), np.int32)
    L = lanes(pw, wpr)
    visits = np.zeros((rh, pw), np.int32)
    for (r0, rstep, q) in L:
        for r in range(r0, rh, rstep):
            visits[r, q] += 1
    assert np.all(visits == 1)
    for bid in range(ntx * nty):
        tx, ty = bid % ntx, bid // ntx
        x0, y0 = tx * tw, ty * th
        gx0, gy0 = x0 - 2 * K, y0 - 2 * K
        assert gx0 % 2 == 0
        U = np.full(ncell, np.nan, np.float32)
        V = np.full(ncell, np.nan, np.float32)
        S = np.full(ncell, np.nan, np.float32)
        G = np.zeros(ncell, np.uint8)
        F = np.zeros((rh + 1) * fs, np.uint8)
        Fset = np.zeros((rh + 1) * fs, bool)
        for (r0, rstep, q) in L:
            for r in range(r0, rh + 1, rstep):
                for b in range(2):
                    gi, gj = gx0 + 2 * q + b, gy0 + r
                    idx = r * fs + 2 * q + b
                    F[idx] = flags[gj, gi] if (0 <= gi < W and 0 <= gj < H) else 0
                    Fset[idx] = True
        for r in range(rh + 1):
            gi, gj = gx0 + rw, gy0 + r
            F[r * fs + rw] = flags[gj, gi] if (0 <= gi < W and 0 <= gj < H) else 0
            Fset[r * fs + rw] = True
        assert Fset.all()
        for (r0, rstep, q) in L:
            for r in range(r0, rh, rstep):
                gj, gi = gy0 + r, gx0 + 2 * q
                for b in range(2):
                    inside = 0 <= gj < H and 0 <= gi + b < W
                    idx = r * rs + b * pwp + q
                    assert idx < ncell
                    U[idx] = u_in[gj, gi + b] if inside else 0.0
                    V[idx] = v_in[gj, gi + b] if inside else 0.0
                    S[idx] = p[gj, gi + b] if inside else 0.0
        for (r0, rstep, q) in L:
            for r in range(r0, rh, rstep):
                for b in range(2):
                    lc = 2 * q + b
                    f, fr, ft = F[r * fs + lc], F[r * fs + lc + 1], F[(r + 1) * fs + lc]
                    g = 0
                    g |= G_UPD if f & CF_UPD else 0
                    g |= G_L if f & CF_FU else 0
                    g |= G_R if fr & CF_FU else 0
                    g |= G_B if f & CF_FV else 0
                    g |= G_T if ft & CF_FV else 0
                    g |= ((int(f) >> 4) & 7) << 5
                    G[r * rs + b * pwp + q] = g
        nphase = 2 * sweeps
        for s in range(nphase):
            c = s & 1
            e = nphase - 1 - s
            lo, chi, rhi = 2 * K - 1 - e, 2 * K + tw + e, 2 * K + th + e
            for (r0, rstep, q) in L:
                r = r0
                if lo > r0:
                    r += ((lo - r0 + rstep - 1) // rstep) * rstep
                while r < rhi:
                    b = (c + gy0 + r) & 1
                    lc = 2 * q + b
                    if lo <= lc < chi:
                        own = r * rs + b * pwp + q
                        g = int(G[own])
                        if g & G_UPD:
                            right = own - pwp + 1 if b else own + pwp
                            top = own + rs
                            assert 0 <= min(own, right, top) and max(own, right, top) < ncell
                            uu, ur, vv, vt = U[own], U[right], V[own], V[top]
                            assert not np.isnan([uu, ur, vv, vt]).any()
                            div = f32(f32(f32(ur - uu) + vt) - vv)
                            d = f32(o_ * f32(div / f32(g >> 5)))
                            if pressure:
                                t = f32(f32(f32(d * f32(cfg.density)) * f32(int(cfg.cell_size)))
                                        / f32(cfg.d_t))
                                S[own] = f32(S[own] + t)
                            if g & G_L:
                                U[own] = f32(uu + d)
                            if g & G_R:
                                U[right] = f32(ur - d)
                            if g & G_B:
                                V[own] = f32(vv + d)
                            if g & G_T:
                                V[top] = f32(vt - d)
                    r += rstep
        for (r0, rstep, q) in L:
            r = r0 + ((2 * K - r0 + rstep - 1) // rstep) * rstep if 2 * K > r0 else r0
            while r < 2 * K + th:
                gj = gy0 + r
                if gj >= H:
                    break
                lc = 2 * q
                gi = gx0 + lc
                for b in range(2):
                    inb = 2 * K <= lc + b < 2 * K + tw and gi + b < W
                    if inb:
                        idx = r * rs + b * pwp + q
                        u_out[gj, gi + b] = U[idx]
                        v_out[gj, gi + b] = V[idx]
                        if pressure:
                            p_out[gj, gi + b] = S[idx]
                        written[gj, gi + b] += 1
                r += rstep
    assert np.all(written == 1)
    assert P == W
    return u_out, v_out, p_out


def emulate_projection(cfg, o, tw, th, K):
    flags = device_flags(o)
    u, v, p = o.u[::-1].copy(), o.v[::-1].copy(), o.p[::-1].copy()
    done = 0
    while done < cfg.projection_n:
        sw = min(K, cfg.projection_n - done)
        u, v, p = emulate_launch(cfg, flags, u, v, p, tw, th, K, sw)
        done += sw
    return u[::-1], v[::-1], p[::-1]


@pytest.mark.parametrize("W,H,n,tw,th,K,pressure", [
    (40, 30, 7, 16, 8, 2, 0),     # several tiles, n not a multiple of k
    (40, 30, 4, 20, 12, 1, 1),    # k = 1, pressure on
    (26, 19, 5, 60, 40, 3, 0),    # one tile larger than the grid
    (33, 21, 6, 10, 6, 2, 1),     # odd width, tiles clipped at both edges
    (300, 12, 3, 240, 4, 2, 0),   # region 248 wide: two wavefronts per row
    (70, 9, 2, 2, 1, 1, 0),       # smallest tile
])
def test_tiled_emulation_matches_oracle(W, H, n, tw, th, K, pressure):
    cfg = make_config(W, H, projection_n=n, enable_pressure=pressure,
                      obstacle_center_x=W // 2 - 3, obstacle_center_y=H // 2 + 1,
                      obstacle_radius=min(4.5, H / 3))
    o = OracleFluid(cfg)
    o.set_state(u=noise((H, W), 21, 0.5), v=noise((H, W), 22, 0.5),
                p=noise((H, W), 23, 0.5) if pressure else None)
    u, v, p = emulate_projection(cfg, o, tw, th, K)
    o.projection()
    assert np.array_equal(u.view(np.uint32), o.u.view(np.uint32))
    assert np.array_equal(v.view(np.uint32), o.v.view(np.uint32))
    if pressure:
        assert np.array_equal(p.view(np.uint32), o.p.view(np.uint32))


def test_lane_mapping_covers_wide_regions():
    for tw, th, K in ((240, 32, 4), (1000, 8, 2), (2, 1, 1), (496, 4, 4)):
        rw, rh, pw, wpr = tuning_geometry(tw, th, K)
        assert pw <= 64 * WAVES
        seen = set()
        for (r0, rstep, q) in lanes(pw, wpr):
            for r in range(r0, rh, rstep):
                assert (r, q) not in seen
                seen.add((r, q))
        assert len(seen) == rh * pw


def test_default_shape_fits_lds():
    assert lds_bytes(240, 32, 4, False) <= LDS_LIMIT
    assert lds_bytes(240, 24, 4, True) <= LDS_LIMIT
