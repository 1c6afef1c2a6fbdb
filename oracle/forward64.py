"""Binary64 restatement of the production timestep, vectorised over particles.

TEST INFRASTRUCTURE ONLY (like the rest of oracle/): the finite-difference
reference of the adjoint tests.  The float32 oracle (oracle.c) and the GPU
kernels round every operation to float32, so central differences of a float32
run with small steps drown in rounding noise and large steps pick up the kinks
of the piecewise-linear interpolants; evaluated in binary64 the same model
resolves steps of 1e-4 grid cells, where both errors are far below 1e-3.

The step follows the same reference lines as oracle.c, in the same order:
advection (advection_dual_wind.wgsl:82-303: trilinear per time level, then the
temporal blend, Petterssen predictor-corrector, clamps), the PBL cell lookup
and compute_hanna (langevin_fused.wgsl:213-293), the horizontal and vertical
Ornstein-Uhlenbeck updates (:350-396) with the vertical sub-steps, PBL
reflection and hanna_short_w (:461-515), and scalar dry deposition
(dry_deposition.wgsl:70-106).  Random normals are the float32 oracle's Philox
blocks (rng.rs; langevin_fused.wgsl:430-451) turned into normals in binary64.
Wet deposition, gridding and compaction are not restated here.

Horizontal positions are continuous grid coordinates (x = cell + fraction);
vertical positions are metres (the level-height mode).
"""
from __future__ import annotations

import numpy as np

import oracle

F32 = lambda v: float(np.float32(v))  # noqa: E731  (the reference's float32 constants, exactly)

USTAR_MIN, TLU_MIN, TLV_MIN, TLW_MIN, SIGMA_FLOOR, H_MIN = 1e-4, 10.0, 10.0, 30.0, F32(1e-2), 1.0
E23, E13, E08, E13P = F32(2.0 / 3.0), F32(-1.0 / 3.0), F32(0.8), F32(1.0 / 3.0)


class Met64:
    """Two time levels of u, v, w (nx, ny, nz) in binary64 and the level heights."""

    def __init__(self, shape, u0, v0, w0, u1, v1, w1, level_heights):
        self.shape = tuple(shape)
        r = lambda a: np.asarray(a, np.float64).reshape(self.shape)  # noqa: E731
        self.f0 = [r(u0), r(v0), r(w0)]
        self.f1 = [r(u1), r(v1), r(w1)]
        self.lh = np.asarray(level_heights, np.float64)


def level_of(z, lh):
    """meters_to_grid_level (advection_dual_wind.wgsl:94-117)."""
    nz = len(lh)
    h0 = lh[0]
    k = np.searchsorted(lh[1:], z, side="left") + 1  # first k >= 1 with z <= lh[k]
    kc = np.minimum(k, nz - 1)
    hp = lh[kc - 1]
    den = lh[kc] - hp
    lev = np.where(den > 0, (kc - 1) + (z - hp) / np.where(den > 0, den, 1.0), kc - 1.0)
    lev = np.where(k >= nz, nz - 1.0, lev)
    below = (z / h0) if h0 > 0 else np.zeros_like(z)
    return np.where(z <= h0, below, lev)


def axis(c, n):
    """clamp_axis (advection_dual_wind.wgsl:130-138)."""
    c = np.clip(c, 0.0, n - 1.0)
    lo = np.floor(c)
    l = lo.astype(np.int64)
    return l, np.minimum(l + 1, n - 1), c - lo


def lerp(a, b, t):
    return a * (1.0 - t) + b * t


def sample(met, alpha, x, y, z):
    """sample_dual_u/v/w (advection_dual_wind.wgsl:142-236)."""
    nx, ny, nz = met.shape
    x0, x1, tx = axis(x, nx)
    y0, y1, ty = axis(y, ny)
    z0, z1, tz = axis(level_of(z, met.lh), nz)
    out = []
    for f0, f1 in zip(met.f0, met.f1):
        v = []
        for f in (f0, f1):
            a = lerp(lerp(f[x0, y0, z0], f[x1, y0, z0], tx), lerp(f[x0, y1, z0], f[x1, y1, z0], tx), ty)
            b = lerp(lerp(f[x0, y0, z1], f[x1, y0, z1], tx), lerp(f[x0, y1, z1], f[x1, y1, z1], tx), ty)
            v.append(lerp(a, b, tz))
        out.append(lerp(v[0], v[1], alpha))
    return out


def _finite(v):
    return np.isfinite(v)


def vertical(ust, wst, ol, h, z_in, neu, uns):
    """sigma_w, dsigma_w/dz, T_Lw at height z (hanna_short_w, :296-344)."""
    z = np.clip(z_in, 0.0, h)
    zeta = np.clip(z / h, 0.0, 1.0)
    with np.errstate(all="ignore"):
        corr = z / ust
        sw_n = 1.3 * ust * np.exp(-2e-4 * corr)
        pa_u = np.power(zeta, E23)
        sq = 1.2 * wst * wst * (1.0 - 0.9 * zeta) * pa_u + (1.8 - 1.4 * zeta) * ust * ust
        sw_u = np.sqrt(sq) + SIGMA_FLOOR
        ds_u = 0.5 / sw_u / h * (-1.4 * ust * ust + wst * wst *
                                 (0.8 * np.power(np.maximum(zeta, F32(1e-3)), E13) - 1.8 * pa_u))
        tl_a = 0.1 * z / (sw_u * (0.55 - 0.38 * np.abs(z / ol)))
        tl_b = 0.59 * z / sw_u
        tl_c = 0.15 * h / sw_u * (1.0 - np.exp(-5.0 * zeta))
        tl_u = np.where(z < np.abs(ol), tl_a, np.where(zeta < 0.1, tl_b, tl_c))
        sw_s = SIGMA_FLOOR + 1.3 * ust * (1.0 - zeta)
        tl_s = 0.1 * h / sw_s * np.power(zeta, E08)
        sigw = np.where(neu, sw_n + SIGMA_FLOOR, np.where(uns, sw_u, sw_s))
        dsw = np.where(neu, -2e-4 * sw_n, np.where(uns, ds_u, -1.3 * ust / h))
        tlw = np.where(neu, 0.5 * z / (sw_n + SIGMA_FLOOR) / (1.0 + 1.5e-3 * corr), np.where(uns, tl_u, tl_s))
    tlw = np.maximum(tlw, TLW_MIN)
    dsw = np.where(dsw == 0.0, 1e-10, dsw)
    sigw = np.where(~_finite(sigw) | (sigw <= 0.0), SIGMA_FLOOR, sigw)
    tlw = np.where(~_finite(tlw), TLW_MIN, tlw)
    dsw = np.where(~_finite(dsw), 1e-10, dsw)
    return sigw, dsw, tlw


def hanna(ust_in, wst_in, ol, h_in, z_in):
    """compute_hanna (langevin_fused.wgsl:224-293): the regime and all profiles."""
    ust = np.maximum(ust_in, USTAR_MIN)
    wst = np.maximum(wst_in, 0.0)
    h = np.maximum(h_in, H_MIN)
    with np.errstate(all="ignore"):
        neu = ~np.isfinite(ol) | (ol == 0.0) | (h / np.abs(ol) < 1.0)
        uns = ~neu & (ol < 0.0)
        z = np.clip(z_in, 0.0, h)
        zeta = np.clip(z / h, 0.0, 1.0)
        corr = z / ust
        sigu_n = SIGMA_FLOOR + 2.0 * ust * np.exp(-3e-4 * corr)
        sigu_u = SIGMA_FLOOR + ust * np.power(12.0 - 0.5 * h / ol, E13P)
        sigu_s = SIGMA_FLOOR + 2.0 * ust * (1.0 - zeta)
        sigv_s = SIGMA_FLOOR + 1.3 * ust * (1.0 - zeta)
        sigw, dsw, tlw = vertical(ust, wst, ol, h, z_in, neu, uns)
        sigw_raw_n = 1.3 * ust * np.exp(-2e-4 * corr) + SIGMA_FLOOR
        tl_n = 0.5 * z / sigw_raw_n / (1.0 + 1.5e-3 * corr)
        sigu = np.where(neu, sigu_n, np.where(uns, sigu_u, sigu_s))
        sigv = np.where(neu, sigw, np.where(uns, sigu_u, sigv_s))
        tlu = np.where(neu, tl_n, np.where(uns, 0.15 * h / sigu_u, 0.15 * h / sigu_s * np.sqrt(zeta)))
        tlv = np.where(neu, tl_n, np.where(uns, tlu, 0.467 * tlu))
    tlu = np.maximum(tlu, TLU_MIN)
    tlv = np.maximum(tlv, TLV_MIN)
    sigu = np.where(~_finite(sigu) | (sigu <= 0.0), SIGMA_FLOOR, sigu)
    sigv = np.where(~_finite(sigv) | (sigv <= 0.0), SIGMA_FLOOR, sigv)
    tlu = np.where(~_finite(tlu), TLU_MIN, tlu)
    tlv = np.where(~_finite(tlv), TLV_MIN, tlv)
    return dict(ust=ust, wst=wst, ol=ol, h=h, neu=neu, uns=uns, sigu=sigu, sigv=sigv, sigw=sigw, dsw=dsw,
                tlu=tlu, tlv=tlv, tlw=tlw)


def _pos(v):
    return np.where(_finite(v) & (v > 0.0), v, 0.0)


def ou_h(prev, sigma, tl, dt, eta):
    """update_horizontal_component (langevin_fused.wgsl:350-364)."""
    s, t = _pos(sigma), _pos(tl)
    with np.errstate(all="ignore"):
        r = dt / t
        e = np.exp(-r)
        nxt = np.where(r < 0.5, (1.0 - r) * prev + eta * s * np.sqrt(np.maximum(2.0 * r, 0.0)),
                       e * prev + eta * s * np.sqrt(np.maximum(1.0 - e * e, 0.0)))
    return np.where((t == 0.0) | ~_finite(nxt), 0.0, nxt)


def ou_v(prev, sigw, dsw, tlw, dt, rho, eta):
    """update_vertical_component (langevin_fused.wgsl:366-389)."""
    s, t = _pos(sigw), _pos(tlw)
    drift = s * (dsw + rho * s)
    with np.errstate(all="ignore"):
        r = dt / t
        e = np.exp(-r)
        nxt = np.where(r < 0.5, (1.0 - r) * prev + eta * s * np.sqrt(np.maximum(2.0 * r, 0.0)) + dt * drift,
                       e * prev + eta * s * np.sqrt(np.maximum(1.0 - e * e, 0.0)) + t * (1.0 - e) * drift)
    return np.where((t == 0.0) | ~_finite(nxt), 0.0, nxt)


def reflect(z, tw, cbt, zmin, h):
    """PBL reflection (langevin_fused.wgsl:481-500)."""
    on = h > zmin
    z, tw, cbt = z.copy(), tw.copy(), cbt.copy()
    for _ in range(10):
        lo = on & (z < zmin)
        hi = on & ~lo & (z > h)
        flip = lo | hi
        z = np.where(lo, 2.0 * zmin - z, np.where(hi, 2.0 * h - z, z))
        tw = np.where(flip, -tw, tw)
        cbt = np.where(flip, -cbt, cbt)
    z = np.where(on, np.clip(z, zmin, h), z)
    return z, tw, cbt


def normals(counter, key, n_slots):
    """The six normals of every slot for one step with 4 sub-steps: Philox
    blocks 2 s and 2 s + 1 (counter + 2 slot), Box-Muller in binary64 on the
    reference's float32 uniforms (box_muller_normals, :186-192)."""
    b = oracle.philox_stream(counter, key, 2 * n_slots).reshape(n_slots, 8).astype(np.uint64)
    u = (b >> np.uint64(8)).astype(np.float64) * 2.0 ** -24
    out = []
    for a, c in ((0, 1), (2, 3), (4, 5)):
        ru = np.clip(u[:, a], F32(1.1754944e-38), F32(0.99999994))
        au = np.clip(u[:, c], 0.0, F32(0.99999994))
        rad = np.sqrt(-2.0 * np.log(ru))
        th = F32(6.2831855) * au
        out += [rad * np.cos(th), rad * np.sin(th)]
    return out  # eta_u, eta_v, eta_w0, eta_w1, eta_w2, eta_w3


def step(state: dict, met: Met64, pbl: dict, prm: dict):
    """One production step of every active particle, in place.  state: x, y
    (grid coordinates), z [m], tu, tv, tw, cbt, mass, active; pbl: ustar,
    wstar, hmix, oli arrays of nx*ny (i*ny + j); prm: dt, alpha, x_scale,
    y_scale, key, counter, n_slots, dry_vd (0 = off), min_height, rho."""
    nx, ny, nz = met.shape
    act = state["active"]
    dt, alpha = prm["dt"], prm["alpha"]
    x0, y0, z0 = state["x"], state["y"], state["z"]
    tu, tv = state["tu"], state["tv"]
    zmax = met.lh[-1]
    s0 = sample(met, alpha, x0, y0, z0)
    u0, v0, w0 = s0[0] + tu, s0[1] + tv, s0[2]
    qx = np.clip(x0 + (u0 * prm["x_scale"]) * dt, 0.0, nx - 1.0)
    qy = np.clip(y0 + (v0 * prm["y_scale"]) * dt, 0.0, ny - 1.0)
    qz = np.clip(z0 + w0 * dt, 0.0, zmax)
    s1 = sample(met, alpha, qx, qy, qz)
    cu, cv, cw = 0.5 * (u0 + s1[0] + tu), 0.5 * (v0 + s1[1] + tv), 0.5 * (w0 + s1[2])
    x1 = np.clip(x0 + (cu * prm["x_scale"]) * dt, 0.0, nx - 1.0)
    y1 = np.clip(y0 + (cv * prm["y_scale"]) * dt, 0.0, ny - 1.0)
    z1 = np.clip(z0 + cw * dt, 0.0, zmax)
    i = np.floor(np.clip(x1, 0.0, nx - 1.0)).astype(np.int64)
    j = np.floor(np.clip(y1, 0.0, ny - 1.0)).astype(np.int64)
    c = i * ny + j
    oli = pbl["oli"][c]
    with np.errstate(divide="ignore"):
        ol = np.where(np.abs(oli) < F32(1e-5), np.inf, 1.0 / oli)
    hn = hanna(pbl["ustar"][c], pbl["wstar"][c], ol, pbl["hmix"][c], z1)
    eu, ev, *ew = [e[prm["slots"]] for e in normals(prm["counter"], prm["key"], prm["n_slots"])]
    tu1 = ou_h(tu, hn["sigu"], hn["tlu"], dt, eu)
    tv1 = ou_h(tv, hn["sigv"], hn["tlv"], dt, ev)
    tw, z, cbt = state["tw"], z1, state["cbt"]
    sigw, dsw, tlw = hn["sigw"], hn["dsw"], hn["tlw"]
    dts = dt / 4.0
    for sub in range(4):
        tw = ou_v(tw, sigw, dsw, tlw, dts, prm.get("rho", 0.0), ew[sub])
        z = z + tw * dts
        z, tw, cbt = reflect(z, tw, cbt, prm.get("min_height", F32(0.01)), hn["h"])
        if sub < 3:
            sigw, dsw, tlw = vertical(hn["ust"], hn["wst"], ol, hn["h"], z, hn["neu"], hn["uns"])
    mass = state["mass"]
    vd = prm.get("dry_vd", 0.0)
    if vd:
        href = max(prm.get("dry_href", 15.0), 0.1)
        sv = np.clip(np.exp(-max(vd, 0.0) * abs(dt) / (2.0 * href)), 0.0, 1.0)
        mass = np.where(z < 2.0 * href, mass * sv, mass)
    for k, v in (("x", x1), ("y", y1), ("z", z), ("tu", tu1), ("tv", tv1), ("tw", tw), ("cbt", cbt),
                 ("mass", mass)):
        state[k] = np.where(act, v, state[k])


def receptor_deposit(state, table, acc, shape, obs_nz, obs_dz):
    """Cloud-in-cell receptor sums of the observation lattice (met columns x
    obs_nz levels obs_dz apart): the adjoint engine's observation."""
    nx, ny = shape[0], shape[1]
    a = state["active"]
    x = np.clip(state["x"][a], 0.0, nx - 1.0)
    y = np.clip(state["y"][a], 0.0, ny - 1.0)
    z = np.clip(state["z"][a], 0.0, (obs_nz - 1) * obs_dz) / obs_dz
    m = state["mass"][a]
    i0 = np.minimum(np.floor(x).astype(np.int64), nx - 2)
    j0 = np.minimum(np.floor(y).astype(np.int64), ny - 2)
    l0 = np.minimum(np.floor(z).astype(np.int64), obs_nz - 2)
    fx, fy, fz = x - i0, y - j0, z - l0
    for k in range(8):
        w = (fx if k & 1 else 1.0 - fx) * (fy if k & 2 else 1.0 - fy) * (fz if k & 4 else 1.0 - fz)
        node = ((i0 + (k & 1)) * ny + j0 + ((k >> 1) & 1)) * obs_nz + l0 + (k >> 2)
        rec = table[node]
        ok = rec >= 0
        np.add.at(acc, rec[ok], m[ok] * w[ok])


def adjoint_forward(case, emission, position, coeff, nodes, obs, n_steps, obs_first, obs_nz, obs_dz,
                    counter, dry_vd=0.0):
    """The differentiable run of point sources in binary64: receptor values
    and the misfit sum of squares.  case: sources (counts), met levels, level
    heights, PBL, wind basis, dt, scales, key (the adjoint engine's inputs)."""
    nx, ny, nz = case.met.shape
    counts = np.asarray(case.counts, np.int64)
    src = np.repeat(np.arange(len(counts)), counts)
    pos = np.asarray(position, np.float64)
    q = np.asarray(emission, np.float64)
    n = len(src)
    state = dict(x=pos[src, 0].copy(), y=pos[src, 1].copy(), z=pos[src, 2].copy(), tu=np.zeros(n), tv=np.zeros(n),
                 tw=np.zeros(n), cbt=np.ones(n), mass=q[src] / counts[src], active=np.ones(n, bool))
    pu = np.asarray(coeff, np.float64) @ np.asarray(case.basis_u, np.float64)
    pv = np.asarray(coeff, np.float64) @ np.asarray(case.basis_v, np.float64)
    lv = []
    for (u, v, w) in case.met.levels:
        lv.append((np.asarray(u, np.float64).reshape(nx * ny, nz) + pu[:, None],
                   np.asarray(v, np.float64).reshape(nx * ny, nz) + pv[:, None], np.asarray(w, np.float64)))
    pbl = {k: np.asarray(getattr(case.pbl, k), np.float64) for k in ("ustar", "wstar", "hmix", "oli")}
    table = np.full(nx * ny * obs_nz, -1, np.int64)
    nd = np.asarray(nodes, np.int64)
    table[(nd[:, 0] * ny + nd[:, 1]) * obs_nz + nd[:, 2]] = np.arange(len(nd))
    vals = np.zeros(len(nd))
    T = case.met.seconds_between_levels
    ctr = [int(v) for v in counter]
    for k in range(n_steps):
        t = k * case.dt_seconds
        b = min(int(t // T), len(lv) - 2)
        met = Met64(case.met.shape, *lv[b], *lv[b + 1], case.met.level_heights)
        alpha = F32(min(max((t - b * T) / T, 0.0), 1.0))
        step(state, met, pbl, dict(dt=case.dt_seconds, alpha=alpha, x_scale=case.x_scale, y_scale=case.y_scale,
                                   key=case.key, counter=ctr, n_slots=n, slots=np.arange(n), dry_vd=dry_vd))
        ctr = [int(v) for v in oracle.counter_add(ctr, 2 * n)]
        if k + 1 >= obs_first:
            receptor_deposit(state, table, vals, case.met.shape, obs_nz, obs_dz)
    obs = np.zeros(len(nd)) if obs is None else np.asarray(obs, np.float64)
    return vals, float(np.sum((vals - obs) ** 2)), state
