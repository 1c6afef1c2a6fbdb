"""Deterministic synthetic inputs for the per-particle timestep.

The reference builds its test and benchmark met fields synthetically
(wind/synthetic.rs: uniform_wind_field, linear_shear_wind_field) and its
benchmark scenario in src/bin/fortran-validation.rs.  This module generates
the two workloads the timestep is measured and checked on, from seeds only,
so that the GPU run and the CPU oracle see identical inputs:

* ``point_release`` - 100,000 particles, one point release at the grid centre
  100 m above ground emitting 1 kg uniformly over the first hour, 24 h at a
  900 s step, a 64 x 64 (1 deg) x 32-level met grid with u = 10 m/s +
  2e-3 s^-1 * z, v = 2 m/s, w = 0, PBL height 1000 m, u* 0.4 m/s,
  w* 1.5 m/s, L = -200 m, dry deposition v_d = 2e-3 m/s, Philox seed 42,
  output 64 x 64 with levels 100/500/1000/3000 m.

* ``global_10m`` - 10,000,000 particles from 16 point sources at
  uniform-random positions with heights U[10, 500] m and lognormal emissions
  (median 1 kg, sigma 1), 24 h at a 600 s step on a 360 x 181 (1 deg global)
  x 60-level met grid with time levels 6 h apart; winds are a zonal jet
  u = 30 sin(2 lat) m/s plus random Fourier modes (wavenumbers <= 8, rms 5
  m/s), w from continuity; hmix lognormal (median 800 m, sigma 0.5), 1/L
  U[-0.02, 0.02] 1/m, u* U[0.1, 0.8] m/s; output 720 x 361 x 5 levels.

Level heights follow a stretched table z_k = z_top * (k/(nz-1))^2 + 10 m,
(strictly increasing, 10 m at the surface level).  w* is 1.5 m/s *
(hmix/1000 m)^(1/3) where 1/L < 0 (convective) and 0 elsewhere.  The
particles of a source are released at t = 0 at the source point (the release
manager's time window is host-side work outside the timestep).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

EARTH_RADIUS_M = 6.371e6
METRES_PER_DEG = EARTH_RADIUS_M * math.pi / 180.0


def level_heights(nz: int, z_top: float) -> np.ndarray:
    k = np.arange(nz, dtype=np.float64)
    return (z_top * (k / (nz - 1)) ** 2 + 10.0).astype(np.float32)


@dataclass
class Met:
    shape: tuple
    levels: list            # list of (u, v, w) float32 arrays, flat ((ix*ny)+iy)*nz+iz
    level_heights: np.ndarray
    seconds_between_levels: float


@dataclass
class Pbl:
    shape: tuple
    ustar: np.ndarray
    wstar: np.ndarray
    hmix: np.ndarray
    oli: np.ndarray
    precip: np.ndarray | None = None


@dataclass
class Workload:
    name: str
    met: Met
    pbl: Pbl
    particles: np.ndarray     # PARTICLE_DTYPE, all slots (capacity)
    dt_seconds: float
    n_steps: int
    x_scale: float
    y_scale: float
    key: tuple
    out_shape: tuple
    outheights: tuple
    refine: tuple
    mass_scale: float
    dry_vd: float = 0.0
    release_seconds: float = 0.0  # emit uniformly over [0, release_seconds]
    extra: dict = field(default_factory=dict)

    def alpha(self, t_seconds: float) -> tuple:
        """(bracket index, alpha) for simulation time t."""
        dtm = self.met.seconds_between_levels
        b = min(int(t_seconds // dtm), len(self.met.levels) - 2)
        return b, float((t_seconds - b * dtm) / dtm)


def make_particles(n: int, x, y, z, mass, capacity: int | None = None) -> np.ndarray:
    """Active particles at grid coordinates (x, y) [grid units], z [m]."""
    from . import PARTICLE_DTYPE as dt
    cap = n if capacity is None else capacity
    p = np.zeros(cap, dt)
    x = np.broadcast_to(np.asarray(x, np.float32), (n,))
    y = np.broadcast_to(np.asarray(y, np.float32), (n,))
    cx = np.floor(x).astype(np.int32)
    cy = np.floor(y).astype(np.int32)
    p["cell_x"][:n] = cx
    p["cell_y"][:n] = cy
    p["pos_x"][:n] = (x - cx).astype(np.float32)
    p["pos_y"][:n] = (y - cy).astype(np.float32)
    p["pos_z"][:n] = np.broadcast_to(np.asarray(z, np.float32), (n,))
    p["mass"][:n, 0] = np.broadcast_to(np.asarray(mass, np.float32), (n,))
    p["flags"][:n] = 1
    p["cbt"][:n] = 1
    return p


def point_release(n_particles: int = 100_000, seed: int = 42) -> Workload:
    nx, ny, nz = 64, 64, 32
    lh = level_heights(nz, 20_000.0)
    z = np.broadcast_to(lh[None, None, :], (nx, ny, nz))
    u = (10.0 + 2.0e-3 * z).astype(np.float32).reshape(-1)
    v = np.full(nx * ny * nz, 2.0, np.float32)
    w = np.zeros(nx * ny * nz, np.float32)
    met = Met((nx, ny, nz), [(u, v, w), (u.copy(), v.copy(), w.copy())], lh, 3 * 3600.0)
    np2 = nx * ny
    pbl = Pbl((nx, ny), np.full(np2, 0.4, np.float32), np.full(np2, 1.5, np.float32),
              np.full(np2, 1000.0, np.float32), np.full(np2, -1.0 / 200.0, np.float32))
    parts = make_particles(n_particles, nx / 2.0, ny / 2.0, 100.0, 1.0 / n_particles)
    parts["flags"][:] = 0  # released over the first hour by the caller
    scale = 1.0 / METRES_PER_DEG
    return Workload("point_release", met, pbl, parts, 900.0, 96, scale, scale, (seed, 0),
                    (64, 64, 4), (100.0, 500.0, 1000.0, 3000.0), (1, 1), 1.0e9,
                    dry_vd=2.0e-3, release_seconds=3600.0)


def _fourier_field(rng, nx, ny, kmax, rms, deg: float = 1.0):
    """sum over 0 <= kx, ky <= kmax, (kx, ky) != 0 of a cos(kx lon + 2 ky lat + ph),
    a ~ N(0, 1), ph ~ U[0, 2 pi) drawn in (kx, ky) order, scaled to the given rms,
    on a global grid of ``deg`` degree spacing (lon from 0, lat from -90).
    Evaluated as Re(Ex^T C Ey) with Ex = exp(i kx lon), Ey = exp(2 i ky lat)."""
    lon = np.deg2rad(np.arange(nx, dtype=np.float64) * deg)
    lat = np.deg2rad(np.arange(ny, dtype=np.float64) * deg - 90.0)
    k = np.arange(kmax + 1)
    coef = np.zeros((kmax + 1, kmax + 1), np.complex128)
    for kx in range(0, kmax + 1):
        for ky in range(0, kmax + 1):
            if kx == 0 and ky == 0:
                continue
            a = rng.normal()
            ph = rng.uniform(0, 2 * np.pi)
            coef[kx, ky] = a * np.exp(1j * ph)
    ex = np.exp(1j * np.outer(lon, k))        # (nx, K)
    ey = np.exp(2j * np.outer(lat, k))        # (ny, K)
    f = np.real(ex @ coef @ ey.T)
    f *= rms / max(np.sqrt(np.mean(f ** 2)), 1e-12)
    return f


def _wind_level(rng, nx, ny, lh, deg: float = 1.0):
    lat = np.deg2rad(np.arange(ny, dtype=np.float64) * deg - 90.0)
    up = _fourier_field(rng, nx, ny, 8, 5.0 / math.sqrt(2.0), deg)
    vp = _fourier_field(rng, nx, ny, 8, 5.0 / math.sqrt(2.0), deg)
    u2 = 30.0 * np.sin(2.0 * lat)[None, :] + up
    v2 = vp
    # horizontal divergence on the sphere (centred differences, periodic in lon)
    coslat = np.maximum(np.cos(lat), 1e-3)[None, :]
    dx = METRES_PER_DEG * deg * coslat
    dudx = (np.roll(u2, -1, 0) - np.roll(u2, 1, 0)) / (2.0 * dx)
    dvdy = np.gradient(v2 * coslat, axis=1) / (METRES_PER_DEG * deg * coslat)
    div = dudx + dvdy
    zf = lh.astype(np.float64)
    g = 0.6 + 0.4 * np.minimum(zf / 10_000.0, 1.0)           # wind profile factor
    G = np.concatenate([[0.0], np.cumsum(0.5 * (g[1:] + g[:-1]) * np.diff(zf))])  # int_0^z g
    gf = g.astype(np.float32)[None, None, :]
    u = u2.astype(np.float32)[:, :, None] * gf
    v = v2.astype(np.float32)[:, :, None] * gf
    w = (-div).astype(np.float32)[:, :, None] * G.astype(np.float32)[None, None, :]  # continuity
    return u.reshape(-1), v.reshape(-1), w.reshape(-1)


def global_met(seed: int = 2, nx: int = 360, ny: int = 181, nz: int = 60,
               n_levels: int = 5, deg: float = 1.0, seconds_between_levels: float = 6 * 3600.0) -> Met:
    lh = level_heights(nz, 25_000.0)
    levels = [_wind_level(np.random.default_rng(seed * 1000 + t), nx, ny, lh, deg)
              for t in range(n_levels)]
    return Met((nx, ny, nz), levels, lh, seconds_between_levels)


def global_pbl(seed: int = 2, nx: int = 360, ny: int = 181) -> Pbl:
    rng = np.random.default_rng(seed * 7919)
    n = nx * ny
    hmix = (800.0 * np.exp(0.5 * rng.normal(size=n))).astype(np.float32)
    oli = rng.uniform(-0.02, 0.02, n).astype(np.float32)
    ustar = rng.uniform(0.1, 0.8, n).astype(np.float32)
    wstar = np.where(oli < 0, 1.5 * np.cbrt(hmix / 1000.0), 0.0).astype(np.float32)
    return Pbl((nx, ny), ustar, wstar, hmix, oli, np.zeros(n, np.float32))


def global_sources(n_particles: int, seed: int = 2, n_sources: int = 16, nx=360, ny=181):
    rng = np.random.default_rng(seed * 104729)
    sx = rng.uniform(0.0, nx - 1.0, n_sources)
    sy = rng.uniform(10.0, ny - 11.0, n_sources)
    sz = rng.uniform(10.0, 500.0, n_sources)
    emis = np.exp(rng.normal(size=n_sources))  # lognormal, median 1 kg, sigma 1
    per = np.full(n_sources, n_particles // n_sources)
    per[: n_particles % n_sources] += 1
    src = np.repeat(np.arange(n_sources), per)
    x = sx[src].astype(np.float32)
    y = sy[src].astype(np.float32)
    z = sz[src].astype(np.float32)
    mass = (emis[src] / per[src]).astype(np.float32)
    return x, y, z, mass


def global_10m(n_particles: int = 10_000_000, seed: int = 2, met: Met | None = None,
               pbl: Pbl | None = None) -> Workload:
    met = global_met(seed) if met is None else met
    pbl = global_pbl(seed) if pbl is None else pbl
    x, y, z, mass = global_sources(n_particles, seed)
    parts = make_particles(n_particles, x, y, z, mass)
    scale = 1.0 / METRES_PER_DEG
    return Workload("global_10m", met, pbl, parts, 600.0, 144, scale, scale, (seed, 0),
                    (720, 361, 5), (100.0, 500.0, 1000.0, 3000.0, 10_000.0), (2, 2), 1.0e8)


def small_random(n_particles: int, shape=(64, 64, 32), seed: int = 7, n_levels: int = 2):
    """Small randomized workload for parity tests: random Fourier winds on a
    reduced grid, random PBL fields covering all three Hanna regimes and
    particles spread over the domain and the boundary layer."""
    nx, ny, nz = shape
    met_rng = np.random.default_rng(seed)
    lh = level_heights(nz, 15_000.0)
    levels = []
    for _ in range(n_levels):
        u = met_rng.normal(5.0, 8.0, (nx, ny, nz)).astype(np.float32)
        v = met_rng.normal(0.0, 6.0, (nx, ny, nz)).astype(np.float32)
        w = met_rng.normal(0.0, 0.05, (nx, ny, nz)).astype(np.float32)
        levels.append((u.reshape(-1), v.reshape(-1), w.reshape(-1)))
    met = Met(shape, levels, lh, 3 * 3600.0)
    n = nx * ny
    hmix = (800.0 * np.exp(0.5 * met_rng.normal(size=n))).astype(np.float32)
    oli = met_rng.uniform(-0.02, 0.02, n).astype(np.float32)
    oli[met_rng.uniform(size=n) < 0.1] = 0.0  # neutral cells
    ustar = met_rng.uniform(0.1, 0.8, n).astype(np.float32)
    wstar = np.where(oli < 0, 1.5 * np.cbrt(hmix / 1000.0), 0.0).astype(np.float32)
    precip = np.where(met_rng.uniform(size=n) < 0.3,
                      2.0 * np.exp(met_rng.normal(size=n)), 0.0).astype(np.float32)
    pbl = Pbl((nx, ny), ustar, wstar, hmix, oli, precip)
    rng = np.random.default_rng(seed + 1)
    x = rng.uniform(0, nx - 1, n_particles).astype(np.float32)
    y = rng.uniform(0, ny - 1, n_particles).astype(np.float32)
    z = rng.uniform(0.0, 2500.0, n_particles).astype(np.float32)
    mass = np.exp(rng.normal(size=n_particles)).astype(np.float32)
    parts = make_particles(n_particles, x, y, z, mass)
    parts["turb_u"] = rng.normal(0, 0.5, n_particles)
    parts["turb_v"] = rng.normal(0, 0.5, n_particles)
    parts["turb_w"] = rng.normal(0, 0.3, n_particles)
    parts["cbt"] = np.where(rng.uniform(size=n_particles) < 0.5, 1, -1)
    parts["flags"][rng.uniform(size=n_particles) < 0.1] = 0  # inactive slots
    scale = 1.0 / 25_000.0
    return Workload("small_random", met, pbl, parts, 600.0, 4, scale, scale, (0xDECAFBAD, 0x12345678),
                    (shape[0] * 2, shape[1] * 2, 4), (100.0, 500.0, 1000.0, 3000.0), (2, 2), 1.0e6,
                    dry_vd=2.0e-3)


@dataclass
class EnsembleSpec:
    """Batched-ensemble workload: member sizes, release steps and Philox keys."""
    counts: np.ndarray        # particles per member
    start_step: np.ndarray    # sync step of each member's release
    keys: list                # (key0, key1) per member
    seed: int
    dt_seconds: float = 600.0


def ensemble_spec(n_members: int = 96, median: float = 200_000.0, sigma: float = 1.0,
                  lo: int = 10_000, hi: int = 4_000_000, release_hours: float = 6.0,
                  dt_seconds: float = 600.0, seed: int = 3) -> EnsembleSpec:
    """96 members with lognormal particle counts (median 200k, sigma 1,
    clipped to [10k, 4M]; about 30M in total) released at U[0, 6 h)."""
    rng = np.random.default_rng(seed * 15485863)
    counts = np.clip(np.round(median * np.exp(sigma * rng.normal(size=n_members))), lo, hi)
    start = rng.uniform(0.0, release_hours * 3600.0, n_members)
    keys = [((seed * 1000 + m) & 0xFFFFFFFF, 0x5EED0000 ^ m) for m in range(n_members)]
    return EnsembleSpec(counts.astype(np.int64), np.floor(start / dt_seconds).astype(np.int64), keys,
                        seed, dt_seconds)


def ensemble_member_particles(spec: EnsembleSpec, m: int) -> np.ndarray:
    """Member m's particles: 4 point sources of its own (global_sources)."""
    n = int(spec.counts[m])
    x, y, z, mass = global_sources(n, seed=spec.seed * 100 + m, n_sources=4)
    return make_particles(n, x, y, z, mass)


def ensemble_member_factor(spec: EnsembleSpec, m: int, nx: int = 360, ny: int = 181,
                           amplitude: float = 0.1) -> np.ndarray:
    """Member m's wind perturbation as a multiplicative (nx, ny) factor
    1 + amplitude * F, F a member-seeded unit-rms Fourier field (wavenumbers
    <= 8) applied to u, v and w at every level of the base met."""
    rng = np.random.default_rng(spec.seed * 7 + 1000 * (m + 1))
    return (1.0 + amplitude * _fourier_field(rng, nx, ny, 8, 1.0)).astype(np.float32)


@dataclass
class AdjointCase:
    """Point sources, wind-perturbation basis and parameters of a
    differentiable run."""
    met: Met
    pbl: Pbl
    counts: np.ndarray          # particles per source
    emission: np.ndarray        # kg per source
    position: np.ndarray        # (S, 3): x, y [grid units], z [m]
    basis_u: np.ndarray         # (K, nx * ny)
    basis_v: np.ndarray
    coeff: np.ndarray           # (K,)
    dt_seconds: float
    x_scale: float
    y_scale: float
    key: tuple


def wind_basis(n_modes: int, nx: int, ny: int, seed: int = 5):
    """K unit-rms Fourier fields (wavenumbers <= 4): the first half perturbs
    u, the second half v."""
    rng = np.random.default_rng(seed * 7907)
    bu = np.zeros((n_modes, nx * ny), np.float32)
    bv = np.zeros((n_modes, nx * ny), np.float32)
    for k in range(n_modes):
        f = _fourier_field(rng, nx, ny, 4, 1.0).astype(np.float32).reshape(-1)
        (bu if k < (n_modes + 1) // 2 else bv)[k] = f
    return bu, bv


def adjoint_case(n_particles: int = 2_000_000, n_sources: int = 1024, n_modes: int = 64,
                 met: Met | None = None, pbl: Pbl | None = None, seed: int = 4,
                 dt_seconds: float = 600.0, x_scale: float | None = None) -> AdjointCase:
    """Sources at uniform-random positions (heights U[10, 500] m) with
    lognormal emissions (median 1 kg, sigma 1), particles split evenly,
    zero wind coefficients.  Defaults: the 2M-particle, 1024-source case on
    the 1 deg global met with 3 time levels (12 h)."""
    met = global_met(2, n_levels=3) if met is None else met
    pbl = global_pbl(2) if pbl is None else pbl
    nx, ny, _ = met.shape
    rng = np.random.default_rng(seed * 6151)
    counts = np.full(n_sources, n_particles // n_sources, np.uint32)
    counts[: n_particles % n_sources] += 1
    margin = min(10.0, (ny - 1) / 4.0)
    pos = np.stack([rng.uniform(1.0, nx - 2.0, n_sources), rng.uniform(margin, ny - 1.0 - margin, n_sources),
                    rng.uniform(10.0, 500.0, n_sources)], axis=1).astype(np.float32)
    q = np.exp(rng.normal(size=n_sources)).astype(np.float32)
    bu, bv = wind_basis(n_modes, nx, ny, seed)
    scale = (1.0 / METRES_PER_DEG) if x_scale is None else x_scale
    return AdjointCase(met, pbl, counts, q, pos, bu, bv, np.zeros(n_modes, np.float32), dt_seconds, scale, scale,
                       (seed, 0x0AD10AD1))


def precip_patches(nx: int, ny: int, cover: float = 0.2, median: float = 2.0, sigma: float = 1.0,
                   seed: int = 5) -> np.ndarray:
    """Precipitation [mm/h] in seeded elliptical cloud patches covering about
    ``cover`` of the (periodic in x) domain, lognormal inside (median
    ``median``, log-sd ``sigma``), zero elsewhere; flat index i * ny + j."""
    rng = np.random.default_rng(seed * 3571)
    mask = np.zeros((nx, ny), bool)
    scale = nx / 360.0
    while mask.mean() < cover:
        cx, cy = rng.uniform(0, nx), rng.uniform(0, ny - 1)
        a, b = rng.uniform(4, 30) * scale, rng.uniform(2, 15) * scale
        th = rng.uniform(0, np.pi)
        r = int(np.ceil(max(a, b))) + 1
        ii = np.arange(int(cx) - r, int(cx) + r + 1)
        jj = np.arange(max(0, int(cy) - r), min(ny, int(cy) + r + 1))
        dx = (ii[:, None] - cx).astype(np.float64)
        dy = (jj[None, :] - cy).astype(np.float64)
        u = (dx * np.cos(th) + dy * np.sin(th)) / a
        v = (-dx * np.sin(th) + dy * np.cos(th)) / b
        inside = u * u + v * v <= 1.0
        mask[np.ix_(ii % nx, jj)] |= inside
    p = np.where(mask, median * np.exp(sigma * rng.normal(size=(nx, ny))), 0.0)
    return p.astype(np.float32).reshape(-1)


def stress_case(n_particles: int = 50_000_000, n_sources: int = 4096, seed: int = 6, deg: float = 0.25,
                nz: int = 60) -> AdjointCase:
    """The 50M-particle deposition stress case: 0.25 deg global met (1440 x
    721 x 60) with two time levels 24 h apart, precipitation in elliptical
    patches over 20 % of the domain, 4096 sources at uniform-random positions
    (heights U[10, 500] m) with lognormal emissions; particles are split in
    proportion to the emissions so that every particle starts with the same
    mass (the loss rule then removes a particle once deposition has taken a
    fixed fraction of it)."""
    nx, ny = int(round(360 / deg)), int(round(180 / deg)) + 1
    met = global_met(seed, nx, ny, nz, n_levels=2, deg=deg, seconds_between_levels=24 * 3600.0)
    pbl = global_pbl(seed, nx, ny)
    pbl.precip = precip_patches(nx, ny, seed=seed)
    rng = np.random.default_rng(seed * 6151)
    q = np.exp(rng.normal(size=n_sources))
    counts = np.maximum(1, np.floor(n_particles * q / q.sum())).astype(np.int64)
    counts[np.argsort(-q)[: n_particles - int(counts.sum())]] += 1
    margin = 10.0 / deg
    pos = np.stack([rng.uniform(1.0, nx - 2.0, n_sources), rng.uniform(margin, ny - 1.0 - margin, n_sources),
                    rng.uniform(10.0, 500.0, n_sources)], axis=1).astype(np.float32)
    scale = 1.0 / (METRES_PER_DEG * deg)
    return AdjointCase(met, pbl, counts.astype(np.uint32), q.astype(np.float32), pos,
                       np.zeros((0, nx * ny), np.float32), np.zeros((0, nx * ny), np.float32),
                       np.zeros(0, np.float32), 600.0, scale, scale, (seed, 0x57E55))


def surface_bracket(nx: int, ny: int, seed: int = 11):
    """Two time levels of the ten surface fields the PBL diagnostics read
    (FP_SURF_* order: pressure, T2m, u10, v10, stress, sshf, ssr, hmix,
    met u*, met 1/L), nx*ny each, index i*ny+j: smooth large-scale patterns
    plus noise, a diurnal heat-flux swing between the two levels, stress
    missing over a third of the cells (log-law u*), hmix missing over a fifth
    (fallback) and met u* / 1/L provided over half."""
    rng = np.random.default_rng(seed)
    lat = np.linspace(-90.0, 90.0, ny, dtype=np.float64)[None, :] * np.ones((nx, 1))
    lon = np.linspace(0.0, 360.0, nx, endpoint=False, dtype=np.float64)[:, None] * np.ones((1, ny))
    levels = []
    for lvl in range(2):
        phase = np.deg2rad(lon) + lvl * np.pi / 2
        t2m = 288.0 - 30.0 * np.sin(np.deg2rad(lat)) ** 2 + rng.normal(0, 2.0, lat.shape)
        ps = 101_325.0 - 1500.0 * np.cos(3 * np.deg2rad(lat)) + rng.normal(0, 300.0, lat.shape)
        u10 = 8.0 * np.sin(2 * np.deg2rad(lat)) + rng.normal(0, 3.0, lat.shape)
        v10 = rng.normal(0, 3.0, lat.shape)
        stress = np.where(rng.uniform(size=lat.shape) < 1 / 3, 0.0, 0.0012 * 1.2 * (u10 ** 2 + v10 ** 2))
        sshf = 150.0 * np.sin(phase) * np.cos(np.deg2rad(lat)) + rng.normal(0, 20.0, lat.shape)
        ssr = np.maximum(0.0, 800.0 * np.sin(phase)) * np.cos(np.deg2rad(lat))
        hmix = np.where(rng.uniform(size=lat.shape) < 0.2, 0.0, 800.0 * np.exp(rng.normal(0, 0.6, lat.shape)))
        ust = np.where(rng.uniform(size=lat.shape) < 0.5, 0.0, rng.uniform(0.1, 0.8, lat.shape))
        oli = np.where(rng.uniform(size=lat.shape) < 0.5, 0.0, rng.uniform(-0.02, 0.02, lat.shape))
        levels.append([np.ascontiguousarray(x.reshape(-1).astype(np.float32))
                       for x in (ps, t2m, u10, v10, stress, sshf, ssr, hmix, ust, oli)])
    return levels
