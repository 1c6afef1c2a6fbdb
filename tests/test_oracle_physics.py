"""Physics stages of the oracle against the reference's closed forms.

Restates the reference's own unit tests: advection (physics/advection.rs:
282-487), Hanna (physics/hanna.rs:280-520), Langevin (physics/langevin.rs:
354-420) and deposition decay (tests/integration/deposition_decay.rs), plus
invariants of the full step.
"""
import numpy as np
import pytest

from conftest import met_dict, pbl_dict


def particle_at(fp, x, y, z, mass=(1.0, 0.0, 0.0, 0.0)):
    p = np.zeros(1, fp.PARTICLE_DTYPE)
    cx, cy = int(np.floor(x)), int(np.floor(y))
    p["cell_x"], p["cell_y"] = cx, cy
    p["pos_x"], p["pos_y"] = np.float32(x) - np.float32(cx), np.float32(y) - np.float32(cy)
    p["pos_z"] = z
    p["mass"][0] = mass
    p["flags"] = 1
    p["cbt"] = 1
    return p


def grid_x(p):
    return p["cell_x"].astype(np.float64) + p["pos_x"].astype(np.float64)


def grid_y(p):
    return p["cell_y"].astype(np.float64) + p["pos_y"].astype(np.float64)


def uniform_met(nx, ny, nz, u, v, w, heights=None):
    n = nx * ny * nz
    f = [np.full(n, c, np.float32) for c in (u, v, w)]
    return dict(shape=(nx, ny, nz), u0=f[0], v0=f[1], w0=f[2], u1=f[0], v1=f[1], w1=f[2],
                level_heights=np.zeros(nz, np.float32) if heights is None else heights)


def shear_met(n, u_ref, v_ref, w, du_dz, dv_dz):
    k = np.arange(n, dtype=np.float32)
    u = np.broadcast_to(np.float32(u_ref) + np.float32(du_dz) * k, (n, n, n)).astype(np.float32).reshape(-1)
    v = np.broadcast_to(np.float32(v_ref) + np.float32(dv_dz) * k, (n, n, n)).astype(np.float32).reshape(-1)
    ww = np.full(n ** 3, w, np.float32)
    return dict(shape=(n, n, n), u0=u, v0=v, w0=ww, u1=u, v1=v, w1=ww,
                level_heights=np.zeros(n, np.float32))


IDENTITY = dict(x_scale=1.0, y_scale=1.0, z_scale=1.0)


# ------------------------------------------------------------------ advection

def test_advection_uniform_wind_matches_analytical_solution(orc, fp):
    # physics/advection.rs:282-330
    dt, u, v, w = 2.75, 0.8, -0.35, 0.15
    met = uniform_met(64, 64, 64, u, v, w)
    for x0, y0, z0 in [(5.25, 8.5, 6.25), (11.0, 13.125, 9.75), (23.9, 17.4, 12.1),
                       (31.375, 4.2, 3.0)]:
        p = particle_at(fp, x0, y0, z0)
        orc.advect(p, met, dict(dt_seconds=dt, **IDENTITY))
        assert abs(grid_x(p)[0] - np.float32(x0 + u * dt)) <= 1e-5
        assert abs(grid_y(p)[0] - np.float32(y0 + v * dt)) <= 1e-5
        assert abs(p["pos_z"][0] - np.float32(z0 + w * dt)) <= 1e-5
        assert abs(p["vel_u"][0] - u) <= np.finfo(np.float32).eps
        assert abs(p["vel_v"][0] - v) <= np.finfo(np.float32).eps
        assert abs(p["vel_w"][0] - w) <= np.finfo(np.float32).eps


def test_advection_backward_direction_reverses_uniform_transport(orc, fp):
    # backward driver: advection with -dt (timeloop.rs:1724-1738)
    met = uniform_met(64, 64, 64, 0.8, -0.35, 0.15)
    p = particle_at(fp, 20.25, 30.5, 10.0)
    orc.advect(p, met, dict(dt_seconds=2.0, direction=-1, **IDENTITY))
    assert abs(grid_x(p)[0] - (20.25 - 1.6)) <= 1e-5
    assert abs(grid_y(p)[0] - (30.5 + 0.7)) <= 1e-5
    assert abs(p["pos_z"][0] - (10.0 - 0.3)) <= 1e-5


def test_petterssen_linear_x_shear_matches_scheme_closed_form(orc, fp):
    # physics/advection.rs:332-363: u = a x; one Petterssen step gives
    # x0 (1 + a dt + (a dt)^2 / 2), closer to x0 e^{a dt} than Euler's x0 (1 + a dt)
    a, dt, x0 = 0.2, 1.0, 2.5
    nx, ny, nz = 64, 4, 4
    u = np.broadcast_to((a * np.arange(nx, dtype=np.float32))[:, None, None], (nx, ny, nz))
    u = u.astype(np.float32).reshape(-1)
    z = np.zeros(nx * ny * nz, np.float32)
    met = dict(shape=(nx, ny, nz), u0=u, v0=z, w0=z, u1=u, v1=z, w1=z,
               level_heights=np.zeros(nz, np.float32))
    p = particle_at(fp, x0, 1.0, 1.0)
    orc.advect(p, met, dict(dt_seconds=dt, **IDENTITY))
    pet = x0 * (1 + a * dt + (a * dt) ** 2 / 2)
    assert abs(grid_x(p)[0] - pet) < 1e-5
    exact = x0 * np.exp(a * dt)
    assert abs(grid_x(p)[0] - exact) < abs(x0 * (1 + a * dt) - exact)


def test_advection_clamps_out_of_domain_position_stably(orc, fp):
    # physics/advection.rs:365-380
    met = uniform_met(3, 3, 3, 10.0, -10.0, 10.0)
    p = particle_at(fp, 1.8, 0.2, 1.0)
    orc.advect(p, met, dict(dt_seconds=1.0, **IDENTITY))
    assert p["cell_x"][0] == 2 and p["pos_x"][0] == 0.0
    assert p["cell_y"][0] == 0 and p["pos_y"][0] == 0.0
    assert p["pos_z"][0] == 2.0


def shear_expected(x0, y0, z0, t, u_ref, v_ref, w, du_dz, dv_dz):
    z = z0 + w * t
    x = x0 + (u_ref + du_dz * z0) * t + 0.5 * du_dz * w * t * t
    y = y0 + (v_ref + dv_dz * z0) * t + 0.5 * dv_dz * w * t * t
    return x, y, z


def test_petterssen_linear_shear_matches_analytical_single_step(orc, fp):
    # physics/advection.rs:382-431
    met = shear_met(96, 1.2, -0.35, 0.2, 0.04, -0.015)
    p = particle_at(fp, 10.25, 11.75, 5.4)
    orc.advect(p, met, dict(dt_seconds=0.6, **IDENTITY))
    x, y, z = shear_expected(10.25, 11.75, 5.4, 0.6, 1.2, -0.35, 0.2, 0.04, -0.015)
    assert abs(grid_x(p)[0] - x) <= 1e-5
    assert abs(grid_y(p)[0] - y) <= 1e-5
    assert abs(p["pos_z"][0] - z) <= 1e-5


def test_petterssen_linear_shear_matches_analytical_multi_step(orc, fp):
    # physics/advection.rs:433-487
    met = shear_met(128, 0.8, 0.5, 0.3, 0.03, -0.02)
    p = particle_at(fp, 8.125, 9.625, 6.5)
    for _ in range(20):
        orc.advect(p, met, dict(dt_seconds=0.25, **IDENTITY))
    x, y, z = shear_expected(8.125, 9.625, 6.5, 5.0, 0.8, 0.5, 0.3, 0.03, -0.02)
    assert abs(grid_x(p)[0] - x) <= 2e-4
    assert abs(grid_y(p)[0] - y) <= 2e-4
    assert abs(p["pos_z"][0] - z) <= 2e-4


def test_temporal_interpolation_blends_the_two_time_levels(orc, fp):
    # advection_dual_wind.wgsl:171 lerp(t0, t1, alpha), alpha clamped to [0, 1]
    met = uniform_met(16, 16, 8, 1.0, 0.0, 0.0)
    met["u1"] = np.full_like(met["u0"], 3.0)
    for alpha, want in [(0.25, 1.5), (0.0, 1.0), (1.0, 3.0), (1.7, 3.0), (-0.5, 1.0)]:
        p = particle_at(fp, 4.5, 4.5, 2.0)
        orc.advect(p, met, dict(dt_seconds=0.5, alpha=alpha, **IDENTITY))
        assert p["vel_u"][0] == np.float32(want)


def test_metre_levels_map_piecewise_linearly_to_level_index(orc, fp):
    # meters_to_grid_level (advection_dual_wind.wgsl:94-117): a field equal to
    # the level index samples to the fractional level of z
    nz = 6
    heights = np.array([10.0, 50.0, 150.0, 400.0, 1000.0, 3000.0], np.float32)
    n = 4 * 4 * nz
    k = np.broadcast_to(np.arange(nz, dtype=np.float32), (4, 4, nz)).reshape(-1).astype(np.float32)
    zero = np.zeros(n, np.float32)
    met = dict(shape=(4, 4, nz), u0=k, v0=zero, w0=zero, u1=k, v1=zero, w1=zero,
               level_heights=heights)
    for z, level in [(5.0, 0.5), (10.0, 1.0 * 10 / 10 - 0.0), (30.0, 0.5), (100.0, 1.5),
                     (2000.0, 4.5), (3000.0, 5.0), (2999.0, 4.9995)]:
        p = particle_at(fp, 1.5, 1.5, z)
        orc.advect(p, met, dict(dt_seconds=1e-3, **IDENTITY))
        assert abs(p["vel_u"][0] - level) < 1e-4, (z, p["vel_u"][0], level)


def test_vertical_position_clamps_to_top_level_height(orc, fp):
    heights = np.array([10.0, 100.0, 1000.0], np.float32)
    met = uniform_met(4, 4, 3, 0.0, 0.0, 5.0, heights)
    p = particle_at(fp, 1.5, 1.5, 990.0)
    orc.advect(p, met, dict(dt_seconds=10.0, **IDENTITY))
    assert p["pos_z"][0] == 1000.0
    p = particle_at(fp, 1.5, 1.5, 5.0)
    met = uniform_met(4, 4, 3, 0.0, 0.0, -5.0, heights)
    orc.advect(p, met, dict(dt_seconds=10.0, **IDENTITY))
    assert p["pos_z"][0] == 0.0


# ------------------------------------------------------------------ Hanna

TABLE = [
    # physics/hanna.rs:359-520 (name, inputs, sigu/u*, sigw/u*, rel tol)
    ("stable_low", (0.30, 0.0, 120.0, 1000.0, 100.0), 1.83, 1.20, 0.20),
    ("stable_mid", (0.28, 0.0, 150.0, 1000.0, 500.0), 1.04, 0.66, 0.22),
    ("neutral_low", (0.30, 0.0, 5000.0, 1000.0, 50.0), 1.93, 1.29, 0.16),
    ("neutral_aloft", (0.32, 0.0, 5000.0, 1000.0, 600.0), 1.17, 0.93, 0.20),
    ("unstable_low", (0.30, 2.0, -80.0, 1000.0, 100.0), 2.66, 3.51, 0.22),
    ("unstable_mid", (0.30, 2.0, -80.0, 1000.0, 600.0), 2.66, 4.32, 0.22),
]


@pytest.mark.parametrize("name,inputs,su,sw,tol", TABLE)
def test_hanna_sigma_matches_table_style_reference(orc, name, inputs, su, sw, tol):
    h = orc.hanna(*inputs)
    assert abs(h["sigu"] / h["ust"] - su) <= tol * su
    assert abs(h["sigw"] / h["ust"] - sw) <= tol * sw
    if name.startswith("unstable"):
        assert abs(h["sigu"] - h["sigv"]) < 1e-6 and h["sigw"] > h["sigu"]
    else:
        assert abs(h["sigv"] - h["sigw"]) < 1e-6


def test_hanna_no_nan_or_inf_for_realistic_ranges(orc):
    # physics/hanna.rs:289-324
    for ust in (0.05, 0.2, 0.6):
        for wst in (0.0, 0.8, 2.5):
            for ol in (-500.0, -80.0, 80.0, 500.0, 2.0e6):
                for hm in (100.0, 800.0, 2500.0):
                    for z in (0.0, 5.0, 50.0, 200.0, 700.0, 2000.0):
                        p = orc.hanna(ust, wst, ol, hm, z)
                        assert all(np.isfinite(v) for v in p.values())
                        assert p["sigu"] > 0 and p["sigv"] > 0 and p["sigw"] > 0
                        assert p["tlu"] >= 10 and p["tlv"] >= 10 and p["tlw"] >= 30


def test_hanna_regime_formulas_closed_form(orc):
    # neutral branch with L = +inf (|1/L| < 1e-5): sigma_u = 0.01 + 2 u* e^{-3e-4 z/u*}
    ust, z, h = 0.3, 120.0, 900.0
    p = orc.hanna(ust, 0.0, np.inf, h, z)
    assert abs(p["sigu"] - (0.01 + 2 * ust * np.exp(-3e-4 * z / ust))) < 1e-6
    sw = 1.3 * ust * np.exp(-2e-4 * z / ust)
    assert abs(p["sigw"] - (sw + 0.01)) < 1e-6
    assert abs(p["dsigwdz"] - (-2e-4 * sw)) < 1e-9
    # stable: sigma_u = 0.01 + 2 u* (1 - z/h), T_Lw = max(0.1 h/sigma_w (z/h)^0.8, 30)
    p = orc.hanna(0.4, 0.0, 100.0, 1000.0, 250.0)
    zeta = 0.25
    sigw = 0.01 + 1.3 * 0.4 * (1 - zeta)
    assert abs(p["sigu"] - (0.01 + 0.8 * (1 - zeta))) < 1e-6
    assert abs(p["tlw"] - max(0.1 * 1000 / sigw * zeta ** 0.8, 30)) < 1e-3
    # unstable: sigma_u = 0.01 + u* (12 - 0.5 h/L)^(1/3)
    p = orc.hanna(0.3, 2.0, -80.0, 1000.0, 100.0)
    assert abs(p["sigu"] - (0.01 + 0.3 * (12 + 0.5 * 1000 / 80) ** (1 / 3))) < 1e-5
    # z above h is clamped to h, h below 1 m raised to 1 m
    assert orc.hanna(0.3, 1.0, -50.0, 500.0, 900.0)["zeta"] == 1.0
    assert orc.hanna(0.3, 1.0, -50.0, 0.2, 0.1)["h"] == 1.0


def test_unstable_has_more_vertical_mixing_than_stable(orc):
    stable = orc.hanna(0.3, 0.0, 120.0, 1000.0, 100.0)
    unstable = orc.hanna(0.3, 2.0, -120.0, 1000.0, 100.0)
    assert unstable["sigw"] > stable["sigw"]
    assert unstable["tlu"] >= stable["tlu"]


# ------------------------------------------------------------------ Langevin

def test_zero_turbulence_yields_stable_deterministic_behavior(orc):
    # physics/langevin.rs:361-381: sigma 0, T_L 50 s, dt 10 s, no noise
    assert abs(orc.update_horizontal(2.0, 0.0, 50.0, 10.0, 0.0) - 1.6) < 1e-6
    assert abs(orc.update_horizontal(-1.0, 0.0, 50.0, 10.0, 0.0) + 0.8) < 1e-6
    assert abs(orc.update_vertical(0.5, 0.0, 0.0, 50.0, 10.0, 0.0, 0.0) - 0.4) < 1e-6


def test_langevin_exact_branch_decay_and_degenerate_time_scale(orc):
    # r = dt/T_L >= 0.5: u' = e^{-r} u (no noise)
    assert abs(orc.update_horizontal(2.0, 0.0, 10.0, 20.0, 0.0) - 2.0 * np.exp(-2.0)) < 1e-6
    # T_L <= 0 or non-finite -> 0
    assert orc.update_horizontal(2.0, 1.0, 0.0, 20.0, 0.3) == 0.0
    assert orc.update_horizontal(2.0, 1.0, np.nan, 20.0, 0.3) == 0.0
    # vertical drift term: sigma_w (dsigma_w/dz + rho' sigma_w) * dt in the small-step branch
    got = orc.update_vertical(0.0, 0.5, 0.01, 1000.0, 10.0, 0.0, 0.0)
    assert abs(got - 10.0 * 0.5 * 0.01) < 1e-7


@pytest.mark.parametrize("dt,tl", [(60.0, 100.0), (10.0, 100.0)])
def test_ornstein_uhlenbeck_update_is_stationary_at_sigma(orc, dt, tl):
    # both branches keep Var(u) at sigma^2 (exactly for r >= 0.5, to O(r^2) below)
    rng = np.random.default_rng(3)
    sigma = 1.3
    u = rng.normal(0.0, sigma, 3000).astype(np.float32)
    for _ in range(25):
        eta = rng.normal(size=u.size).astype(np.float32)
        u = np.array([orc.update_horizontal(float(a), sigma, tl, dt, float(e))
                      for a, e in zip(u, eta)], np.float32)
    tol = 0.08 if dt / tl >= 0.5 else 0.15
    assert abs(u.var() / sigma ** 2 - 1.0) < tol


# ------------------------------------------------------------------ deposition

STEPS, DT, VD, LAM = 12, 60.0, 0.024, 1.6e-3
MASS0 = (1.0, 0.4, 0.2, 0.1)


def _evolve(orc, fp, params):
    p = particle_at(fp, 2.25, 3.75, 5.0)
    p["mass"][0] = MASS0
    out = [float(p["mass"][0].astype(np.float64).sum())]
    for _ in range(STEPS):
        orc.deposit(p, dict(dt_seconds=DT, **params))
        out.append(float(p["mass"][0].astype(np.float64).sum()))
    return np.array(out)


@pytest.mark.parametrize("kind", ["dry", "wet", "combined"])
def test_deposition_mass_evolution_matches_analytical_exponential_decay(orc, fp, kind):
    # tests/integration/deposition_decay.rs:10-30,222-268 (CPU tolerance 2e-6)
    params = {}
    rate = 0.0
    if kind in ("dry", "combined"):
        params.update(dry_mode=1, dry_vd=VD, dry_href=15.0)
        rate += VD / (2 * 15.0)
    if kind in ("wet", "combined"):
        params.update(wet_mode=1, wet_lambda=LAM, wet_fraction=1.0)
        rate += LAM
    m = _evolve(orc, fp, params)
    t = np.arange(STEPS + 1) * DT
    want = sum(MASS0) * np.exp(-rate * t)
    assert np.all(np.diff(m) < 0)
    assert np.all(np.abs(m - want) <= 2e-6 + 2e-6 * want)


def test_dry_deposition_only_below_twice_reference_height(orc, fp):
    p = particle_at(fp, 2.0, 2.0, 30.0)
    orc.deposit(p, dict(dt_seconds=600.0, dry_mode=1, dry_vd=0.01, dry_href=15.0))
    assert p["mass"][0, 0] == 1.0
    p["pos_z"] = 29.99
    orc.deposit(p, dict(dt_seconds=600.0, dry_mode=1, dry_vd=0.01, dry_href=15.0))
    assert abs(p["mass"][0, 0] - np.exp(-0.01 * 600 / 30)) < 1e-6


def test_wet_scavenging_from_precip_field_uses_a_p_to_the_b(orc, fp):
    # below_cloud_scavenging_coefficient_gas: lambda = A * P^B (wet_scavenging.rs:213-225)
    pbl = dict(shape=(4, 4), ustar=np.full(16, 0.3, np.float32), wstar=np.zeros(16, np.float32),
               hmix=np.full(16, 800.0, np.float32), oli=np.zeros(16, np.float32),
               precip=np.where(np.arange(16) == 2 * 4 + 1, 2.0, 0.0).astype(np.float32))
    wet = dict(dt_seconds=600.0, wet_mode=3, wet_a=1e-4, wet_b=0.8, wet_fraction=1.0)
    p = particle_at(fp, 2.5, 1.5, 500.0)
    orc.deposit(p, wet, pbl)
    lam = 1e-4 * 2.0 ** 0.8
    assert abs(p["mass"][0, 0] - np.exp(-lam * 600.0)) < 1e-6
    q = particle_at(fp, 1.5, 1.5, 500.0)  # dry cell
    orc.deposit(q, wet, pbl)
    assert q["mass"][0, 0] == 1.0


# ------------------------------------------------------------------ full step

@pytest.fixture(scope="module")
def small(workloads):
    return workloads.small_random(3000, shape=(48, 40, 24))


def _params(w, **kw):
    d = dict(dt_seconds=w.dt_seconds, alpha=0.3, x_scale=w.x_scale, y_scale=w.y_scale,
             key=w.key, counter=(5, 0, 0, 0))
    d.update(kw)
    return d


def test_step_conserves_mass_without_deposition_and_leaves_inactive_slots(orc, small):
    p = small.particles.copy()
    m0 = p["mass"].copy()
    orc.step(p, met_dict(small.met), pbl_dict(small.pbl), _params(small))
    assert np.array_equal(p["mass"].view(np.uint32), m0.view(np.uint32))
    inact = small.particles["flags"] == 0
    assert np.array_equal(p[inact].view(np.uint32), small.particles[inact].view(np.uint32))


def test_step_keeps_particles_in_domain_and_inside_the_boundary_layer(orc, small):
    p = small.particles.copy()
    nx, ny, nz = small.met.shape
    for k in range(3):
        orc.step(p, met_dict(small.met), pbl_dict(small.pbl), _params(small, counter=(k * 6000, 0, 0, 0)))
    a = p["flags"] == 1
    assert np.all((p["cell_x"][a] >= 0) & (p["cell_x"][a] <= nx - 1))
    assert np.all((p["pos_x"][a] >= 0) & (p["pos_x"][a] < 1))
    hm = np.maximum(small.pbl.hmix, 1.0)
    i = np.clip(p["cell_x"][a], 0, nx - 1)
    j = np.clip(p["cell_y"][a], 0, ny - 1)
    h = hm[i * ny + j]
    assert np.all(p["pos_z"][a] >= 0.01)
    assert np.all(p["pos_z"][a] <= h * (1 + 1e-6))
    assert np.all(np.isfinite(p["turb_w"][a]))


@pytest.mark.parametrize("n_sub,blocks", [(0, 1), (1, 1), (2, 1), (3, 2), (4, 2)])
def test_step_advances_philox_counter_by_dispatch_times_blocks(orc, small, n_sub, blocks):
    # gpu/langevin_fused.rs:229-233
    p = small.particles.copy()
    nxt = orc.step(p, met_dict(small.met), pbl_dict(small.pbl),
                   _params(small, n_substeps=n_sub, counter=(0xFFFFFFF0, 7, 0, 0)))
    want = orc.counter_add([0xFFFFFFF0, 7, 0, 0], len(p) * blocks)
    assert list(nxt) == list(want)


def test_step_is_deterministic_and_thread_count_independent(orc, small):
    a = small.particles.copy()
    b = small.particles.copy()
    orc.step(a, met_dict(small.met), pbl_dict(small.pbl), _params(small))
    orc.step(b, met_dict(small.met), pbl_dict(small.pbl), _params(small, n_threads=4))
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_step_random_stream_is_indexed_by_slot(orc, small):
    # moving a particle to another slot changes its draws (counter = base + slot*B)
    p = small.particles[:64].copy()
    p["flags"] = 1
    q = p[::-1].copy()
    orc.step(p, met_dict(small.met), pbl_dict(small.pbl), _params(small))
    orc.step(q, met_dict(small.met), pbl_dict(small.pbl), _params(small))
    assert not np.array_equal(p["turb_u"], q["turb_u"][::-1])


def test_deterministic_math_oracle_tracks_the_libm_oracle(orc, small):
    # the fp_math build against the glibc build: same trajectories to well
    # inside the grid-unit tolerance (a last-bit difference can only move a
    # particle by a rounding-sized amount unless it flips a regime branch)
    a = small.particles.copy()
    b = small.particles.copy()
    for k in range(4):
        prm = _params(small, counter=(k * 6000, 0, 0, 0), dry_mode=1, dry_vd=2e-3)
        orc.step(a, met_dict(small.met), pbl_dict(small.pbl), prm)
        orc.step(b, met_dict(small.met), pbl_dict(small.pbl), prm, variant="libm")
    act = a["flags"] == 1
    dx = np.abs((a["cell_x"] + a["pos_x"].astype(np.float64)) - (b["cell_x"] + b["pos_x"].astype(np.float64)))[act]
    dz = np.abs(a["pos_z"].astype(np.float64) - b["pos_z"])[act]
    assert np.quantile(dx, 0.999) < 1e-4
    assert np.quantile(dz, 0.99) < 1e-2
    assert np.mean(dz < 1e-2) > 0.99


def test_level_lookup_known_answers(orc):
    """meters_to_grid_level (advection_dual_wind.wgsl:94-117) at hand-worked
    points: z / h0 below the first level, k - 1 + (z - h[k-1]) / (h[k] -
    h[k-1]) at the first k >= 1 with z <= h[k], the top index above the
    table, the first of a run of equal heights, and z itself in level-index
    mode (top height not positive)."""
    lh = np.array([10.0, 20.0, 40.0, 80.0], np.float32)
    z = np.array([5.0, 10.0, 15.0, 20.0, 30.0, 70.0, 80.0, 90.0, -4.0, np.nan], np.float32)
    want = np.array([0.5, 1.0, 0.5, 1.0, 1.5, 2.75, 3.0, 3.0, -0.4, 3.0], np.float32)
    assert np.array_equal(orc.level_eval(lh, z), want)
    ties = np.array([10.0, 10.0, 20.0, 20.0, 50.0], np.float32)
    z = np.array([10.0, 15.0, 20.0, 35.0], np.float32)
    assert np.array_equal(orc.level_eval(ties, z), np.array([1.0, 1.5, 2.0, 3.5], np.float32))
    # a first level at 0 m: heights at or below it map to level 0
    ground = np.array([0.0, 100.0], np.float32)
    assert np.array_equal(orc.level_eval(ground, np.array([-3.0, 0.0, 50.0], np.float32)),
                          np.array([0.0, 0.0, 0.5], np.float32))
    index = np.zeros(4, np.float32)
    assert np.array_equal(orc.level_eval(index, np.array([1.25, -2.0], np.float32)),
                          np.array([1.25, -2.0], np.float32))
