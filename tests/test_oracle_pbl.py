"""The oracle's PBL-from-surface restatement (io/temporal.rs:122-214 followed by
shaders/pbl_diagnostics.wgsl:220-305), pinned by the reference's own unit
tests (io/pbl_params.rs:704-927, gpu/pbl.rs:482-629, io/temporal.rs tests)
and by the closed forms of the diagnostics evaluated in binary64:

    rho = p / (R T),  u* = sqrt(tau / rho)  or  kappa |U10| / ln(z_ref / z0),
    L = -rho c_p T u*^3 / (kappa g H),  w* = (g/T * H / (rho c_p) * h)^(1/3).
"""
import math

import numpy as np
import pytest

R_AIR, GA, CPA, KAPPA = 287.05, 9.81, 1004.6, 0.4
HMIX_MIN, HMIX_MAX = 100.0, 4500.0


def cell(**kw):
    """One cell's ten surface fields (FP_SURF_* order), zeros by default."""
    names = ("ps", "t2m", "u10", "v10", "stress", "sshf", "ssr", "hmix", "ustar", "oli")
    return [np.array([kw.get(k, 0.0)], np.float32) for k in names]


def run(orc, s, alpha=0.0, s1=None, **opt):
    out = orc.pbl_from_surface(s, s if s1 is None else s1, alpha, opt)
    return {k: v[0] for k, v in out.items()} if len(out["ustar"]) == 1 else out


def closed_form(ps, t2m, u10, v10, stress, sshf, hmix, z0=0.1, zref=10.0):
    rho = ps / (R_AIR * t2m)
    ust = math.sqrt(stress / rho) if stress > 0 else KAPPA * math.hypot(u10, v10) / math.log(zref / z0)
    ust = max(ust, 1e-4)
    ol = -(rho * CPA * t2m * ust ** 3) / (KAPPA * GA * sshf) if abs(sshf) >= 1.0 else math.inf
    oli = 0.0 if (not math.isfinite(ol) or abs(ol) > 1e5) else 1.0 / ol
    h = min(max(hmix if hmix > 0 else 800.0, HMIX_MIN), HMIX_MAX)
    wst = (GA / t2m * sshf / (rho * CPA) * h) ** (1.0 / 3.0) if sshf > 0 else 0.0
    return dict(ustar=ust, wstar=wst, hmix=h, oli=oli)


def test_friction_velocity_from_surface_stress(orc):
    # pbl_params.rs:704-716: tau = 0.48, rho = 1.2 -> u* = sqrt(0.4)
    ps = 1.2 * R_AIR * 300.0
    r = run(orc, cell(ps=ps, t2m=300.0, stress=0.48))
    rho = np.float32(ps) / (np.float32(R_AIR) * np.float32(300.0))
    assert r["ustar"] == pytest.approx(math.sqrt(0.48 / float(rho)), rel=1e-6)
    assert r["ustar"] == pytest.approx(math.sqrt(0.4), abs=1e-6)


def test_friction_velocity_log_law_fallback(orc):
    # pbl_params.rs:718-731: no stress, U10 = 5, z0 = 0.1, z_ref = 10
    r = run(orc, cell(ps=101325.0, t2m=300.0, u10=5.0))
    assert r["ustar"] == pytest.approx(KAPPA * 5.0 / math.log(10.0 / 0.1), abs=1e-6)
    # the log-law uses the wind speed: u and v enter symmetrically
    r2 = run(orc, cell(ps=101325.0, t2m=300.0, u10=3.0, v10=4.0))
    assert r2["ustar"] == pytest.approx(r["ustar"], rel=1e-6)
    # options reach the log-law (z0 = 0.5 m, z_ref = 20 m)
    r3 = run(orc, cell(ps=101325.0, t2m=300.0, u10=5.0), roughness_length_m=0.5, wind_reference_height_m=20.0)
    assert r3["ustar"] == pytest.approx(KAPPA * 5.0 / math.log(20.0 / 0.5), rel=1e-6)
    # zero wind and no stress: the u* floor
    assert run(orc, cell(ps=101325.0, t2m=300.0))["ustar"] == np.float32(1e-4)


def test_obukhov_sign_follows_the_heat_flux(orc):
    # pbl_params.rs:733-761: H > 0 unstable (L < 0), H < 0 stable, H = 0 neutral
    base = dict(ps=101325.0, t2m=300.0, stress=0.4 ** 2 * 1.2)
    assert run(orc, cell(sshf=150.0, **base))["oli"] < 0.0
    assert run(orc, cell(sshf=-150.0, **base))["oli"] > 0.0
    assert run(orc, cell(sshf=0.0, **base))["oli"] == 0.0
    # |H| below the neutral threshold (1 W/m^2) is neutral too
    assert run(orc, cell(sshf=0.9, **base))["oli"] == 0.0
    assert run(orc, cell(sshf=0.9, **base), heat_flux_neutral_threshold_w_m2=0.5)["oli"] < 0.0


def test_gridded_regimes_and_mixing_height_clamps(orc):
    # pbl_params.rs:855-927 without the profile cue (the GPU path has none)
    cells = [dict(ps=101325.0, t2m=300.0, u10=4.0, v10=1.0, stress=0.4, sshf=120.0, ssr=350.0, hmix=1500.0),
             dict(ps=100900.0, t2m=295.0, u10=2.0, v10=1.0, stress=0.2, sshf=-80.0, ssr=50.0, hmix=50.0),
             dict(ps=101100.0, t2m=299.0, u10=3.0, v10=0.0, stress=0.0, sshf=0.0, ssr=100.0, hmix=6000.0),
             dict(ps=101000.0, t2m=301.0, u10=1.0, v10=1.0, stress=0.0, sshf=0.0, ssr=100.0, hmix=0.0)]
    fields = [np.concatenate([cell(**c)[k] for c in cells]) for k in range(10)]
    r = run(orc, fields)
    assert r["oli"][0] < 0 and r["oli"][1] > 0 and r["oli"][2] == 0 and r["oli"][3] == 0
    assert r["hmix"][1] == np.float32(HMIX_MIN)
    assert r["hmix"][2] == np.float32(HMIX_MAX)
    assert r["hmix"][3] == np.float32(800.0)           # fallback mixing height
    assert r["sshf"][0] == 120.0 and r["ssr"][0] == 350.0 and r["surfstr"][0] == np.float32(0.4)
    assert r["wstar"][1] == 0.0 and r["wstar"][2] == 0.0 and r["wstar"][0] > 0.0


# gpu/pbl.rs:482-545: the reference's GPU-vs-CPU fixture (cells (i, j) in
# i-major order) with the closed forms as the expected values
GPU_FIXTURE = [
    dict(ps=101325.0, t2m=300.0, u10=4.0, v10=1.0, stress=0.4, sshf=120.0, ssr=350.0, hmix=1500.0),
    dict(ps=101000.0, t2m=301.0, u10=1.0, v10=1.0, sshf=0.0, ssr=100.0, hmix=0.0),
    dict(ps=100900.0, t2m=295.0, u10=2.0, v10=1.0, stress=0.2, sshf=-80.0, ssr=50.0, hmix=50.0),
    dict(ps=101300.0, t2m=298.0, u10=5.0, v10=2.0, stress=0.3, sshf=50.0, ssr=200.0, hmix=1200.0,
         ustar=0.42, oli=-0.005),
    dict(ps=101100.0, t2m=299.0, u10=3.0, v10=0.0, sshf=0.0, ssr=100.0, hmix=6000.0),
    dict(ps=101000.0, t2m=300.0, u10=0.0, v10=0.0, sshf=10.0, ssr=50.0, hmix=500.0),
]


@pytest.mark.parametrize("k", range(len(GPU_FIXTURE)))
def test_reference_gpu_fixture_matches_closed_forms(orc, k):
    c = GPU_FIXTURE[k]
    r = run(orc, cell(**c))
    want = closed_form(c["ps"], c["t2m"], c.get("u10", 0.0), c.get("v10", 0.0), c.get("stress", 0.0),
                       c["sshf"], c["hmix"])
    if "ustar" in c:  # met-provided u* and 1/L are taken as they are
        want["ustar"], want["oli"] = c["ustar"], c["oli"]
        rho = c["ps"] / (R_AIR * c["t2m"])
        want["wstar"] = (GA / c["t2m"] * c["sshf"] / (rho * CPA) * c["hmix"]) ** (1.0 / 3.0)
    for name in ("ustar", "wstar", "hmix", "oli"):
        assert float(r[name]) == pytest.approx(want[name], rel=2e-6, abs=1e-30), name
    # the reference test's own tolerance against its CPU path
    assert r["ssr"] == np.float32(c["ssr"]) and r["surfstr"] == np.float32(c.get("stress", 0.0))


def test_met_provided_values_and_their_validity_rules(orc):
    base = dict(ps=101325.0, t2m=300.0, u10=4.0, stress=0.3, sshf=100.0, hmix=900.0)
    assert run(orc, cell(ustar=0.35, **base))["ustar"] == np.float32(0.35)
    assert run(orc, cell(ustar=5e-5, **base))["ustar"] == np.float32(1e-4)      # floored
    est = run(orc, cell(**base))["ustar"]
    for bad in (-0.2, np.nan, np.inf):
        assert run(orc, cell(ustar=bad, **base))["ustar"] == est               # not provided
    assert run(orc, cell(oli=-0.02, **base))["oli"] == np.float32(-0.02)
    own = run(orc, cell(**base))["oli"]
    for bad in (5e-6, -9e-6, np.nan, -np.inf):                                 # |1/L| < 1e-5: not provided
        assert run(orc, cell(oli=bad, **base))["oli"] == own


def test_sanitizing_of_invalid_surface_values(orc):
    r = run(orc, cell(ps=np.nan, t2m=-5.0, u10=np.inf, stress=np.nan, sshf=np.nan, ssr=-3.0, hmix=np.inf))
    # p -> 101325, T -> 300, stress -> 0, |U| = inf -> u* = kappa*inf/ln(100) = inf, not positive-finite -> floor
    assert r["ustar"] == np.float32(1e-4)
    assert r["hmix"] == np.float32(800.0) and r["sshf"] == 0.0 and r["ssr"] == 0.0 and r["surfstr"] == 0.0
    assert r["wstar"] == 0.0 and r["oli"] == 0.0


def surface_with_offset(nx, ny, offset):
    """io/temporal.rs test fields: field k = (k+1) * (i + 10 j) + offset."""
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    base = (i + 10.0 * j).reshape(-1)
    return [((k + 1) * base + offset).astype(np.float32) for k in range(10)]


def test_temporal_interpolation_endpoints_are_exact_copies(orc):
    s0, s1 = surface_with_offset(3, 2, 10.0), surface_with_offset(3, 2, 14.0)
    s1_nan = [x.copy() for x in s1]
    s1_nan[5][:] = np.nan                                   # t1 sshf NaN: alpha = 0 must not see it
    at0 = run(orc, s0, 0.0, s1_nan)
    ref0 = run(orc, s0, 0.0, s0)
    for k in at0:
        assert np.array_equal(at0[k], ref0[k]), k
    s0_inf = [x.copy() for x in s0]
    s0_inf[6][:] = np.inf
    at1 = run(orc, s0_inf, 1.0, s1)
    ref1 = run(orc, s1, 0.0, s1)
    for k in at1:
        assert np.array_equal(at1[k], ref1[k]), k


def test_temporal_interpolation_midpoint_and_fma_lerp(orc):
    s0, s1 = surface_with_offset(3, 2, 10.0), surface_with_offset(3, 2, 14.0)
    mid = run(orc, s0, 0.5, s1)
    # sshf (field 5), ssr (6), stress (4) pass through sanitized: exact midpoints
    assert np.array_equal(mid["sshf"], surface_with_offset(3, 2, 12.0)[5])
    assert np.array_equal(mid["ssr"], surface_with_offset(3, 2, 12.0)[6])
    assert np.array_equal(mid["surfstr"], surface_with_offset(3, 2, 12.0)[4])
    # a general weight: start.mul_add(1 - alpha, end * alpha) (temporal.rs:300-303), one rounding of the fma
    rng = np.random.default_rng(3)
    a = rng.uniform(-300, 300, 64).astype(np.float32)
    b = rng.uniform(-300, 300, 64).astype(np.float32)
    z = [np.zeros(64, np.float32)] * 10
    t0 = list(z); t0[5] = a
    t1 = list(z); t1[5] = b
    al = np.float32(0.3)
    got = orc.pbl_from_surface(t0, t1, float(al))["sshf"]
    from fractions import Fraction
    for x, y, g in zip(a, b, got):
        exact = Fraction(float(x)) * Fraction(float(np.float32(1) - al)) + Fraction(float(np.float32(y * al)))
        assert g == np.float32(float(exact)) or abs(float(g) - float(exact)) <= abs(float(exact)) * 2 ** -24
