"""Closed-form pins for the step pieces that the reference's own tests leave open.

The reference checks PBL reflection only through the invariant z in [z_min, h]
(tests/integration/physics_validation.rs) and never evaluates the unstable
vertical profile's height derivative, its three T_Lw branches or the large-ratio
drift term of the vertical Ornstein-Uhlenbeck update.  Each test here derives
the expected value from the mathematics of the formula in
langevin_fused.wgsl (mirror images, the analytic derivative of sigma_w(z),
the branch formulas, the exact OU solution) in binary64, independently of the
oracle's float32 code, so that a dropped term, a wrong sign or coefficient, or
a swapped branch in the oracle fails one of them:

- reflection (langevin_fused.wgsl:481-500): position after k mirror images
  folds z into [z_min, h] with period 2 (h - z_min); turb_w and cbt change sign
  once per image; at most 10 images, then the clamp;
- unstable sigma_w and dsigma_w/dz (:252-262): the closed form of sigma_w and a
  central difference of the oracle's own sigma_w(z), with the reference's
  0.5 / (sigma_w + 0.01 floor) / h factor (an inherited quirk: the derivative of
  sqrt(S) is 0.5 S' / sqrt(S), the reference divides by the floored sigma_w);
- stable dsigma_w/dz = -1.3 u* / h, the exact derivative of its sigma_w(z);
- unstable T_Lw in each of its three branches (:318-327);
- the vertical OU update in both branches, with noise, memory and drift
  (:366-389), the r >= 0.5 branch's T_L (1 - e^{-r}) drift term included;
- hanna_short_w equal to compute_hanna's vertical part (:296-344 vs :224-293);
- the wiring of the Philox/Box-Muller normals into the step (:430-451,461-515):
  eta_u, eta_v from block 0 words (x, y), the sub-step normals from block 0
  (z, w) and block 1 (x, y).
"""
import math

import numpy as np
import pytest

F32 = np.float32


def f32(x):
    return float(np.float32(x))


# ------------------------------------------------------------------ reflection

def folded(z, zmin, h):
    """Mirror images of z about z_min and h: (final z, number of images)."""
    span = h - zmin
    k = math.floor((z - zmin) / span)
    images = k if k > 0 else -k
    y = (z - zmin) % (2.0 * span)
    zf = zmin + (y if y <= span else 2.0 * span - y)
    return zf, images


REFLECT_CASES = [
    # (z, z_min, h): one image below, one above, two, four, nine, exactly ten
    (-5.0, 0.01, 1000.0),
    (1100.0, 0.01, 1000.0),
    (-1500.0, 0.01, 1000.0),
    (450.0, 0.01, 100.0),
    (0.01 + 9.5 * 99.99, 0.01, 100.0),
    (0.01 + 10.5 * 99.99, 0.01, 100.0),
    (-(9.25 * 499.95) + 0.01, 0.01, 500.0),
    (0.004, 0.01, 2000.0),
    (2000.5, 0.01, 2000.0),
]


@pytest.mark.parametrize("z,zmin,h", REFLECT_CASES)
def test_reflection_mirrors_position_and_flips_signs_per_image(orc, z, zmin, h):
    zf, images = folded(z, zmin, h)
    assert 1 <= images <= 10
    for tw, cbt in [(2.5, 3), (-0.75, -1)]:
        got_z, got_tw, got_cbt = orc.reflect(f32(z), tw, cbt, zmin, h)
        sign = -1 if images % 2 else 1
        assert got_tw == sign * tw, (z, images)
        assert got_cbt == sign * cbt, (z, images)
        assert abs(got_z - zf) <= 2e-6 * max(abs(z), h), (z, got_z, zf)
        assert zmin <= got_z <= h


@pytest.mark.parametrize("z,zmin,h", [(0.01 + 11.5 * 99.99, 0.01, 100.0),
                                      (0.01 - 10.5 * 99.99, 0.01, 100.0),
                                      (1.0e6, 0.01, 100.0)])
def test_reflection_stops_after_ten_images_and_clamps(orc, z, zmin, h):
    # more than ten images needed: ten are taken (even: signs restored), then the
    # position is clamped to the boundary it is still beyond
    _, images = folded(z, zmin, h)
    assert images > 10
    got_z, got_tw, got_cbt = orc.reflect(f32(z), 1.5, 7, zmin, h)
    assert got_tw == 1.5 and got_cbt == 7
    z10 = z - 10 * (h - zmin) if z > h else z + 10 * (h - zmin)
    assert got_z == f32(h if z10 > h else zmin)


def test_reflection_inside_the_layer_and_degenerate_layer_are_no_ops(orc):
    assert orc.reflect(50.0, 1.25, -2, 0.01, 100.0) == (50.0, 1.25, -2)
    assert orc.reflect(f32(0.01), 1.25, -2, 0.01, 100.0) == (f32(0.01), 1.25, -2)
    assert orc.reflect(100.0, 1.25, -2, 0.01, 100.0) == (100.0, 1.25, -2)
    # h <= z_min: no reflection and no clamp at all
    assert orc.reflect(-3.0, 1.25, -2, 0.01, 0.01) == (-3.0, 1.25, -2)
    assert orc.reflect(7.0, 1.25, -2, 0.5, 0.25) == (7.0, 1.25, -2)


# ------------------------------------------------------------------ Hanna vertical profile

UNSTABLE = [
    # (u*, w*, L, h)
    (0.3, 2.0, -80.0, 1000.0),
    (0.6, 1.2, -25.0, 2500.0),
    (0.15, 0.6, -400.0, 800.0),
]


def s_unstable(ust, wst, zeta):
    return 1.2 * wst ** 2 * (1 - 0.9 * zeta) * zeta ** (2 / 3) + (1.8 - 1.4 * zeta) * ust ** 2


@pytest.mark.parametrize("ust,wst,ol,h", UNSTABLE)
def test_unstable_sigma_w_closed_form(orc, ust, wst, ol, h):
    for zeta in (0.0, 0.02, 0.1, 0.37, 0.8, 1.0):
        p = orc.hanna(ust, wst, ol, h, zeta * h)
        want = math.sqrt(s_unstable(ust, wst, p["zeta"])) + 0.01
        assert abs(p["sigw"] - want) <= 2e-6 * want, zeta


@pytest.mark.parametrize("ust,wst,ol,h", UNSTABLE)
def test_unstable_dsigw_dz_is_the_derivative_with_the_floored_denominator(orc, ust, wst, ol, h):
    for zeta in (0.02, 0.1, 0.25, 0.5, 0.75, 0.95):
        z = zeta * h
        p = orc.hanna(ust, wst, ol, h, z)
        # analytic: dS/dz = (w*^2 (0.8 zeta^-1/3 - 1.8 zeta^2/3) - 1.4 u*^2) / h,
        # reference: dsigw/dz = 0.5 dS/dz / sigma_w(floored)
        ds = (wst ** 2 * (0.8 * zeta ** (-1 / 3) - 1.8 * zeta ** (2 / 3)) - 1.4 * ust ** 2) / h
        want = 0.5 * ds / (math.sqrt(s_unstable(ust, wst, zeta)) + 0.01)
        assert abs(p["dsigwdz"] - want) <= 1e-4 * abs(want) + 1e-9, zeta
        # and against a central difference of the oracle's own sigma_w(z)
        dz = 2e-3 * h
        fd = (orc.hanna(ust, wst, ol, h, z + dz)["sigw"] - orc.hanna(ust, wst, ol, h, z - dz)["sigw"]) / (2 * dz)
        raw = p["sigw"] - 0.01
        assert abs(p["dsigwdz"] * p["sigw"] / raw - fd) <= 2e-3 * abs(fd) + 2e-7, zeta


def test_unstable_dsigw_dz_floors_zeta_in_the_inverse_cube_root_only(orc):
    ust, wst, ol, h = 0.3, 2.0, -80.0, 1000.0
    for zeta in (2e-4, 5e-4, 9e-4):
        p = orc.hanna(ust, wst, ol, h, zeta * h)
        z = p["zeta"]
        ds = (wst ** 2 * (0.8 * 10.0 - 1.8 * z ** (2 / 3)) - 1.4 * ust ** 2) / h  # 1e-3^(-1/3) = 10
        want = 0.5 * ds / (math.sqrt(s_unstable(ust, wst, z)) + 0.01)
        assert abs(p["dsigwdz"] - want) <= 1e-4 * abs(want)


@pytest.mark.parametrize("ust,ol,h", [(0.4, 100.0, 1000.0), (0.2, 30.0, 600.0)])
def test_stable_dsigw_dz_is_the_exact_derivative(orc, ust, ol, h):
    for zeta in (0.05, 0.3, 0.7):
        z = zeta * h
        p = orc.hanna(ust, 0.0, ol, h, z)
        assert abs(p["dsigwdz"] - (-1.3 * ust / h)) <= 1e-6 * 1.3 * ust / h
        dz = 1e-2 * h
        fd = (orc.hanna(ust, 0.0, ol, h, z + dz)["sigw"] - orc.hanna(ust, 0.0, ol, h, z - dz)["sigw"]) / (2 * dz)
        assert abs(p["dsigwdz"] - fd) <= 1e-3 * abs(fd)


def tlw_unstable(ust, wst, ol, h, z):
    zeta = min(max(z / h, 0.0), 1.0)
    sw = math.sqrt(s_unstable(ust, wst, zeta)) + 0.01
    if z < abs(ol):
        t = 0.1 * z / (sw * (0.55 - 0.38 * abs(z / ol)))
    elif zeta < 0.1:
        t = 0.59 * z / sw
    else:
        t = 0.15 * h / sw * (1.0 - math.exp(-5.0 * zeta))
    return max(t, 30.0), ("a" if z < abs(ol) else ("b" if zeta < 0.1 else "c"))


@pytest.mark.parametrize("ust,wst,ol,h,z,branch", [
    (0.3, 1.5, -500.0, 2000.0, 300.0, "a"),   # z < |L|
    (0.3, 1.5, -500.0, 2000.0, 480.0, "a"),
    (0.3, 1.5, -50.0, 2000.0, 150.0, "b"),    # |L| <= z, zeta < 0.1
    (0.5, 2.5, -20.0, 3000.0, 250.0, "b"),
    (0.3, 1.5, -50.0, 2000.0, 1000.0, "c"),   # |L| <= z, zeta >= 0.1
    (0.4, 1.0, -100.0, 1500.0, 1400.0, "c"),
])
def test_unstable_tlw_each_branch_closed_form(orc, ust, wst, ol, h, z, branch):
    want, b = tlw_unstable(ust, wst, ol, h, z)
    assert b == branch and want > 30.0
    for got in (orc.hanna(ust, wst, ol, h, z)["tlw"], orc.hanna_short_w(ust, wst, ol, h, z)["tlw"]):
        assert abs(got - want) <= 5e-6 * want


def test_hanna_short_w_equals_the_vertical_part_of_compute_hanna(orc):
    rng = np.random.default_rng(11)
    for _ in range(3000):
        ust = f32(rng.uniform(0.0, 0.9))
        wst = f32(rng.uniform(0.0, 3.0))
        ol = f32(rng.choice([-1, 1]) * 10 ** rng.uniform(0.5, 4.0))
        h = f32(10 ** rng.uniform(0.0, 3.6))
        z = f32(rng.uniform(-10.0, 1.2 * h))
        full = orc.hanna(ust, wst, ol, h, z)
        short = orc.hanna_short_w(ust, wst, ol, h, z)
        assert (short["sigw"], short["dsigwdz"], short["tlw"]) == (full["sigw"], full["dsigwdz"], full["tlw"])


# ------------------------------------------------------------------ vertical OU

@pytest.mark.parametrize("prev,sigw,dsw,tlw,dt,rho,eta", [
    (0.0, 0.8, 2e-3, 40.0, 150.0, 0.0, 0.0),      # r = 3.75: drift only
    (0.0, 0.8, -1e-3, 60.0, 150.0, -1.2e-4, 0.0),  # r = 2.5, density gradient
    (0.35, 1.1, 5e-4, 100.0, 60.0, 0.0, 0.7),     # r = 0.6: memory, noise, drift
    (-0.4, 0.6, 1e-3, 30.0, 15.0, -1e-4, -1.3),   # r = 0.5 exactly: exact branch
])
def test_vertical_ou_large_ratio_branch_with_drift(orc, prev, sigw, dsw, tlw, dt, rho, eta):
    r = dt / tlw
    assert r >= 0.5
    e = math.exp(-r)
    drift = sigw * (dsw + rho * sigw)
    want = e * prev + eta * sigw * math.sqrt(1.0 - e * e) + tlw * (1.0 - e) * drift
    got = orc.update_vertical(prev, sigw, dsw, tlw, dt, rho, eta)
    assert abs(got - want) <= 2e-6 * (abs(prev) + abs(eta) * sigw + tlw * abs(drift))
    # the drift term's own contribution (no memory, no noise)
    d = orc.update_vertical(0.0, sigw, dsw, tlw, dt, rho, 0.0)
    assert abs(d - tlw * (1.0 - e) * drift) <= 2e-6 * tlw * abs(drift)


@pytest.mark.parametrize("prev,sigw,dsw,tlw,dt,rho,eta", [
    (0.35, 1.1, 5e-4, 100.0, 30.0, 0.0, 0.7),
    (-0.4, 0.6, 1e-3, 300.0, 15.0, -1e-4, -1.3),
])
def test_vertical_ou_small_ratio_branch_with_drift(orc, prev, sigw, dsw, tlw, dt, rho, eta):
    r = dt / tlw
    assert r < 0.5
    drift = sigw * (dsw + rho * sigw)
    want = (1.0 - r) * prev + eta * sigw * math.sqrt(2.0 * r) + dt * drift
    got = orc.update_vertical(prev, sigw, dsw, tlw, dt, rho, eta)
    assert abs(got - want) <= 2e-6 * (abs(prev) + abs(eta) * sigw + dt * abs(drift))


# ------------------------------------------------------------------ normals into the step

def test_step_draws_the_normals_from_the_documented_philox_words(orc, fp):
    """One particle, no wind, 4 sub-steps: the step equals the composition of the
    oracle's primitives with eta_u, eta_v <- block 0 (x, y), sub-step normals
    <- block 0 (z, w), block 1 (x, y), and the sub-step order OU -> move ->
    reflect -> profile at the new height (not after the last)."""
    nx, ny, nz = 8, 8, 4
    n = nx * ny * nz
    zero = np.zeros(n, np.float32)
    met = dict(shape=(nx, ny, nz), u0=zero, v0=zero, w0=zero, u1=zero, v1=zero, w1=zero,
               level_heights=np.array([100.0, 500.0, 2000.0, 8000.0], np.float32))
    ust, wst, oli, hmix = 0.35, 1.4, -1.0 / 60.0, 900.0
    pbl = dict(shape=(nx, ny), ustar=np.full(nx * ny, ust, np.float32),
               wstar=np.full(nx * ny, wst, np.float32), hmix=np.full(nx * ny, hmix, np.float32),
               oli=np.full(nx * ny, oli, np.float32))
    dt, key, ctr = 600.0, (0x1234, 0xABCD), (7, 0, 0, 0)
    for slot, z0, tw0 in [(0, 40.0, 0.3), (3, 850.0, -0.2), (5, 5.0, -1.5)]:
        p = np.zeros(6, fp.PARTICLE_DTYPE)
        p["flags"] = 1
        p["cell_x"], p["cell_y"] = 3, 4
        p["pos_x"], p["pos_y"] = 0.5, 0.25
        p["pos_z"] = z0
        p["turb_u"], p["turb_v"], p["turb_w"] = 0.1, -0.05, tw0
        p["cbt"] = 1
        p["mass"][:, 0] = 1.0
        got = p.copy()
        orc.step(got, met, pbl, dict(dt_seconds=dt, key=key, counter=ctr, n_substeps=4))

        b0 = orc.philox4x32_10(orc.counter_add(ctr, 2 * slot), key)
        b1 = orc.philox4x32_10(orc.counter_add(ctr, 2 * slot + 1), key)
        u = [orc.u32_to_uniform01(int(v)) for v in (*b0, *b1)]
        eu, ev = orc.box_muller(u[0], u[1])
        ew = [*orc.box_muller(u[2], u[3]), *orc.box_muller(u[4], u[5])]
        ol = F32(1.0) / F32(oli)
        hn = orc.hanna(ust, wst, ol, hmix, z0)
        tu = orc.update_horizontal(0.1, hn["sigu"], hn["tlu"], dt, float(eu))
        tv = orc.update_horizontal(-0.05, hn["sigv"], hn["tlv"], dt, float(ev))
        z, tw, cbt = F32(z0), tw0, 1
        sw, dsw, tl = hn["sigw"], hn["dsigwdz"], hn["tlw"]
        dts = dt / 4
        for sub in range(4):
            tw = orc.update_vertical(tw, sw, dsw, tl, dts, 0.0, float(ew[sub]))
            z = float(F32(z) + F32(tw) * F32(dts))
            z, tw, cbt = orc.reflect(z, tw, cbt, 0.01, max(hmix, 1.0))
            if sub < 3:
                s = orc.hanna_short_w(ust, wst, ol, hmix, z)
                sw, dsw, tl = s["sigw"], s["dsigwdz"], s["tlw"]
        q = got[slot]
        assert (float(q["turb_u"]), float(q["turb_v"])) == (tu, tv)
        assert (float(q["turb_w"]), float(q["pos_z"]), int(q["cbt"])) == (tw, z, cbt)
