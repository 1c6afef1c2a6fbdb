"""PBL state from the resident surface bracket on the GPU (fp_surface_upload +
fp_pbl_from_surface) against the oracle's restatement of
interpolate_surface_fields_linear (io/temporal.rs:122-214) and
pbl_diagnostics.wgsl:220-305: bit-exact u*, w*, hmix, 1/L and pass-through
fields, and particle steps driven by it identical to steps driven by the
oracle's PBL uploaded with fp_pbl_upload."""
import sys

import numpy as np
import pytest

from conftest import assert_particles_equal, met_dict

pytestmark = pytest.mark.gpu

T0, T1 = 21_600, 43_200


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def surface_fields(n, seed, specials=True):
    """Ten surface fields covering all branches of the diagnostics: stress or
    log-law u*, provided or computed u* and 1/L, missing and clamped hmix,
    both flux signs, and non-finite or negative values in every field."""
    r = np.random.default_rng(seed)
    f = [r.uniform(9.0e4, 1.05e5, n),                                   # pressure
         r.uniform(250.0, 310.0, n),                                    # T2m
         r.normal(0.0, 5.0, n), r.normal(0.0, 5.0, n),                  # u10, v10
         np.where(r.uniform(size=n) < 0.5, 0.0, r.uniform(0.0, 1.0, n)),   # stress
         r.normal(0.0, 120.0, n),                                       # sshf
         r.uniform(0.0, 800.0, n),                                      # ssr
         np.where(r.uniform(size=n) < 0.3, 0.0, 800.0 * np.exp(r.normal(0, 1.2, n))),  # hmix
         np.where(r.uniform(size=n) < 0.5, 0.0, r.uniform(0.02, 1.0, n)),  # ustar
         np.where(r.uniform(size=n) < 0.5, 0.0, r.uniform(-0.05, 0.05, n))]  # 1/L
    f = [x.astype(np.float32) for x in f]
    if specials:
        bad = np.array([np.nan, np.inf, -np.inf, -1.0, 0.0, 1e-38, 3e38], np.float32)
        for k in range(10):
            idx = r.choice(n, size=len(bad) * 3, replace=False)
            f[k][idx] = np.tile(bad, 3)
    return f


def oracle_pbl(orc, s0, s1, alpha, opts=None):
    return orc.pbl_from_surface(s0, s1, alpha, opts)


@pytest.fixture(scope="module")
def small(workloads):
    return workloads.small_random(20_000, shape=(48, 40, 24))


def make_ctx(fp, w):
    ctx = fp.GpuTimestep(len(w.particles), w.met.shape, n_species=4, out_shape=w.out_shape)
    u0, v0, w0 = w.met.levels[0]
    u1, v1, w1 = w.met.levels[1]
    ctx.upload_met_bracket(u0, v0, w0, u1, v1, w1, w.met.level_heights)
    ctx.upload_store(w.particles)
    return ctx


@pytest.mark.parametrize("now,strict", [(T0, True), (T1, True), (T0 + 7_777, True), (T0 - 100, False),
                                        (T1 + 5, False)])
def test_pbl_from_surface_matches_oracle_bitwise(fp, orc, small, now, strict):
    say(f"pbl from surface: now={now} strict={strict}")
    import torch
    nx, ny = small.met.shape[:2]
    n = nx * ny
    s0, s1 = surface_fields(n, 1), surface_fields(n, 2)
    alpha = fp.interpolation_alpha(T0, T1, now, strict=False)
    want = oracle_pbl(orc, s0, s1, alpha)
    with make_ctx(fp, small) as ctx:
        ctx.upload_surface(0, s0, T0)
        ctx.upload_surface(1, s1, T1)
        pas = torch.full((3 * n,), -1.0, dtype=torch.float32, device="cuda")
        ctx.pbl_from_surface(fp.PblOptions(), now, strict=strict, passthrough=pas)
        ctx.sync()
        got = ctx.download_pbl()
        pas = pas.cpu().numpy().reshape(3, n)
    for k in ("ustar", "wstar", "hmix", "oli"):
        assert np.array_equal(got[k].view(np.uint32), want[k].view(np.uint32)), k
    for i, k in enumerate(("sshf", "ssr", "surfstr")):
        assert np.array_equal(pas[i].view(np.uint32), want[k].view(np.uint32)), k
    assert np.isfinite(got["hmix"]).all() and (got["ustar"] >= np.float32(1e-4)).all()


def test_options_and_device_upload_path(fp, orc, small):
    say("pbl from surface: non-default options, device-pointer upload")
    import torch
    nx, ny = small.met.shape[:2]
    n = nx * ny
    s0, s1 = surface_fields(n, 5), surface_fields(n, 6)
    opts = dict(roughness_length_m=0.03, wind_reference_height_m=12.0, heat_flux_neutral_threshold_w_m2=5.0,
                fallback_mixing_height_m=1200.0, hmix_min_m=50.0, hmix_max_m=3000.0)
    want = oracle_pbl(orc, s0, s1, fp.interpolation_alpha(T0, T1, T0 + 3_600), opts)
    with make_ctx(fp, small) as ctx:
        ctx.upload_surface(0, [torch.from_numpy(x).cuda() for x in s0], T0)
        ctx.upload_surface(1, [torch.from_numpy(x).cuda() for x in s1], T1)
        ctx.pbl_from_surface(fp.PblOptions(**opts), T0 + 3_600)
        got = ctx.download_pbl()
    for k in ("ustar", "wstar", "hmix", "oli"):
        assert np.array_equal(got[k].view(np.uint32), want[k].view(np.uint32)), k


def test_steps_driven_by_surface_pbl_equal_uploaded_oracle_pbl(fp, orc, small):
    """The step kernel reads the PBL records written by the surface kernel
    exactly as those packed from fp_pbl_upload of the oracle's fields."""
    say("pbl from surface: 3 steps vs fp_pbl_upload of the oracle's PBL and vs the oracle step")
    nx, ny = small.met.shape[:2]
    n = nx * ny
    s0, s1 = surface_fields(n, 8, specials=False), surface_fields(n, 9, specials=False)
    from test_gpu_parity import base_cfg, oracle_params
    states = []
    want = small.particles.copy()
    for mode in ("surface", "upload"):
        with make_ctx(fp, small) as ctx:
            ctx.upload_surface(0, s0, T0)
            ctx.upload_surface(1, s1, T1)
            counter = (11, 0, 0, 0)
            for k in range(3):
                now = T0 + 600 * k
                o = oracle_pbl(orc, s0, s1, fp.interpolation_alpha(T0, T1, now))
                if mode == "surface":
                    ctx.pbl_from_surface(None, now)
                else:
                    ctx.upload_pbl(o["ustar"], o["wstar"], o["hmix"], o["oli"])
                cfg = base_cfg(fp, small, counter=counter, dry_mode=1, dry_vd=2e-3)
                if mode == "surface":
                    pbl = dict(shape=(nx, ny), ustar=o["ustar"], wstar=o["wstar"], hmix=o["hmix"], oli=o["oli"])
                    orc.step(want, met_dict(small.met), pbl, oracle_params(small, cfg))
                counter = ctx.step(cfg)
            states.append(ctx.download_particles())
    assert_particles_equal(states[0], states[1], "surface PBL vs uploaded PBL")
    assert_particles_equal(states[0], want, "surface PBL vs oracle")


def test_call_order_and_time_errors(fp, small):
    say("pbl from surface: errors")
    nx, ny = small.met.shape[:2]
    n = nx * ny
    s = surface_fields(n, 3, specials=False)
    with make_ctx(fp, small) as ctx:
        with pytest.raises(fp.FpError):
            ctx.pbl_from_surface(None, T0)           # no surface bracket yet
        with pytest.raises(fp.FpError):
            ctx.download_pbl()                       # no PBL state yet
        ctx.upload_surface(0, s, T0)
        with pytest.raises(fp.FpError):
            ctx.pbl_from_surface(None, T0)           # t1 missing
        ctx.upload_surface(1, s, T1)
        with pytest.raises(fp.FpError):
            ctx.pbl_from_surface(None, T1 + 1, strict=True)
        with pytest.raises(fp.FpError):
            ctx.upload_surface(2, s, T1)
        with pytest.raises(ValueError):
            ctx.upload_surface(0, s[:9], T0)
        ctx.upload_surface(1, s, T0)                 # t1 == t0: invalid bracket
        with pytest.raises(fp.FpError):
            ctx.pbl_from_surface(None, T0)
        ctx.upload_surface(1, s, T1)
        ctx.pbl_from_surface(None, T1 + 1)           # clamped
        assert np.isfinite(ctx.download_pbl()["ustar"]).all()
