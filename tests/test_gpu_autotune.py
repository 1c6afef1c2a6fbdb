"""The autotuner picks a valid launch shape and re-sort interval without
touching the simulation: state and grid after tuning are those before it."""
import os

import numpy as np
import pytest

from conftest import assert_particles_equal, met_dict, pbl_dict

pytestmark = pytest.mark.gpu


def test_autotune_preserves_state_and_caches(fp, orc, workloads, tmp_path):
    w = workloads.small_random(40_000, shape=(48, 40, 24))
    u0, v0, w0 = w.met.levels[0]
    u1, v1, w1 = w.met.levels[1]
    with fp.GpuTimestep(len(w.particles), w.met.shape, out_shape=w.out_shape) as ctx:
        ctx.upload_met_bracket(u0, v0, w0, u1, v1, w1, w.met.level_heights)
        ctx.upload_pbl(w.pbl.ustar, w.pbl.wstar, w.pbl.hmix, w.pbl.oli, w.pbl.precip)
        ctx.upload_store(w.particles)
        ctx.configure_grid(w.outheights, w.refine, 0, w.mass_scale)
        ctx.clear_grid()
        cfg = fp.StepConfig(dt_seconds=w.dt_seconds, alpha=0.3, x_scale=w.x_scale,
                            y_scale=w.y_scale, key=w.key, dry_mode=1, dry_vd=2e-3, grid_fused=True)
        ctx.step(cfg)
        before = ctx.download_particles()
        gc, gm = ctx.download_grid()
        cache = os.path.join(str(tmp_path), "tune.json")
        r = ctx.autotune(cfg, sort_intervals=(0, 2, 5), warmup_steps=1, measured_steps=3,
                         sort_trial_steps=6, cache_path=cache)
        assert r["block_size"] in (64, 128, 256, 512)
        assert r["particles_per_thread"] in (1, 2, 4)
        assert r["sort_interval"] in (0, 2, 5)
        assert r["step_ms"] > 0 and not r["from_cache"]
        assert_particles_equal(ctx.download_particles(), before, "state after autotune")
        gc2, gm2 = ctx.download_grid()
        assert np.array_equal(gc, gc2) and np.array_equal(gm, gm2)
        r2 = ctx.autotune(cfg, sort_intervals=(0, 2, 5), cache_path=cache)
        assert r2["from_cache"] and r2["block_size"] == r["block_size"]
        assert r2["sort_interval"] == r["sort_interval"]
        # the chosen launch shape computes the same step as the oracle
        cfg.counter = (12345, 0, 0, 0)
        ctx.step(cfg)
        want = before.copy()
        orc.step(want, met_dict(w.met), pbl_dict(w.pbl), dict(
            dt_seconds=cfg.dt_seconds, alpha=cfg.alpha, x_scale=cfg.x_scale, y_scale=cfg.y_scale,
            key=cfg.key, counter=cfg.counter, dry_mode=1, dry_vd=2e-3))
        assert_particles_equal(ctx.download_particles(), want, "step after autotune")


@pytest.mark.parametrize("block,ppt", [(64, 1), (128, 2), (512, 4), (1024, 1), (256, 16)])
def test_every_launch_shape_computes_the_same_step(fp, orc, workloads, block, ppt):
    w = workloads.small_random(30_001, shape=(40, 32, 20))
    u0, v0, w0 = w.met.levels[0]
    u1, v1, w1 = w.met.levels[1]
    with fp.GpuTimestep(len(w.particles), w.met.shape) as ctx:
        ctx.set_launch(block, ppt)
        ctx.upload_met_bracket(u0, v0, w0, u1, v1, w1, w.met.level_heights)
        ctx.upload_pbl(w.pbl.ustar, w.pbl.wstar, w.pbl.hmix, w.pbl.oli)
        ctx.upload_store(w.particles)
        cfg = fp.StepConfig(dt_seconds=w.dt_seconds, alpha=0.6, x_scale=w.x_scale,
                            y_scale=w.y_scale, key=w.key, counter=(7, 0, 0, 0))
        ctx.step(cfg)
        want = w.particles.copy()
        orc.step(want, met_dict(w.met), pbl_dict(w.pbl), dict(
            dt_seconds=cfg.dt_seconds, alpha=cfg.alpha, x_scale=cfg.x_scale, y_scale=cfg.y_scale,
            key=cfg.key, counter=cfg.counter))
        assert_particles_equal(ctx.download_particles(), want, f"launch {block}x{ppt}")
