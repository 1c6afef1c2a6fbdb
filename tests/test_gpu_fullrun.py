"""The production run end to end: the global workload's 24 h (144 steps of
600 s) with met bracket switches every 6 h, fused gridding cleared every hour
and dry deposition, on the device through fp_step, against the oracle's step
sequence on the same particles: every record, the Philox counter and the last
hour's concentration grid bit for bit.

The default size (1M particles) keeps the oracle to seconds on 16 threads;
FP_FULLRUN_PARTICLES=10000000 runs the 10M configuration itself (about 45 s
of oracle time on 16 cores)."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = int(os.environ.get("FP_FULLRUN_PARTICLES", "1000000"))
STEPS = 144


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("ppt", [1, 2])
def test_24h_run_matches_oracle_bitwise(fp, orc, workloads, ppt):
    import torch
    from test_gpu_parity import oracle_params

    say(f"24 h production run: {N} particles, {STEPS} steps, {ppt} particle(s) per thread, device vs oracle")
    w = workloads.global_10m(N)
    dev = torch.device("cuda:0")
    levels = [[torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in lvl] for lvl in w.met.levels]
    torch.cuda.synchronize()
    sph = int(round(3600.0 / w.dt_seconds))
    dtm = int(w.met.seconds_between_levels)

    def cfg_at(k, counter):
        t = int(k * w.dt_seconds)
        b = w.alpha(t)[0]
        return b, fp.StepConfig(dt_seconds=w.dt_seconds, alpha=fp.interpolation_alpha(b * dtm, (b + 1) * dtm, t),
                                x_scale=w.x_scale, y_scale=w.y_scale, key=w.key, counter=counter,
                                dry_mode=1, dry_vd=2e-3, grid_fused=True)

    with fp.GpuTimestep(N, w.met.shape, n_species=1, out_shape=w.out_shape) as ctx:
        ctx.upload_pbl(w.pbl.ustar, w.pbl.wstar, w.pbl.hmix, w.pbl.oli)
        ctx.configure_grid(w.outheights, w.refine, 0, w.mass_scale)
        ctx.set_launch(256, ppt)  # the single-row kernel and the row loop
        ctx.upload_store(w.particles)
        bracket, counter = -1, (0, 0, 0, 0)
        for k in range(STEPS):
            b, cfg = cfg_at(k, counter)
            if b != bracket:  # upload_dual_wind_if_bracket_changed (timeloop.rs:1378-1402)
                ptrs = [levels[b][c].data_ptr() for c in range(3)] + [levels[b + 1][c].data_ptr() for c in range(3)]
                ctx.upload_met_bracket_dev(ptrs, w.met.level_heights)
                bracket = b
            if k % sph == 0:
                ctx.clear_grid()
            counter = ctx.step(cfg)
        got = ctx.download_particles()
        gcounts, gmass = ctx.download_grid()
    del levels

    parts = w.particles.copy()
    pbl = dict(shape=w.pbl.shape, ustar=w.pbl.ustar, wstar=w.pbl.wstar, hmix=w.pbl.hmix, oli=w.pbl.oli)
    ocounter = (0, 0, 0, 0)
    oc = om = None
    threads = max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))))
    for k in range(STEPS):
        b, cfg = cfg_at(k, ocounter)
        u0, v0, w0 = w.met.levels[b]
        u1, v1, w1 = w.met.levels[b + 1]
        met = dict(shape=w.met.shape, u0=u0, v0=v0, w0=w0, u1=u1, v1=v1, w1=w1, level_heights=w.met.level_heights)
        p = oracle_params(w, cfg)
        p["n_threads"] = threads
        ocounter = tuple(int(v) for v in orc.step(parts, met, pbl, p))
        if k // sph == (STEPS - 1) // sph:  # the last hour, as the device grid after its hourly clears
            oc, om = orc.grid_accumulate(parts, w.out_shape, w.outheights, w.refine, 0, w.mass_scale, oc, om)
        if k % 48 == 47:
            say(f"  oracle step {k + 1}")
    assert tuple(int(v) for v in counter) == ocounter
    differ = np.count_nonzero(np.any(got.view(np.uint32).reshape(N, -1) != parts.view(np.uint32).reshape(N, -1),
                                     axis=1))
    assert differ == 0, f"{differ} of {N} records differ after {STEPS} steps"
    assert int(oc.astype(np.int64).sum()) == N * sph
    assert np.array_equal(gcounts, oc) and np.array_equal(gmass, om)
