"""The 10M-particle global workload on the GPU.

The oracle cannot step 10M particles in seconds, but every particle-step is
independent and its random stream is indexed by its slot, so the oracle
checks a sample of slots exactly: the same state array with every other slot
deactivated gives the sampled slots the same counters and the same inputs.
The whole run is checked through size-independent properties.
"""
import sys

import numpy as np
import pytest

from conftest import assert_particles_equal, met_dict, pbl_dict

pytestmark = pytest.mark.gpu

N = 10_000_000


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def big(workloads):
    return workloads.global_10m(N)


def run(fp, w, n_steps, sort_every=0, grid=True, tag=False):
    ctx = fp.GpuTimestep(N, w.met.shape, n_species=1, out_shape=w.out_shape)
    u0, v0, w0 = w.met.levels[0]
    u1, v1, w1 = w.met.levels[1]
    ctx.upload_met_bracket(u0, v0, w0, u1, v1, w1, w.met.level_heights)
    ctx.upload_pbl(w.pbl.ustar, w.pbl.wstar, w.pbl.hmix, w.pbl.oli)
    parts = w.particles
    if tag:  # a unique id per particle in the padding word
        parts = parts.copy()
        parts["pad0"] = np.arange(N, dtype=np.uint32)
    ctx.upload_store(parts)
    ctx.configure_grid(w.outheights, w.refine, 0, w.mass_scale)
    ctx.clear_grid()
    counter = (0, 0, 0, 0)
    cfgs = []
    for k in range(n_steps):
        if sort_every and k % sort_every == 0:
            ctx.sort_morton(10, want_map=False)
        cfg = fp.StepConfig(dt_seconds=w.dt_seconds, alpha=k * w.dt_seconds / 21600.0,
                            x_scale=w.x_scale, y_scale=w.y_scale, key=w.key, counter=counter,
                            dry_mode=1, dry_vd=2e-3, grid_fused=grid)
        cfgs.append(cfg)
        counter = ctx.step(cfg)
    return ctx, cfgs, counter


def test_sampled_slots_match_oracle_bitwise_at_full_size(fp, orc, big):
    say("full size: 10M particles, 360x181x60 met, 3 steps, 4096 sampled slots")
    ctx, cfgs, counter = run(fp, big, 3, grid=False)
    got = ctx.download_particles()
    ctx.close()
    assert counter == (N * 2 * 3, 0, 0, 0)
    rng = np.random.default_rng(17)
    sample = np.sort(rng.choice(N, 4096, replace=False))
    want = big.particles.copy()
    keep = np.zeros(N, bool)
    keep[sample] = True
    want["flags"][~keep] = 0
    for cfg in cfgs:
        prm = dict(dt_seconds=cfg.dt_seconds, alpha=cfg.alpha, x_scale=cfg.x_scale,
                   y_scale=cfg.y_scale, key=cfg.key, counter=cfg.counter, dry_mode=1,
                   dry_vd=2e-3)
        orc.step(want, met_dict(big.met), pbl_dict(big.pbl), prm)
    assert_particles_equal(got[sample], want[sample], "sampled full-size slots")


def test_full_size_invariants_and_run_to_run_determinism(fp, big):
    say("full size: 10M particles, 6 steps with Morton sort every 3 and fused grid, twice")
    ctx, _, _ = run(fp, big, 6, sort_every=3, tag=True)
    a = ctx.download_particles()
    ca, ma = ctx.download_grid()
    ctx.close()
    nx, ny, nz = big.met.shape
    act = a["flags"] == 1
    assert act.all()
    assert np.all((a["cell_x"] >= 0) & (a["cell_x"] < nx) & (a["cell_y"] >= 0) & (a["cell_y"] < ny))
    assert np.all((a["pos_z"] >= 0.01) & np.isfinite(a["pos_z"]))
    # each of the 6 fused gridding passes counts every active particle once
    assert int(ca.astype(np.int64).sum()) == 6 * N
    # dry deposition only removes mass; the multiset of particle ids is preserved by the sort
    assert np.all(a["mass"][:, 0] <= big.particles["mass"][:, 0].max())
    assert np.array_equal(np.sort(a["pad0"]), np.arange(N, dtype=np.uint32))
    ctx, _, _ = run(fp, big, 6, sort_every=3, tag=True)
    b = ctx.download_particles()
    cb, mb = ctx.download_grid()
    ctx.close()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(ca, cb) and np.array_equal(ma, mb)
