"""The benchmark configurations end to end on the GPU, at their stated sizes.

* Point release (100k particles, 64 x 64 x 32 met, 900 s steps, 24 h): the
  particles enter through fp_release_inject over the first hour (the linear
  window of release/mod.rs:181-196), every step grids the active particles and
  adds the grid to the output interval, and every hour fp_output_finalize
  gives the hourly means and concentrations.  All 24 hourly fields, the final
  records and the Philox counter equal the oracle's sequence
  (orc_release_inject, orc_step, orc_grid_accumulate, orc_output_finalize)
  bit for bit.
* The 96-member ensemble (about 32M particles, per-member 360 x 181 x 60 met
  perturbation, staggered releases over 6 h) through the captured graph:
  every member's active count, grid total (active particles x gridded steps),
  finiteness and domain, and 1024 sampled slots of three members bit for bit
  against the oracle with the member's own met, key, counters and temporal
  weights.
* The differentiable 2M-particle configuration (1024 sources, 64 wind
  coefficients, 2000 receptors, 72 steps): 2048 sampled slots of the forward
  state bit for bit against the oracle, the emission gradient of four
  sources against the exact linear decomposition, and gradients independent
  of the checkpoint interval.
* The stress configuration's 0.25 deg met (1440 x 721 x 60) with deposition,
  precipitation patches, the loss rule and compaction after every step, at a
  reduced particle count: active counts, compacted records and receptor sums
  equal the oracle's step + compaction sequence.

Sampled-slot checks rely on the step being independent per slot: the oracle
steps the same state with every other slot deactivated, so the sampled slots
get the same counters and inputs.
"""
import os
import sys

import numpy as np
import pytest

from conftest import assert_particles_equal, met_dict, pbl_dict

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def threads():
    return max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))))


# ------------------------------------------------------------------ point release, 24 h

def test_point_release_24h_hourly_means_match_the_oracle(fp, orc, workloads):
    from test_gpu_parity import oracle_params
    w = workloads.point_release()
    n = len(w.particles)
    nx, ny, nz = w.met.shape
    steps, sph = 96, 4
    grid = dict(xlon0=-32.0, ylat0=13.0, dx=1.0, dy=1.0, nx=nx, ny=ny)
    lon, lat = grid["xlon0"] + nx / 2.0, grid["ylat0"] + ny / 2.0
    rel = [dict(lon_min=lon, lon_max=lon, lat_min=lat, lat_max=lat, z_min=100.0, z_max=100.0, mass_kg=1.0,
                particle_count=n, start_s=0, end_s=3600, release_point=0, class_id=0)]
    geo = fp.OutputGeometry(xlon0=grid["xlon0"], ylat0=grid["ylat0"], dx=1.0, dy=1.0)
    (u0, v0, w0), (u1, v1, w1) = w.met.levels
    T = int(w.met.seconds_between_levels)
    say(f"point release: {n} particles released over hour 1, {steps} steps of {w.dt_seconds:.0f} s, "
        f"hourly means on {w.out_shape}")

    def cfg_at(k, counter):
        return fp.StepConfig(dt_seconds=w.dt_seconds, alpha=fp.interpolation_alpha(0, T, int(k * w.dt_seconds)),
                             x_scale=w.x_scale, y_scale=w.y_scale, key=w.key, counter=counter, dry_mode=1,
                             dry_vd=w.dry_vd, grid_fused=True)

    got_hours = []
    with fp.GpuTimestep(n, w.met.shape, n_species=1, out_shape=w.out_shape) as ctx:
        ctx.upload_met_bracket(u0, v0, w0, u1, v1, w1, w.met.level_heights)
        ctx.upload_pbl(w.pbl.ustar, w.pbl.wstar, w.pbl.hmix, w.pbl.oli)
        ctx.upload_store(w.particles)  # all slots inactive until released
        ctx.configure_grid(w.outheights, w.refine, 0, w.mass_scale)
        ctx.output_reset()
        done = np.zeros(1, np.uint64)
        cursor, counter = 0, (0, 0, 0, 0)
        for k in range(steps):
            cursor += ctx.inject_releases(grid, rel, done, int(k * w.dt_seconds), first_slot=cursor)
            ctx.clear_grid()
            counter = ctx.step(cfg_at(k, counter))
            ctx.output_sample()
            if (k + 1) % sph == 0:
                got_hours.append(ctx.output_finalize(geo, concentration=True))
                ctx.output_reset()
        got = ctx.download_particles()
    assert cursor == n and int(done[0]) == n

    want = w.particles.copy()
    vol = orc.output_volumes(grid["ylat0"], 1.0, 1.0, w.outheights, ny)
    met = dict(shape=w.met.shape, u0=u0, v0=v0, w0=w0, u1=u1, v1=v1, w1=w1, level_heights=w.met.level_heights)
    ncell = int(np.prod(w.out_shape))
    acc_c, acc_m = np.zeros(ncell, np.uint64), np.zeros(ncell, np.uint64)
    odone, ocursor, ocounter = np.zeros(1, np.uint64), 0, (0, 0, 0, 0)
    for k in range(steps):
        t = int(k * w.dt_seconds)
        new, st = orc.release_inject(grid, rel, odone, t, t)
        assert st == 0
        want[ocursor:ocursor + len(new)] = new
        ocursor += len(new)
        p = oracle_params(w, cfg_at(k, ocounter))
        p["n_threads"] = threads()
        ocounter = tuple(int(v) for v in orc.step(want, met, pbl_dict(w.pbl), p))
        oc, om = orc.grid_accumulate(want, w.out_shape, w.outheights, w.refine, 0, w.mass_scale)
        acc_c += oc.reshape(-1).astype(np.uint64)
        acc_m += om.reshape(-1).astype(np.uint64)
        if (k + 1) % sph == 0:
            ref = orc.output_finalize(acc_c, acc_m, sph, w.mass_scale, w.out_shape, vol)
            g = got_hours[k // sph]
            h = k // sph
            assert g["n_samples"] == sph
            assert np.array_equal(g["count"], ref["count"]), f"hour {h}: counts"
            assert np.array_equal(g["mass_kg"].view(np.uint32), ref["mass_kg"].view(np.uint32)), f"hour {h}: mass"
            assert np.array_equal(g["conc_kg_m3"].view(np.uint32), ref["conc_kg_m3"].view(np.uint32)), \
                f"hour {h}: concentration"
            acc_c[:] = 0
            acc_m[:] = 0
    assert counter == ocounter
    assert_particles_equal(got, want, "point release after 24 h")
    # every particle is counted once per step of its hour once released
    assert int(got_hours[-1]["count"].astype(np.int64).sum()) == sph * n
    assert int(got_hours[0]["count"].astype(np.int64).sum()) == n // 4 * (1 + 2 + 3)


# ------------------------------------------------------------------ 96-member ensemble

def test_ensemble_96_members_full_size_through_the_graph(fp, orc, workloads):
    import torch
    from test_gpu_parity import oracle_params

    W = workloads
    spec = W.ensemble_spec()
    M = len(spec.counts)
    caps = spec.counts.astype(np.uint64)
    met = W.global_met(2, n_levels=3)
    pbl = W.global_pbl(2)
    nx, ny, nz = met.shape
    out_shape = (nx, ny, 5)
    outheights = (100.0, 500.0, 1000.0, 3000.0, 10_000.0)
    bracket_s = met.seconds_between_levels
    dt = spec.dt_seconds
    window = int(round(bracket_s / dt))
    extra = 6
    say(f"ensemble: {M} members, {int(caps.sum())} particles, releases over {window} steps, "
        f"then {extra} steps in the next bracket, captured graph")
    dev = torch.device("cuda:0")
    levels = [[torch.from_numpy(np.ascontiguousarray(a)).to(dev).view(nx, ny, nz) for a in lvl]
              for lvl in met.levels]
    factors = [torch.from_numpy(W.ensemble_member_factor(spec, m, nx, ny)).to(dev)[:, :, None] for m in range(M)]
    keys = np.array(spec.keys, np.uint32)
    cfg = fp.StepConfig(dt_seconds=dt, x_scale=1.0 / W.METRES_PER_DEG, y_scale=1.0 / W.METRES_PER_DEG,
                        dry_mode=1, dry_vd=2e-3, grid_fused=True)
    order = np.argsort(caps)
    sample_members = [int(order[0]), int(order[M // 2]), int(order[-1])]
    history = {m: [] for m in sample_members}  # (bracket, alpha, counter) of each step the member took

    ens = fp.GpuEnsemble(caps, met.shape, n_species=1, out_shape=out_shape)
    with ens:
        ens.upload_pbl(pbl.ustar, pbl.wstar, pbl.hmix, pbl.oli)
        ens.configure_grid(outheights, (1, 1), 0, 1.0e8)
        ens.clear_grid()

        def upload_bracket(b):
            for m in range(M):
                ts = [(levels[b + t][c] * factors[m]).contiguous() for t in (0, 1) for c in range(3)]
                torch.cuda.synchronize()
                ens.upload_met_member_dev(m, [t.data_ptr() for t in ts], met.level_heights)
                del ts

        def record(ctr, released, b, t0, j):
            alpha = np.float32(min(max((t0 + j * dt) / bracket_s, 0.0), 1.0))
            for m in sample_members:
                if released[m]:
                    history[m].append((b, float(alpha), tuple(int(v) for v in ctr[m])))

        upload_bracket(0)
        ctr = np.zeros((M, 4), np.uint32)
        released = np.zeros(M, bool)
        t0, j = 0.0, 0
        for k in range(window):
            newly = np.nonzero(spec.start_step == k)[0]
            for m in newly:
                ens.upload_member(int(m), W.ensemble_member_particles(spec, int(m)))
                released[m] = True
            if len(newly):
                ens.configure(cfg, keys=keys, counters=ctr, dispatch=np.where(released, caps, 0),
                              time_in_bracket_s=k * dt, bracket_s=bracket_s)
                t0, j = k * dt, 0
            record(ctr, released, 0, t0, j)
            ctr = ens.step(1, use_graph=True)
            j += 1
        assert released.all()
        upload_bracket(1)
        ens.configure(cfg, keys=keys, counters=ctr, time_in_bracket_s=0.0, bracket_s=bracket_s)
        ens.clear_grid()
        for j in range(extra):
            record(ctr, released, 1, 0.0, j)
            ctr = ens.step(1, use_graph=True)
        states = {}
        for m in range(M):
            p = ens.download_member(m)
            counts, _ = ens.download_grid_member(m)
            cap = int(caps[m])
            act = (p["flags"] & 1) == 1
            assert act.sum() == cap, f"member {m}: active {act.sum()} of {cap}"
            assert int(counts.astype(np.int64).sum()) == cap * extra, f"member {m}: grid total"
            assert np.all((p["cell_x"] >= 0) & (p["cell_x"] < nx) & (p["cell_y"] >= 0) & (p["cell_y"] < ny))
            assert np.all(np.isfinite(p["pos_z"]) & (p["pos_z"] >= 0.01))
            assert np.all(np.isfinite(p["turb_u"]) & np.isfinite(p["turb_v"]) & np.isfinite(p["turb_w"]))
            assert np.all(p["mass"][:, 0] > 0)
            if m in sample_members:
                states[m] = p
    say(f"  all {M} members: active counts, grid totals, domain; oracle on members {sample_members}")
    rng = np.random.default_rng(23)
    for m in sample_members:
        cap = int(caps[m])
        sample = np.sort(rng.choice(cap, min(1024, cap), replace=False))
        want = W.ensemble_member_particles(spec, m)
        keep = np.zeros(cap, bool)
        keep[sample] = True
        want["flags"][~keep] = 0
        f = W.ensemble_member_factor(spec, m, nx, ny)[:, :, None]
        mm = {}
        for b, alpha, counter in history[m]:
            if b not in mm:
                lv = [[(np.asarray(a).reshape(nx, ny, nz) * f).astype(np.float32).reshape(-1) for a in met.levels[b + t]]
                      for t in (0, 1)]
                mm[b] = dict(shape=met.shape, u0=lv[0][0], v0=lv[0][1], w0=lv[0][2], u1=lv[1][0], v1=lv[1][1],
                             w1=lv[1][2], level_heights=met.level_heights)
            p = oracle_params(None, fp.StepConfig(**{**cfg.__dict__, "alpha": alpha, "key": tuple(int(x) for x in keys[m]),
                                                      "counter": counter}))
            p["n_threads"] = threads()
            orc.step(want, mm[b], pbl_dict(pbl), p)
        assert len(history[m]) == window + extra - int(spec.start_step[m])
        assert_particles_equal(states[m][sample], want[sample], f"ensemble member {m}")


# ------------------------------------------------------------------ 2M-particle adjoint

@pytest.fixture(scope="module")
def adjoint_2m(fp, workloads):
    c = workloads.adjoint_case()
    nx, ny, _ = c.met.shape
    return c


def make_adj_2m(fp, c, ckpt, n_steps=72, obs_nz=11, obs_dz=500.0):
    adj = fp.GpuAdjoint(c.counts, c.met.shape, n_met_levels=len(c.met.levels), n_modes=len(c.coeff),
                        obs_nz=obs_nz, obs_dz_m=obs_dz, n_steps=n_steps, obs_first_step=n_steps - 11,
                        checkpoint_every=ckpt)
    adj.upload_met(c.met.levels, c.met.level_heights, c.met.seconds_between_levels)
    adj.upload_basis(c.basis_u, c.basis_v)
    adj.upload_pbl(c.pbl.ustar, c.pbl.wstar, c.pbl.hmix, c.pbl.oli)
    return adj


def test_adjoint_2m_forward_sampled_slots_match_the_oracle(fp, orc, adjoint_2m):
    from test_gpu_adjoint import initial_particles, perturbed_levels
    from test_gpu_parity import oracle_params
    c = adjoint_2m
    n_steps = 72
    say(f"adjoint 2M: {int(c.counts.sum())} particles, {len(c.counts)} sources, {len(c.coeff)} coefficients, "
        f"{n_steps} steps; 2048 sampled slots vs oracle")
    cfg = fp.StepConfig(dt_seconds=c.dt_seconds, x_scale=c.x_scale, y_scale=c.y_scale, key=c.key,
                        dry_mode=1, dry_vd=2e-3)
    with make_adj_2m(fp, c, 8) as adj:
        nx, ny, _ = c.met.shape
        adj.set_receptors(np.array([[nx // 2, ny // 2, 0]]))
        adj.forward(cfg, c.emission, c.position, c.coeff)
        got = adj.download_state()
    parts = initial_particles(fp, c)
    n = len(parts)
    sample = np.sort(np.random.default_rng(29).choice(n, 2048, replace=False))
    keep = np.zeros(n, bool)
    keep[sample] = True
    parts["flags"][~keep] = 0
    lv = perturbed_levels(c, c.coeff)
    T = int(c.met.seconds_between_levels)
    counter = tuple(cfg.counter)
    for k in range(n_steps):
        t = int(k * c.dt_seconds)
        b = min(t // T, len(lv) - 2)
        met = dict(shape=c.met.shape, u0=lv[b][0], v0=lv[b][1], w0=lv[b][2], u1=lv[b + 1][0],
                   v1=lv[b + 1][1], w1=lv[b + 1][2], level_heights=c.met.level_heights)
        cfg.alpha = fp.interpolation_alpha(b * T, (b + 1) * T, t)
        cfg.counter = counter
        p = oracle_params(c, cfg)
        p["n_threads"] = threads()
        counter = tuple(int(v) for v in orc.step(parts, met, pbl_dict(c.pbl), p))
    assert_particles_equal(got[sample], parts[sample], "adjoint 2M forward, sampled slots")


@pytest.fixture(scope="module")
def receptors_2m(fp, adjoint_2m):
    """2000 lattice nodes reached by a perturbed-truth run, observed there."""
    c = adjoint_2m
    nx, ny, _ = c.met.shape
    cfg = fp.StepConfig(dt_seconds=c.dt_seconds, x_scale=c.x_scale, y_scale=c.y_scale, key=c.key,
                        dry_mode=1, dry_vd=2e-3)
    rng = np.random.default_rng(17)
    q_true = (c.emission * np.exp(0.3 * rng.normal(size=len(c.emission)))).astype(np.float32)
    r_true = c.position.copy()
    r_true[:, :2] += rng.normal(0.0, 0.3, (len(c.emission), 2)).astype(np.float32)
    c_true = rng.normal(0.0, 0.5, len(c.coeff)).astype(np.float32)
    i, j, l = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(11), indexing="ij")
    nodes = np.stack([i.ravel(), j.ravel(), l.ravel()], axis=1)
    with make_adj_2m(fp, c, 8) as adj:
        adj.set_receptors(nodes)
        truth, _ = adj.forward(cfg, q_true, r_true, c_true)
    hit = np.nonzero(truth > 0)[0]
    pick = np.sort(rng.choice(hit, size=min(2000, len(hit)), replace=False))
    return cfg, nodes[pick], truth[pick]


def test_adjoint_2m_emission_gradient_is_the_linear_decomposition(fp, adjoint_2m, receptors_2m):
    c = adjoint_2m
    cfg, nodes, obs = receptors_2m
    say(f"adjoint 2M: emission gradient of 4 sources vs unit-source forward runs ({len(nodes)} receptors)")
    with make_adj_2m(fp, c, 8) as adj:
        adj.set_receptors(nodes, obs)
        vals, _ = adj.forward(cfg, c.emission, c.position, c.coeff)
        gq, _, _ = adj.backward()
        picks = list(np.argsort(-np.abs(gq))[:2]) + list(np.argsort(np.abs(gq))[len(gq) // 2:len(gq) // 2 + 2])
        for s in picks:
            unit = np.zeros(len(c.emission), np.float32)
            unit[s] = 1.0
            col, _ = adj.forward(cfg, unit, c.position, c.coeff)
            want = float(np.sum(2.0 * (vals - obs) * col))
            assert gq[s] == pytest.approx(want, rel=2e-5, abs=1e-6 * float(np.max(np.abs(gq)))), int(s)


def test_adjoint_2m_gradients_do_not_depend_on_the_checkpoint_interval(fp, adjoint_2m, receptors_2m):
    c = adjoint_2m
    cfg, nodes, obs = receptors_2m
    say("adjoint 2M: gradients with checkpoints every 8, 3 and 72 steps")
    grads = []
    for ckpt in (8, 3, 72):
        with make_adj_2m(fp, c, ckpt) as adj:
            adj.set_receptors(nodes, obs)
            adj.forward(cfg, c.emission, c.position, c.coeff)
            grads.append(adj.backward())
    for g in grads[1:]:
        for a, b in zip(g, grads[0]):
            np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12 * max(1.0, float(np.max(np.abs(b)))))


# ------------------------------------------------------------------ 0.25 deg stress met

def test_stress_run_on_the_quarter_degree_met_matches_the_oracle_sequence(fp, orc, workloads):
    from test_gpu_adjoint import cic_values_lattice, initial_particles
    from test_gpu_parity import oracle_params
    c = workloads.stress_case(200_000, 512, deg=0.25)
    n_steps, obs_first = 36, 31
    m0 = float(np.median(c.emission / c.counts.astype(np.float32)))
    cfg = fp.StepConfig(dt_seconds=c.dt_seconds, x_scale=c.x_scale, y_scale=c.y_scale, key=c.key,
                        dry_mode=1, dry_vd=1e-2, wet_mode=3, wet_a=1e-4, wet_b=0.8, wet_fraction=1.0,
                        loss_mode=1, loss_min_mass=0.9 * m0)
    nx, ny, _ = c.met.shape
    src = np.floor(c.position[:300, :2]).astype(np.int64)
    nodes = np.unique(np.array([(i + di, j + dj, l) for i, j in src for di in (0, 1) for dj in (0, 1)
                                for l in (0, 1, 2) if i + di < nx and j + dj < ny]), axis=0)
    say(f"stress on {c.met.shape}: {int(c.counts.sum())} particles, {n_steps} steps, compaction every step")
    with fp.GpuAdjoint(c.counts, c.met.shape, n_met_levels=2, n_modes=0, obs_nz=11, obs_dz_m=500.0,
                       n_steps=n_steps, obs_first_step=obs_first, checkpoint_every=8, compact=True,
                       emission_only=True) as adj:
        adj.upload_met(c.met.levels, c.met.level_heights, c.met.seconds_between_levels)
        adj.upload_pbl(c.pbl.ustar, c.pbl.wstar, c.pbl.hmix, c.pbl.oli, c.pbl.precip)
        adj.set_receptors(nodes)
        vals, _ = adj.forward(cfg, c.emission, c.position)
        counts = adj.active_counts()
        got = adj.download_state(int(counts[-1]))
    parts = initial_particles(fp, c)
    (u0, v0, w0), (u1, v1, w1) = c.met.levels
    met = dict(shape=c.met.shape, u0=u0, v0=v0, w0=w0, u1=u1, v1=v1, w1=w1, level_heights=c.met.level_heights)
    T = int(c.met.seconds_between_levels)
    table = np.full(nx * ny * 11, -1, np.int64)
    table[(nodes[:, 0] * ny + nodes[:, 1]) * 11 + nodes[:, 2]] = np.arange(len(nodes))
    want_vals = np.zeros(len(nodes))
    counter = tuple(cfg.counter)
    n_cur = len(parts)
    want_counts = [n_cur]
    for k in range(n_steps):
        cfg.alpha = fp.interpolation_alpha(0, T, int(k * c.dt_seconds))
        cfg.counter = counter
        live = parts[:n_cur].copy()
        p = oracle_params(c, cfg)
        p["n_threads"] = threads()
        counter = tuple(int(v) for v in orc.step(live, met, pbl_dict(c.pbl), p))
        if k + 1 >= obs_first:
            cic_values_lattice(c, live, table, want_vals, 11, 500.0)
        n_cur, out, _ = orc.compact(live)
        parts[: len(out)] = out
        want_counts.append(n_cur)
    say(f"  active {want_counts[0]} -> {want_counts[-1]}")
    assert list(counts) == want_counts
    assert want_counts[-1] < 0.9 * want_counts[0], "the run must lose particles"
    assert got.tobytes() == parts[: want_counts[-1]].tobytes()
    assert np.count_nonzero(want_vals) > 0
    np.testing.assert_allclose(vals, want_vals, rtol=1e-9, atol=1e-12 * max(1.0, float(np.max(want_vals))))
