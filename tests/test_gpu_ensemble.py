"""Batched ensembles (fp_ens_*) vs the oracle, through the C ABI.

Every member of an ensemble must come out bit-identical to the oracle's step
of that member alone (its own particles, met bracket, Philox key and counter,
and the temporal weight of the shared clock), whatever the launch shape, for
direct launches and for replays of the captured graph, with ragged member
sizes (a single particle, less than a wave, non-multiples of the block) and a
dispatch count below a member's capacity.
"""
import sys

import numpy as np
import pytest

from conftest import assert_particles_equal, pbl_dict

pytestmark = pytest.mark.gpu

CAPS = [3000, 257, 64, 5000, 1, 1200]
DISPATCH = [3000, 257, 64, 4000, 1, 1200]
T0, BRACKET = 1800.0, 3 * 3600.0


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def small(workloads):
    return workloads.small_random(20_000, shape=(48, 40, 24))


def member_levels(w, m):
    """Member m's met bracket: the base winds scaled by 1 + 0.05 m."""
    f = np.float32(1.0 + 0.05 * m)
    return [tuple((a * f).astype(np.float32) for a in w.met.levels[b]) for b in (0, 1)]


def member_particles(w, m):
    off = 1500 * m
    return w.particles[off:off + CAPS[m]].copy()


def member_key(w, m):
    return ((w.key[0] ^ (0x9E3779B9 * m)) & 0xFFFFFFFF, (w.key[1] + m) & 0xFFFFFFFF)


def member_counter(m):
    if m == 3:
        return (0xFFFFFFF0, 0xFFFFFFFF, 3, 0)  # carries out of the low words on the first step
    return (1000 * m + 7, 0, 0, 0)


CASES = {
    "production": dict(dry_mode=1, dry_vd=2e-3, grid_fused=True),
    "backward_wet_two_substeps": dict(direction=-1, n_substeps=2, wet_mode=3, wet_a=1e-4, wet_b=0.8,
                                      wet_fraction=1.0),
    "no_substeps_loss": dict(n_substeps=0, dry_mode=1, dry_vd=5e-2, loss_mode=1, loss_min_mass=0.5),
}


def build_ensemble(fp, w, device_met=False):
    """Ensemble of the CAPS members; device_met uploads the met brackets from
    torch device tensors (fp_ens_met_upload_member_dev) instead of host arrays."""
    ens = fp.GpuEnsemble(CAPS, w.met.shape, n_species=4, out_shape=w.out_shape)
    for m in range(len(CAPS)):
        lv = member_levels(w, m)
        if device_met:
            import torch
            ts = [torch.from_numpy(a).cuda() for b in lv for a in b]
            torch.cuda.synchronize()
            ens.upload_met_member_dev(m, [t.data_ptr() for t in ts], w.met.level_heights)
            del ts
        else:
            ens.upload_met_member(m, *lv[0], *lv[1], w.met.level_heights)
        ens.upload_member(m, member_particles(w, m))
    ens.upload_pbl(w.pbl.ustar, w.pbl.wstar, w.pbl.hmix, w.pbl.oli, w.pbl.precip)
    ens.configure_grid(w.outheights, w.refine, 0, w.mass_scale)
    ens.clear_grid()
    return ens


def oracle_member(fp, orc, w, m, cfg, n_steps):
    """Oracle states after each step, counters after each step and the
    member's fused grid sum over the steps (u32 wrapping, as on the device)."""
    from test_gpu_parity import oracle_params
    parts = member_particles(w, m)
    lv = member_levels(w, m)
    met = dict(shape=w.met.shape, u0=lv[0][0], v0=lv[0][1], w0=lv[0][2], u1=lv[1][0], v1=lv[1][1],
               w1=lv[1][2], level_heights=w.met.level_heights)
    counter = member_counter(m)
    dt = -cfg.dt_seconds if cfg.direction < 0 else cfg.dt_seconds
    states, counters = [], []
    ncell = int(np.prod(w.out_shape))
    gc = np.zeros(ncell, np.uint32)
    gm = np.zeros(ncell, np.uint32)
    for k in range(n_steps):
        # the clock's weight is interpolation_alpha at the step's start time
        alpha = fp.interpolation_alpha(0, int(BRACKET), int(T0 + k * dt))
        cfg.alpha, cfg.counter, cfg.key = alpha, counter, member_key(w, m)
        nxt = orc.step(parts, met, pbl_dict(w.pbl), oracle_params(w, cfg), n_dispatch=DISPATCH[m])
        counter = tuple(int(v) for v in nxt)
        states.append(parts.copy())
        counters.append(counter)
        if cfg.grid_fused:
            oc, om = orc.grid_accumulate(parts[:DISPATCH[m]], w.out_shape, w.outheights, w.refine, 0,
                                         w.mass_scale)
            gc += oc.astype(np.uint32)
            gm += om.astype(np.uint32)
    return states, counters, (gc, gm)


@pytest.mark.parametrize("launch", [(256, 1), (64, 4), (1024, 2)])
@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("case", sorted(CASES))
def test_ensemble_members_match_oracle_bitwise(fp, orc, small, case, use_graph, launch):
    from test_gpu_parity import base_cfg
    say(f"ensemble parity: case={case} graph={use_graph} launch={launch} members={len(CAPS)}")
    n_steps = 3
    want = [oracle_member(fp, orc, small, m, base_cfg(fp, small, **CASES[case]), n_steps)
            for m in range(len(CAPS))]
    # the device-pointer met path on one case, the host path on the others
    device_met = case == "production" and launch == (256, 1)
    with build_ensemble(fp, small, device_met) as ens:
        ens.set_launch(*launch)
        cfg = base_cfg(fp, small, **CASES[case])
        keys = [member_key(small, m) for m in range(len(CAPS))]
        ctrs = [member_counter(m) for m in range(len(CAPS))]
        ens.configure(cfg, keys=keys, counters=ctrs, dispatch=DISPATCH, time_in_bracket_s=T0,
                      bracket_s=BRACKET)
        for k in range(n_steps):
            got_ctr = ens.step(1, use_graph=use_graph)
            for m in range(len(CAPS)):
                assert tuple(int(v) for v in got_ctr[m]) == want[m][1][k], f"member {m} counter"
                assert_particles_equal(ens.download_member(m), want[m][0][k],
                                       f"{case} member {m} step {k}")
        if CASES[case].get("grid_fused"):
            for m in range(len(CAPS)):
                counts, mass = ens.download_grid_member(m)
                assert np.array_equal(counts, want[m][2][0]), f"member {m} grid counts"
                assert np.array_equal(mass, want[m][2][1]), f"member {m} grid mass"


def test_graph_replays_many_steps_equal_direct_launches(fp, small):
    from test_gpu_parity import base_cfg
    say("ensemble: 12 graph replays vs 12 direct launches")
    outs = []
    for use_graph in (False, True):
        with build_ensemble(fp, small) as ens:
            cfg = base_cfg(fp, small, **CASES["production"])
            ens.configure(cfg, keys=[member_key(small, m) for m in range(len(CAPS))],
                          counters=[member_counter(m) for m in range(len(CAPS))], dispatch=DISPATCH,
                          time_in_bracket_s=T0, bracket_s=BRACKET)
            ens.step(5, use_graph=use_graph)
            ctr = ens.step(7, use_graph=use_graph)
            outs.append(([ens.download_member(m) for m in range(len(CAPS))], ctr,
                         [ens.download_grid_member(m) for m in range(len(CAPS))]))
    (pa, ca, ga), (pb, cb, gb) = outs
    assert np.array_equal(ca, cb)
    for m in range(len(CAPS)):
        assert_particles_equal(pa[m], pb[m], f"member {m} graph vs direct")
        assert np.array_equal(ga[m][0], gb[m][0]) and np.array_equal(ga[m][1], gb[m][1])


def test_multi_step_graphs_equal_direct_launches(fp, small):
    """Multi-step graphs (a 36-step group plus one-step remainders, a 6-step
    output interval replayed twice) kept across re-configurations with the
    same launch table and captured again after a launch-shape or dispatch
    change, against the same calls with direct launches."""
    from test_gpu_parity import base_cfg
    say("ensemble: multi-step graphs vs direct launches over re-configurations")
    keys = [member_key(small, m) for m in range(len(CAPS))]
    outs = []
    for use_graph in (False, True):
        with build_ensemble(fp, small) as ens:
            cfg = base_cfg(fp, small, **CASES["production"])
            ens.configure(cfg, keys=keys, counters=[member_counter(m) for m in range(len(CAPS))],
                          dispatch=DISPATCH, time_in_bracket_s=T0, bracket_s=BRACKET)
            ctr = ens.step(40, use_graph=use_graph)          # 36-step graph + 4 one-step replays
            ens.configure(cfg, keys=keys, counters=ctr, dispatch=DISPATCH, time_in_bracket_s=T0,
                          bracket_s=BRACKET)                  # same table: graphs kept
            ens.step(6, use_graph=use_graph)
            ctr = ens.step(6, use_graph=use_graph)
            ens.set_launch(128, 2)                            # new shape: captured again
            ctr = ens.step(3, use_graph=use_graph)
            ens.configure(cfg, keys=keys, counters=ctr, dispatch=CAPS, time_in_bracket_s=T0 + 600.0,
                          bracket_s=BRACKET)                  # new dispatch: captured again
            ctr = ens.step(3, use_graph=use_graph)
            outs.append(([ens.download_member(m) for m in range(len(CAPS))], ctr,
                         [ens.download_grid_member(m) for m in range(len(CAPS))]))
    (pa, ca, ga), (pb, cb, gb) = outs
    assert np.array_equal(ca, cb)
    for m in range(len(CAPS)):
        assert_particles_equal(pa[m], pb[m], f"member {m} multi-step graphs vs direct")
        assert np.array_equal(ga[m][0], gb[m][0]) and np.array_equal(ga[m][1], gb[m][1])


def test_member_equals_single_run_context(fp, small):
    """A member advanced in an ensemble equals the same member stepped alone
    by fp_step with the clock's alpha."""
    from test_gpu_parity import base_cfg
    say("ensemble member vs fp_step")
    m = 3
    with build_ensemble(fp, small) as ens:
        cfg = base_cfg(fp, small, **CASES["production"])
        ens.configure(cfg, keys=[member_key(small, j) for j in range(len(CAPS))],
                      counters=[member_counter(j) for j in range(len(CAPS))], dispatch=DISPATCH,
                      time_in_bracket_s=T0, bracket_s=BRACKET)
        ens.step(4, use_graph=True)
        got = ens.download_member(m)
    lv = member_levels(small, m)
    with fp.GpuTimestep(CAPS[m], small.met.shape, n_species=4) as ctx:
        ctx.upload_met_bracket(*lv[0], *lv[1], small.met.level_heights)
        ctx.upload_pbl(small.pbl.ustar, small.pbl.wstar, small.pbl.hmix, small.pbl.oli, small.pbl.precip)
        ctx.upload_store(member_particles(small, m))
        ctx.set_dispatch_count(DISPATCH[m])
        counter = member_counter(m)
        for k in range(4):
            c = base_cfg(fp, small, **CASES["production"])
            c.grid_fused = False
            c.alpha = fp.interpolation_alpha(0, int(BRACKET), int(T0 + k * small.dt_seconds))
            c.key, c.counter = member_key(small, m), counter
            counter = ctx.step(c)
        alone = ctx.download_particles()
    assert_particles_equal(got, alone, "ensemble member vs fp_step")


def test_ensemble_errors_are_reported(fp, small):
    from test_gpu_parity import base_cfg
    with pytest.raises(fp.FpError) as e:
        fp.GpuEnsemble([], small.met.shape)
    assert e.value.code == fp.FP_E_INVALID
    with fp.GpuEnsemble([10, 20], small.met.shape, out_shape=small.out_shape) as ens:
        with pytest.raises(fp.FpError) as e:
            ens.configure(base_cfg(fp, small))
        assert e.value.code == fp.FP_E_STATE
        with pytest.raises(fp.FpError) as e:
            ens.step(1)
        assert e.value.code == fp.FP_E_STATE
        with pytest.raises(fp.FpError) as e:
            ens.upload_member(2, small.particles[:5])
        assert e.value.code == fp.FP_E_INVALID
        with pytest.raises(fp.FpError) as e:
            ens.upload_member(0, small.particles[:11])
        assert e.value.code == fp.FP_E_INVALID
        lv = member_levels(small, 0)
        ens.upload_met_member(0, *lv[0], *lv[1], small.met.level_heights)
        with pytest.raises(fp.FpError):
            ens.upload_met_member(1, *lv[0], *lv[1], small.met.level_heights * 1.5)
        ens.upload_met_member(1, *lv[0], *lv[1], small.met.level_heights)
        ens.upload_pbl(small.pbl.ustar, small.pbl.wstar, small.pbl.hmix, small.pbl.oli)
        with pytest.raises(fp.FpError) as e:
            ens.configure(base_cfg(fp, small, wet_mode=3, wet_a=1e-4, wet_b=0.8, wet_fraction=1.0))
        assert e.value.code == fp.FP_E_STATE
        for bad in (dict(dry_mode=2), dict(wet_mode=2), dict(n_substeps=5), dict(direction=0)):
            with pytest.raises(fp.FpError) as e:
                ens.configure(base_cfg(fp, small, **bad))
            assert e.value.code == fp.FP_E_INVALID
        with pytest.raises(fp.FpError) as e:
            ens.configure(base_cfg(fp, small), dispatch=[11, 20])
        assert e.value.code == fp.FP_E_INVALID
        ens.configure(base_cfg(fp, small))
        ens.step(2)
        with pytest.raises(fp.FpError):
            ens.set_launch(100, 1)


def configured(fp, small, ens, launch=None):
    from test_gpu_parity import base_cfg
    if launch:
        ens.set_launch(*launch)
    ens.configure(base_cfg(fp, small, **CASES["production"]), keys=[member_key(small, m) for m in range(len(CAPS))],
                  counters=[member_counter(m) for m in range(len(CAPS))], dispatch=DISPATCH,
                  time_in_bracket_s=T0, bracket_s=BRACKET)


def test_member_sort_equals_the_single_context_sort(fp, small):
    """fp_ens_sort_morton leaves every member in the order fp_sort_morton gives
    for that member alone (its dispatch range; the slots beyond it untouched),
    and the steps after it equal the single context's steps after its sort."""
    say("ensemble Morton sort vs fp_sort_morton per member")
    from test_gpu_parity import base_cfg
    with build_ensemble(fp, small) as ens:
        configured(fp, small, ens)
        ctr3 = ens.step(3, use_graph=True)
        before = [ens.download_member(m) for m in range(len(CAPS))]
        ens.sort_morton(bits_per_axis=10)
        sorted_ = [ens.download_member(m) for m in range(len(CAPS))]
        ens.step(2, use_graph=True)   # the graph captured before the sort stays valid
        after = [ens.download_member(m) for m in range(len(CAPS))]
    for m in range(len(CAPS)):
        lv = member_levels(small, m)
        with fp.GpuTimestep(CAPS[m], small.met.shape, n_species=4) as ctx:
            ctx.upload_met_bracket(*lv[0], *lv[1], small.met.level_heights)
            ctx.upload_pbl(small.pbl.ustar, small.pbl.wstar, small.pbl.hmix, small.pbl.oli, small.pbl.precip)
            ctx.upload_store(before[m])
            ctx.set_dispatch_count(DISPATCH[m])
            ctx.sort_morton(10, want_map=False)
            assert_particles_equal(sorted_[m], ctx.download_particles(n=CAPS[m]), f"member {m} sorted")
            c = base_cfg(fp, small, **CASES["production"])
            c.grid_fused = False
            c.key = member_key(small, m)
            ctr = tuple(int(v) for v in ctr3[m])
            for k in range(3, 5):
                c.alpha = fp.interpolation_alpha(0, int(BRACKET), int(T0 + k * small.dt_seconds))
                c.counter = ctr
                ctr = ctx.step(c)
            assert_particles_equal(after[m], ctx.download_particles(n=CAPS[m]), f"member {m} stepped after sort")


def test_autotune_keeps_the_run_unchanged(fp, small):
    """fp_ens_autotune times launch shapes and re-sort intervals on copies of
    the state: afterwards the ensemble continues exactly like one that was
    not tuned (results do not depend on the launch shape)."""
    say("ensemble autotune: state, grids and clock restored")
    outs = []
    for tune in (False, True):
        with build_ensemble(fp, small) as ens:
            configured(fp, small, ens, launch=(256, 1))
            ens.step(2, use_graph=True)
            if tune:
                r = ens.autotune(sort_intervals=(0, 2), warmup_steps=1, measured_steps=2, sort_trial_steps=3)
                say(f"  autotune: {r}")
                assert r["block_size"] in (128, 256, 512) and r["particles_per_thread"] in (1, 2, 4)
                assert r["sort_interval"] in (0, 2) and r["step_ms"] > 0.0
            ctr = ens.step(3, use_graph=True)
            outs.append(([ens.download_member(m) for m in range(len(CAPS))], ctr,
                         [ens.download_grid_member(m) for m in range(len(CAPS))]))
    (pa, ca, ga), (pb, cb, gb) = outs
    assert np.array_equal(ca, cb)
    for m in range(len(CAPS)):
        assert_particles_equal(pa[m], pb[m], f"member {m} tuned vs untuned")
        assert np.array_equal(ga[m][0], gb[m][0]) and np.array_equal(ga[m][1], gb[m][1])
    with build_ensemble(fp, small) as ens:
        with pytest.raises(fp.FpError) as e:
            ens.autotune()
        assert e.value.code == fp.FP_E_STATE
