"""HIP path vs the CPU oracle, through the C ABI.

Integer work (Philox blocks, counters, compaction survivor maps, Morton
permutations, fixed-point grids) must be bit-exact; so must the float state,
because kernel and oracle evaluate the same operation sequence with the same
deterministic transcendentals (fp_math.h) and no contraction.
"""
import sys

import numpy as np
import pytest

from conftest import assert_particles_equal, met_dict, pbl_dict

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def small(workloads):
    return workloads.small_random(20_000, shape=(48, 40, 24))


def make_ctx(fp, w, n_species=4, out=True):
    ctx = fp.GpuTimestep(len(w.particles), w.met.shape, n_species=n_species,
                         out_shape=w.out_shape if out else None)
    u0, v0, w0 = w.met.levels[0]
    u1, v1, w1 = w.met.levels[1]
    ctx.upload_met_bracket(u0, v0, w0, u1, v1, w1, w.met.level_heights)
    ctx.upload_pbl(w.pbl.ustar, w.pbl.wstar, w.pbl.hmix, w.pbl.oli, w.pbl.precip)
    ctx.upload_store(w.particles)
    return ctx


def oracle_params(w, cfg):
    d = dict(dt_seconds=cfg.dt_seconds, alpha=cfg.alpha, x_scale=cfg.x_scale,
             y_scale=cfg.y_scale, z_scale=cfg.z_scale, direction=cfg.direction, key=cfg.key,
             counter=cfg.counter, n_substeps=cfg.n_substeps, min_height_m=cfg.min_height_m,
             rho_grad_over_rho=cfg.rho_grad_over_rho, dry_mode=cfg.dry_mode, dry_vd=cfg.dry_vd,
             dry_href=cfg.dry_href, wet_mode=cfg.wet_mode, wet_lambda=cfg.wet_lambda,
             wet_fraction=cfg.wet_fraction, wet_a=cfg.wet_a, wet_b=cfg.wet_b,
             loss_mode=cfg.loss_mode, loss_min_mass=cfg.loss_min_mass)
    return d


def base_cfg(fp, w, **kw):
    c = fp.StepConfig(dt_seconds=w.dt_seconds, alpha=0.35, x_scale=w.x_scale, y_scale=w.y_scale,
                      key=w.key, counter=(11, 0, 0, 0))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def test_philox_blocks_match_oracle_and_known_answers(fp, orc, small):
    say("philox blocks: 100000 blocks at a carrying counter")
    with fp.GpuTimestep(16, (4, 4, 4), n_species=1) as ctx:
        base, key = [0xFFFFFF00, 0xFFFFFFFF, 7, 0], [0xDECAFBAD, 0x12345678]
        got = ctx.philox_blocks(base, key, 100_000)
        assert np.array_equal(got, orc.philox_stream(base, key, 100_000))
        assert ctx.philox_blocks([0, 0, 0, 0], [0, 0], 1)[0].tolist() == [
            0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
        assert ctx.philox_blocks([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, 1)[0].tolist() == [
            0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]


def test_upload_download_roundtrip_is_bitwise(fp, small):
    say("roundtrip: 20000 particles")
    with make_ctx(fp, small) as ctx:
        assert_particles_equal(ctx.download_particles(), small.particles, "roundtrip")
        slots = np.array([5, 17, 19999, 0], np.uint32)
        new = small.particles[[1, 2, 3, 4]].copy()
        new["time"] = [101, 102, 103, 104]
        ctx.upload_particle_slots(slots, new)
        want = small.particles.copy()
        want[slots] = new
        assert_particles_equal(ctx.download_particles(), want, "slot upload")


STEP_CASES = {
    "production": {},
    "dry_wet_uniform": dict(dry_mode=1, dry_vd=2e-3, wet_mode=1, wet_lambda=1e-4, wet_fraction=0.7),
    "wet_precip_field": dict(wet_mode=3, wet_a=1e-4, wet_b=0.8, wet_fraction=1.0, dry_mode=1,
                             dry_vd=1e-2),
    "no_substeps": dict(n_substeps=0),
    "one_substep": dict(n_substeps=1),
    "two_substeps": dict(n_substeps=2),
    "three_substeps": dict(n_substeps=3),
    "backward": dict(direction=-1),
    "density_gradient": dict(rho_grad_over_rho=-1.2e-4),
    "alpha_clamped": dict(alpha=1.5),
    "loss_rule": dict(dry_mode=1, dry_vd=5e-2, loss_mode=1, loss_min_mass=0.5),
}


@pytest.mark.parametrize("case", sorted(STEP_CASES))
def test_step_matches_oracle_bitwise(fp, orc, small, case):
    say(f"step parity: case={case} n=20000 met=48x40x24 steps=3")
    kw = STEP_CASES[case]
    want = small.particles.copy()
    with make_ctx(fp, small) as ctx:
        counter = (11, 0, 0, 0)
        for k in range(3):
            cfg = base_cfg(fp, small, counter=counter, **kw)
            nxt = ctx.step(cfg)
            onxt = orc.step(want, met_dict(small.met), pbl_dict(small.pbl), oracle_params(small, cfg))
            assert list(nxt) == [int(v) for v in onxt]
            counter = nxt
        got = ctx.download_particles()
    assert_particles_equal(got, want, f"step {case}")


@pytest.mark.parametrize("min_height", [0.01, 2.0 ** -40, 0.0])
def test_step_at_the_quotient_range_check_matches_oracle(fp, orc, workloads, min_height):
    """Wavefronts whose cells and heights pass the range check evaluate the
    turbulence quotients without range scaling; the others (and every
    wavefront when z_min is below 2^-40) take the IEEE quotients.  PBL cells
    at, just inside and just outside each bound of the check (u*, w*, h at
    2^20, |L| at 2^-20) and beyond it, and particle heights at 0, 2^-41,
    2^-40 and ordinary values, mixed within wavefronts: bit for bit."""
    say(f"step parity at the quotient range check: z_min={min_height}")
    w = workloads.small_random(8192, shape=(16, 12, 24), seed=5)
    rng = np.random.default_rng(17)
    n = w.pbl.ustar.size
    edge = np.float32(2.0 ** 20)
    above = np.nextafter(edge, np.float32(np.inf))
    pick = lambda vals, p: rng.choice(np.array(vals, np.float32), n, p=p)
    ustar = np.where(rng.uniform(size=n) < 0.6, w.pbl.ustar, pick([edge, above, 1e25], [0.4, 0.4, 0.2]))
    wstar = np.where(rng.uniform(size=n) < 0.7, w.pbl.wstar, pick([edge, above, 3e30], [0.4, 0.4, 0.2]))
    hmix = np.where(rng.uniform(size=n) < 0.6, w.pbl.hmix, pick([edge, above, 1e30], [0.4, 0.4, 0.2]))
    oli = np.where(rng.uniform(size=n) < 0.6, w.pbl.oli,
                   pick([-edge, -above, -1e20, -edge * 0.5], [0.3, 0.3, 0.2, 0.2]))
    parts = w.particles.copy()
    z = parts["pos_z"]
    sel = rng.uniform(size=z.size)
    z[sel < 0.05] = 0.0
    z[(sel >= 0.05) & (sel < 0.1)] = np.float32(2.0 ** -41)
    z[(sel >= 0.1) & (sel < 0.15)] = np.float32(2.0 ** -40)
    z[(sel >= 0.15) & (sel < 0.2)] = np.float32(1e-30)
    met = met_dict(w.met)
    pbl = dict(shape=w.pbl.shape, ustar=ustar.astype(np.float32), wstar=wstar.astype(np.float32),
               hmix=hmix.astype(np.float32), oli=oli.astype(np.float32))
    want = parts.copy()
    with fp.GpuTimestep(len(parts), w.met.shape, n_species=4) as ctx:
        u0, v0, w0 = w.met.levels[0]
        u1, v1, w1 = w.met.levels[1]
        ctx.upload_met_bracket(u0, v0, w0, u1, v1, w1, w.met.level_heights)
        ctx.upload_pbl(pbl["ustar"], pbl["wstar"], pbl["hmix"], pbl["oli"])
        ctx.upload_store(parts)
        counter = (7, 0, 0, 0)
        for _ in range(3):
            cfg = base_cfg(fp, w, counter=counter, min_height_m=min_height)
            nxt = ctx.step(cfg)
            onxt = orc.step(want, met, pbl, oracle_params(w, cfg))
            assert list(nxt) == [int(v) for v in onxt]
            counter = nxt
        got = ctx.download_particles()
    assert_particles_equal(got, want, f"range check z_min={min_height}")


def test_step_in_a_set_beyond_32bit_offsets_matches_oracle(fp, orc, small):
    """A particle set of more than 2^32 bytes (capacity above 44,739,242
    slots) is addressed with 64-bit field bases, a smaller one with 32-bit
    offsets from the set's base: the production step (4 sub-steps, one
    particle per thread) in a context of each kind, bit for bit against the
    oracle."""
    w = small
    parts = w.particles[:4096].copy()
    met, pbl = met_dict(w.met), pbl_dict(w.pbl)
    for cap in (len(parts), 44_739_243):
        say(f"step parity: capacity {cap}")
        want = parts.copy()
        with fp.GpuTimestep(cap, w.met.shape, n_species=4) as ctx:
            u0, v0, w0 = w.met.levels[0]
            u1, v1, w1 = w.met.levels[1]
            ctx.upload_met_bracket(u0, v0, w0, u1, v1, w1, w.met.level_heights)
            ctx.upload_pbl(w.pbl.ustar, w.pbl.wstar, w.pbl.hmix, w.pbl.oli, w.pbl.precip)
            ctx.upload_store(parts)
            ctx.set_dispatch_count(len(parts))
            ctx.set_launch(256, 1)
            counter = (5, 0, 0, 0)
            for _ in range(2):
                cfg = base_cfg(fp, w, counter=counter, n_substeps=4, dry_mode=1, dry_vd=2e-3)
                nxt = ctx.step(cfg)
                onxt = orc.step(want, met, pbl, oracle_params(w, cfg))
                assert list(nxt) == [int(v) for v in onxt]
                counter = nxt
            got = ctx.download_particles(len(parts))
        assert_particles_equal(got, want, f"capacity {cap}")


@pytest.mark.parametrize("case", ["production", "three_substeps"])
@pytest.mark.parametrize("ppt", [1, 2, 4])
@pytest.mark.parametrize("block", [64, 128, 256, 512, 1024])
def test_every_launch_shape_matches_oracle_bitwise(fp, orc, small, block, ppt, case):
    """Each block size the library compiles (each its own register
    allocation under the 128-VGPR cap), at 1, 2 and 4 particles per thread,
    with the sub-step count fixed at compile time (production, 4) and read
    at run time (3): two steps with dry and wet deposition and fused
    gridding, bit for bit against the oracle."""
    say(f"step parity: launch {block}x{ppt}, case={case}")
    kw = dict(STEP_CASES[case], dry_mode=1, dry_vd=2e-3, wet_mode=1, wet_lambda=1e-4, wet_fraction=0.7,
              grid_fused=True)
    want = small.particles.copy()
    oc = om = None
    with make_ctx(fp, small) as ctx:
        ctx.set_launch(block, ppt)
        ctx.configure_grid(small.outheights, small.refine, 0, small.mass_scale)
        ctx.clear_grid()
        counter = (3, 0, 0, 0)
        for _ in range(2):
            cfg = base_cfg(fp, small, counter=counter, **kw)
            counter = ctx.step(cfg)
            orc.step(want, met_dict(small.met), pbl_dict(small.pbl), oracle_params(small, cfg))
            oc, om = orc.grid_accumulate(want, small.out_shape, small.outheights, small.refine, 0,
                                         small.mass_scale, oc, om)
        got = ctx.download_particles()
        counts, mass = ctx.download_grid()
    assert_particles_equal(got, want, f"launch {block}x{ppt} {case}")
    assert np.array_equal(counts, oc) and np.array_equal(mass, om)


def test_step_per_slot_forcing_matches_oracle(fp, orc, small):
    say("step parity: per-slot forcing")
    rng = np.random.default_rng(5)
    n = len(small.particles)
    vd = rng.uniform(0, 0.02, n).astype(np.float32)
    lam = rng.uniform(0, 1e-3, n).astype(np.float32)
    frac = rng.uniform(-0.2, 1.2, n).astype(np.float32)
    want = small.particles.copy()
    with make_ctx(fp, small) as ctx:
        ctx.upload_forcing(vd, lam, frac)
        cfg = base_cfg(fp, small, dry_mode=2, wet_mode=2)
        ctx.step(cfg)
        prm = oracle_params(small, cfg)
        prm.update(dry_vd_slots=vd, wet_lambda_slots=lam, wet_fraction_slots=frac)
        orc.step(want, met_dict(small.met), pbl_dict(small.pbl), prm)
        assert_particles_equal(ctx.download_particles(), want, "per-slot forcing")


def test_step_level_index_mode_matches_oracle(fp, orc, small):
    say("step parity: level-index vertical coordinate")
    w = small
    parts = w.particles.copy()
    parts["pos_z"] = np.random.default_rng(2).uniform(0, w.met.shape[2] - 1, len(parts))
    met = met_dict(w.met)
    met["level_heights"] = np.zeros(w.met.shape[2], np.float32)
    with fp.GpuTimestep(len(parts), w.met.shape, out_shape=None) as ctx:
        u0, v0, w0 = w.met.levels[0]
        u1, v1, w1 = w.met.levels[1]
        ctx.upload_met_bracket(u0, v0, w0, u1, v1, w1, met["level_heights"])
        ctx.upload_pbl(w.pbl.ustar, w.pbl.wstar, w.pbl.hmix, w.pbl.oli)
        ctx.upload_store(parts)
        cfg = base_cfg(fp, w, x_scale=1e-3, y_scale=1e-3, z_scale=1e-3)
        ctx.step(cfg)
        orc.step(parts, met, pbl_dict(w.pbl), oracle_params(w, cfg))
        assert_particles_equal(ctx.download_particles(), parts, "level index")


def test_single_species_context_matches_oracle(fp, orc, small):
    say("step parity: n_species=1 context")
    want = small.particles.copy()
    with make_ctx(fp, small, n_species=1) as ctx:
        cfg = base_cfg(fp, small, dry_mode=1, dry_vd=3e-3)
        ctx.step(cfg)
        orc.step(want, met_dict(small.met), pbl_dict(small.pbl), oracle_params(small, cfg))
        assert_particles_equal(ctx.download_particles(), want, "one species")


def test_species_beyond_context_count_are_rejected(fp, small):
    parts = small.particles[:10].copy()
    parts["mass"][:, 2] = 1.0
    with fp.GpuTimestep(10, small.met.shape, n_species=1) as ctx:
        with pytest.raises(fp.FpError):
            ctx.upload_store(parts)


@pytest.mark.parametrize("fused", [False, True])
def test_concentration_grid_matches_oracle(fp, orc, small, fused):
    say(f"grid parity: fused={fused}")
    want = small.particles.copy()
    with make_ctx(fp, small) as ctx:
        ctx.configure_grid(small.outheights, small.refine, 0, small.mass_scale)
        ctx.clear_grid()
        cfg = base_cfg(fp, small, grid_fused=fused)
        ctx.step(cfg)
        if not fused:
            ctx.accumulate_grid()
        counts, mass = ctx.download_grid()
    orc.step(want, met_dict(small.met), pbl_dict(small.pbl), oracle_params(small, cfg))
    oc, om = orc.grid_accumulate(want, small.out_shape, small.outheights, small.refine, 0,
                                 small.mass_scale)
    assert counts.sum() == np.count_nonzero(want["flags"] & 1)
    assert np.array_equal(counts, oc)
    assert np.array_equal(mass, om)


def test_grid_level_index_and_saturating_quantisation(fp, orc):
    say("grid parity: clustered cells, level-index output levels")
    from flexpart_gpu_amd.workloads import make_particles
    rng = np.random.default_rng(9)
    n = 50_000
    x = rng.integers(0, 3, n) + rng.uniform(0, 1, n)
    y = rng.integers(0, 2, n) + rng.uniform(0, 1, n)
    p = make_particles(n, x, y, rng.uniform(0, 5, n), rng.uniform(0, 20, n))
    p["flags"][rng.uniform(size=n) < 0.2] = 0
    with fp.GpuTimestep(n, (8, 8, 4), n_species=1, out_shape=(8, 8, 3)) as ctx:
        ctx.upload_store(p)
        ctx.configure_grid(None, (1, 1), 0, 3.0e8)
        ctx.clear_grid()
        ctx.accumulate_grid()
        counts, mass = ctx.download_grid()
    oc, om = orc.grid_accumulate(p, (8, 8, 3), None, (1, 1), 0, 3.0e8)
    assert np.array_equal(counts, oc) and np.array_equal(mass, om)


@pytest.mark.parametrize("block,ppt", [(256, 1), (256, 8), (1024, 2)])
def test_grid_block_tables_with_hot_cells_and_overflow_match_oracle(fp, orc, small, block, ppt):
    """Half the particles sit in three hot cells (every block's table absorbs
    them), the other half are spread over the whole output grid, so a block
    meets far more distinct cells than its 512-slot table holds and the
    overflow goes to the global grid directly; fused and standalone paths."""
    say(f"grid parity: hot cells + table overflow, block={block} ppt={ppt}")
    from flexpart_gpu_amd.workloads import make_particles
    rng = np.random.default_rng(31)
    n = len(small.particles)
    nx, ny, _ = small.met.shape
    hot = rng.integers(0, 3, n)
    spread = rng.uniform(size=n) < 0.5
    x = np.where(spread, rng.uniform(0, nx - 1, n), 10.2 + 7.0 * hot + rng.uniform(0, 0.05, n))
    y = np.where(spread, rng.uniform(0, ny - 1, n), 12.3 + 5.0 * hot + rng.uniform(0, 0.05, n))
    z = np.where(spread, rng.uniform(0, 12_000.0, n), 50.0 + rng.uniform(0, 5.0, n))
    p = make_particles(n, x, y, z, np.exp(rng.normal(size=n)))
    p["flags"][rng.uniform(size=n) < 0.05] = 0
    for fused in (False, True):
        want = p.copy()
        with make_ctx(fp, small) as ctx:
            ctx.set_launch(block, ppt)
            ctx.upload_store(p)
            ctx.configure_grid(small.outheights, small.refine, 0, small.mass_scale)
            ctx.clear_grid()
            cfg = base_cfg(fp, small, grid_fused=fused)
            if fused:
                ctx.step(cfg)
                orc.step(want, met_dict(small.met), pbl_dict(small.pbl), oracle_params(small, cfg))
            else:
                ctx.accumulate_grid()
            counts, mass = ctx.download_grid()
        oc, om = orc.grid_accumulate(want, small.out_shape, small.outheights, small.refine, 0,
                                     small.mass_scale)
        assert oc.max() > 1000 and np.count_nonzero(oc) > 2000
        assert np.array_equal(counts, oc), f"counts differ (fused={fused})"
        assert np.array_equal(mass, om), f"mass differs (fused={fused})"


@pytest.mark.parametrize("pattern", ["every_third", "all_active", "all_inactive", "single",
                                     "random"])
def test_compaction_matches_oracle(fp, orc, small, pattern):
    say(f"compaction parity: pattern={pattern}")
    p = small.particles.copy()
    n = len(p)
    i = np.arange(n)
    flags = {"every_third": i % 3 != 0, "all_active": np.ones(n, bool),
             "all_inactive": np.zeros(n, bool), "single": i == 12345,
             "random": np.random.default_rng(4).uniform(size=n) < 0.4}[pattern]
    p["flags"] = flags.astype(np.uint32)
    with make_ctx(fp, small) as ctx:
        ctx.upload_store(p)
        a, smap = ctx.compact()
        oa, oout, omap = orc.compact(p)
        assert a == oa and ctx.particle_count() == a
        assert np.array_equal(smap, omap)
        assert_particles_equal(ctx.download_particles(), oout, "compaction")


@pytest.mark.parametrize("n", [1, 63, 64, 65, 2047, 2048, 2049, 4097])
def test_compaction_tile_edges_match_oracle(fp, orc, small, n):
    """Sizes around the wave (64) and the scatter's tile (2048 slots): a
    partial last tile, a single slot, exact multiples; then a compaction of
    the already compacted store (dispatch narrowed, possibly to zero)."""
    say(f"compaction tile edges: n={n}")
    p = small.particles[:n].copy()
    p["flags"] = (np.random.default_rng(n).uniform(size=n) < 0.5).astype(np.uint32)
    with fp.GpuTimestep(n, small.met.shape, n_species=4) as ctx:
        ctx.upload_store(p)
        a, smap = ctx.compact()
        oa, oout, omap = orc.compact(p)
        assert a == oa and np.array_equal(smap, omap)
        assert_particles_equal(ctx.download_particles(), oout, f"compaction n={n}")
        a2, smap2 = ctx.compact()
        oa2, oout2, omap2 = orc.compact(oout[:oa])
        assert a2 == oa2 == oa and np.array_equal(smap2, omap2)
        assert_particles_equal(ctx.download_particles(n=oa), oout2, f"second compaction n={n}")


def test_compaction_then_steps_then_release_into_free_slots(fp, orc, small):
    say("compaction in a loop: step, compact, release, step")
    want = small.particles.copy()
    with make_ctx(fp, small) as ctx:
        counter = (0, 0, 0, 0)
        for k in range(3):
            cfg = base_cfg(fp, small, counter=counter, dry_mode=1, dry_vd=4e-2, loss_mode=1,
                           loss_min_mass=0.6)
            counter = ctx.step(cfg)
            n = ctx.particle_count()
            orc.step(want, met_dict(small.met), pbl_dict(small.pbl), oracle_params(small, cfg),
                     n_dispatch=n)
            a, smap = ctx.compact()
            oa, out, omap = orc.compact(want[:n])
            want[:n] = out
            assert a == oa and np.array_equal(smap, omap)
        # release 100 new particles right after the active prefix (O-07)
        a = ctx.particle_count()
        fresh = small.particles[:100].copy()
        fresh["flags"] = 1
        slots = np.arange(a, a + 100, dtype=np.uint32)
        ctx.upload_particle_slots(slots, fresh)
        ctx.set_dispatch_count(a + 100)
        want[a:a + 100] = fresh
        cfg = base_cfg(fp, small, counter=counter)
        ctx.step(cfg)
        orc.step(want, met_dict(small.met), pbl_dict(small.pbl), oracle_params(small, cfg),
                 n_dispatch=a + 100)
        assert_particles_equal(ctx.download_particles(), want, "compact+release")


@pytest.mark.parametrize("bits,bounds", [(10, None), (4, None), (21, None),
                                         (8, (0.0, 48.0, 0.0, 40.0, 0.0, 3000.0))])
def test_morton_sort_matches_oracle(fp, orc, small, bits, bounds):
    say(f"morton parity: bits={bits} explicit={bounds is not None}")
    with make_ctx(fp, small) as ctx:
        n2o = ctx.sort_morton(bits, bounds)
        out, on2o = orc.sort_morton(small.particles, bits, bounds)
        assert np.array_equal(n2o, on2o)
        assert_particles_equal(ctx.download_particles(), out, "morton")


def test_sort_then_step_matches_oracle(fp, orc, small):
    say("sort + step parity")
    with make_ctx(fp, small) as ctx:
        ctx.sort_morton(10, want_map=False)
        cfg = base_cfg(fp, small)
        ctx.step(cfg)
        want, _ = orc.sort_morton(small.particles, 10)
        orc.step(want, met_dict(small.met), pbl_dict(small.pbl), oracle_params(small, cfg))
        assert_particles_equal(ctx.download_particles(), want, "sort+step")


def test_call_order_and_argument_errors_are_reported(fp, small):
    with fp.GpuTimestep(100, small.met.shape) as ctx:
        with pytest.raises(fp.FpError) as e:
            ctx.step(base_cfg(fp, small))
        assert e.value.code == fp.FP_E_STATE
    with make_ctx(fp, small) as ctx:
        for bad in (dict(n_substeps=5), dict(dt_seconds=0.0), dict(direction=0),
                    dict(wet_mode=9)):
            with pytest.raises(fp.FpError) as e:
                ctx.step(base_cfg(fp, small, **bad))
            assert e.value.code == fp.FP_E_INVALID
        with pytest.raises(fp.FpError):
            ctx.set_launch(100, 1)
        with pytest.raises(fp.FpError):
            ctx.sort_morton(0)
    with pytest.raises(fp.FpError):
        fp.GpuTimestep(0, (4, 4, 4))
