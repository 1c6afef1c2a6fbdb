"""Source-term injection on the device (fp_release_inject) against the
oracle's restatement of ReleaseManager::inject_for_time (release/mod.rs):
particles bit-identical (binary64 placement, f32 storage), emitted counts,
slot placement (consecutive after compaction, or an explicit free-slot
order), and the reference's error behaviour (partial writes before an
out-of-domain sample or a full store)."""
import numpy as np
import pytest

from conftest import assert_particles_equal

pytestmark = pytest.mark.gpu

GRID = dict(xlon0=-180.0, ylat0=-90.0, dx=1.0, dy=1.0, nx=360, ny=181)


def releases():
    return [
        dict(lon_min=10.0, lon_max=14.5, lat_min=40.0, lat_max=43.0, z_min=0.0, z_max=500.0, mass_kg=3.0,
             particle_count=1000, start_s=0, end_s=3600, release_point=0, class_id=0),
        dict(lon_min=-70.25, lon_max=-70.25, lat_min=-33.5, lat_max=-33.5, z_min=100.0, z_max=100.0,
             mass_kg=1.0, particle_count=27, start_s=600, end_s=600, release_point=1, class_id=0),
        dict(lon_min=100.0, lon_max=101.0, lat_min=10.0, lat_max=10.5, z_min=10.0, z_max=20.0, mass_kg=0.5,
             particle_count=64, start_s=1800, end_s=7200, release_point=2, class_id=1),
        dict(lon_min=-179.5, lon_max=178.5, lat_min=-80.0, lat_max=80.0, z_min=0.0, z_max=3000.0,
             mass_kg=1.0e3, particle_count=5000, start_s=0, end_s=86400, release_point=3, class_id=3),
    ]


def test_injection_matches_oracle_over_a_sequence_of_times(fp, orc):
    rel = releases()
    got_done = np.zeros(len(rel), np.uint64)
    want_done = np.zeros(len(rel), np.uint64)
    cursor = 100
    with fp.GpuTimestep(20_000, (8, 8, 4), n_species=1) as ctx:
        for t in (0, 600, 1200, 3600, 5400, 7200, 43200):
            n = ctx.inject_releases(GRID, rel, got_done, t, first_slot=cursor)
            want, st = orc.release_inject(GRID, rel, want_done, t, t)
            assert st == 0 and n == len(want), (t, n, len(want))
            assert np.array_equal(got_done, want_done)
            if n:
                got = ctx.download_particles(cursor + n)[cursor:]
                assert_particles_equal(got, want, f"injection at t={t}")
            cursor += n
        assert cursor - 100 == int(want_done.sum()) and want_done[1] == 27 and want_done[2] == 64


def test_injection_into_a_free_slot_list(fp, orc):
    rel = releases()[:3]
    slots = np.random.default_rng(3).permutation(4000)[:1200].astype(np.uint32)
    done = np.zeros(len(rel), np.uint64)
    with fp.GpuTimestep(4000, (8, 8, 4), n_species=1) as ctx:
        n = ctx.inject_releases(GRID, rel, done, 3600, slots=slots)
        want, st = orc.release_inject(GRID, rel, np.zeros(len(rel), np.uint64), 3600, 3600)
        assert st == 0 and n == len(want)
        got = ctx.download_particles()[slots[:n]]
        assert_particles_equal(got, want, "slot-list injection")


def test_reference_counts_in_a_one_hour_window(fp):
    # correct_count_released_inside_window (release/mod.rs:407-428)
    r = dict(lon_min=10.0, lon_max=12.0, lat_min=45.0, lat_max=46.0, z_min=100.0, z_max=200.0, mass_kg=1.0,
             particle_count=12, start_s=3600, end_s=7200, release_point=0, class_id=0)
    g = dict(xlon0=0.0, ylat0=0.0, dx=1.0, dy=1.0, nx=40, ny=80)
    done = np.zeros(1, np.uint64)
    with fp.GpuTimestep(32, (8, 8, 4), n_species=1) as ctx:
        counts = []
        cursor = 0
        for t in (3599, 7201, 3600 + 900, 3600 + 2700, 7200):
            n = ctx.inject_releases(g, [r], done, t, first_slot=cursor)
            counts.append(n)
            cursor += n
    assert counts == [0, 0, 3, 6, 3] and done[0] == 12


def test_out_of_domain_and_full_store_write_the_particles_before(fp, orc):
    ok = dict(lon_min=10.0, lon_max=12.0, lat_min=40.0, lat_max=41.0, z_min=0.0, z_max=10.0, mass_kg=1.0,
              particle_count=50, start_s=0, end_s=0, release_point=0, class_id=0)
    bad = dict(ok, lon_min=178.0, lon_max=179.9, release_point=1)
    with fp.GpuTimestep(1000, (8, 8, 4), n_species=1) as ctx:
        done = np.zeros(2, np.uint64)
        with pytest.raises(fp.FpError) as e:
            ctx.inject_releases(GRID, [ok, bad], done, 0, first_slot=0)
        assert e.value.code == fp.FP_E_INVALID
        want, st = orc.release_inject(GRID, [ok, bad], np.zeros(2, np.uint64), 0, 0)
        assert st == 1 and ctx.last_injected == len(want) and list(done) == [50, 0]
        assert_particles_equal(ctx.download_particles(len(want)), want, "partial injection")
        done = np.zeros(1, np.uint64)
        with pytest.raises(fp.FpError):
            ctx.inject_releases(GRID, [ok], done, 0, first_slot=980)
        assert ctx.last_injected == 20 and done[0] == 0
