"""The oracle's source-term injection (release/mod.rs) against the
reference's own release tests and the closed form of stratified sampling.

Reference tests restated: no_release_outside_time_window (:391-405),
correct_count_released_inside_window (:407-428: 3, 6, 3 of 12 at a quarter,
three quarters and the end of a one-hour window),
released_positions_stay_within_bounds (:430-455).
"""
import numpy as np

GRID = dict(xlon0=0.0, ylat0=0.0, dx=1.0, dy=1.0, nx=40, ny=80)
T0 = 3600  # 01:00 of the reference's window, as seconds from midnight


def release(**kw):
    r = dict(lon_min=10.0, lon_max=12.0, lat_min=45.0, lat_max=46.0, z_min=100.0, z_max=200.0,
             mass_kg=1.0, particle_count=12, start_s=T0, end_s=T0 + 3600, release_point=0, class_id=0)
    r.update(kw)
    return r


def test_no_release_outside_time_window(orc):
    done = np.zeros(1, np.uint64)
    for t in (T0 - 1, T0 + 3601):
        p, st = orc.release_inject(GRID, [release()], done, t, t)
        assert st == 0 and len(p) == 0
    assert done[0] == 0


def test_counts_inside_the_window_are_the_reference_ones(orc):
    done = np.zeros(1, np.uint64)
    got = []
    for t in (T0 + 900, T0 + 2700, T0 + 3600):
        p, st = orc.release_inject(GRID, [release()], done, t, t)
        assert st == 0
        got.append(len(p))
    assert got == [3, 6, 3] and done[0] == 12


def test_positions_stay_within_bounds_and_fields_are_initialised(orc):
    done = np.zeros(1, np.uint64)
    p, st = orc.release_inject(GRID, [release(release_point=7, class_id=2)], done, T0 + 3600, 123)
    assert st == 0 and len(p) == 12
    lon = p["cell_x"] + p["pos_x"].astype(np.float64)
    lat = p["cell_y"] + p["pos_y"].astype(np.float64)
    assert np.all((lon >= 10) & (lon <= 12)) and np.all((lat >= 45) & (lat <= 46))
    assert np.all((p["pos_z"] >= 100) & (p["pos_z"] <= 200))
    assert np.all(p["flags"] == 1) and np.all(p["mass"][:, 0] == np.float32(1.0 / 12))
    assert np.all(p["mass"][:, 1:] == 0) and np.all(p["time"] == 123) and np.all(p["time_mem"] == 123)
    assert np.all(p["release_point"] == 7) and np.all(p["class_id"] == 2) and np.all(p["cbt"] == 0)


def test_stratified_sampling_closed_form(orc):
    """8 particles: ceil(cbrt 8) = 2 strata per axis, cell centres at 1/4 and
    3/4, index order x fastest, then y, then z."""
    done = np.zeros(1, np.uint64)
    r = release(particle_count=8, lon_min=10.0, lon_max=14.0, lat_min=40.0, lat_max=48.0, z_min=0.0,
                z_max=400.0, start_s=0, end_s=0)
    p, st = orc.release_inject(GRID, [r], done, 0, 0)
    assert st == 0 and len(p) == 8
    k = np.arange(8)
    want_lon = 10.0 + (k % 2 + 0.5) / 2 * 4.0
    want_lat = 40.0 + ((k // 2) % 2 + 0.5) / 2 * 8.0
    want_z = 0.0 + (k // 4 + 0.5) / 2 * 400.0
    assert np.array_equal(p["cell_x"] + p["pos_x"].astype(np.float64), want_lon)
    assert np.array_equal(p["cell_y"] + p["pos_y"].astype(np.float64), want_lat)
    assert np.array_equal(p["pos_z"].astype(np.float64), want_z)


def test_out_of_domain_and_full_store(orc):
    done = np.zeros(2, np.uint64)
    ok = release(particle_count=4, start_s=0, end_s=0)
    bad = release(particle_count=4, lon_min=38.0, lon_max=40.0, start_s=0, end_s=0, release_point=1)
    p, st = orc.release_inject(GRID, [ok, bad], done, 0, 0)
    # the first release completes; the second writes its in-domain samples and stops
    assert st == 1 and done[0] == 4 and done[1] == 0
    assert len(p) > 4 and np.all(p["release_point"][4:] == 1)
    done = np.zeros(1, np.uint64)
    p, st = orc.release_inject(GRID, [release()], done, T0 + 3600, 0, max_out=5)
    assert st == 2 and len(p) == 5 and done[0] == 0
