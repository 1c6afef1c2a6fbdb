"""Gridding, compaction and Morton sort of the oracle.

Known-answer vectors from the reference: gridding cells (gpu/gridding.rs:
446-501), compaction patterns (gpu/compaction.rs:625-740), Morton order
(particles/mod.rs:919-992); brute force for the rest.
"""
import numpy as np
import pytest


def particle_for_cell(fp, cx, cy, z, mass, active=True):
    p = np.zeros(1, fp.PARTICLE_DTYPE)
    p["cell_x"], p["cell_y"], p["pos_z"] = cx, cy, z
    p["mass"][0, 0] = mass
    p["flags"] = 1 if active else 0
    return p


def test_gridding_simple_particle_pattern_matches_expected_cells(orc, fp):
    # gpu/gridding.rs:446-501
    parts = np.concatenate([
        particle_for_cell(fp, 1, 0, 0.2, 1.5),
        particle_for_cell(fp, 1, 0, 0.8, 0.5),
        particle_for_cell(fp, 2, 1, 1.2, 2.0),
        particle_for_cell(fp, 0, 2, 0.0, 3.0, active=False),
        particle_for_cell(fp, 3, 2, 1.99, 4.25),
    ])
    counts, mass = orc.grid_accumulate(parts, (4, 3, 2), mass_scale=1000.0)
    counts = counts.reshape(4, 3, 2)
    mass = mass.reshape(4, 3, 2) / 1000.0
    expected = {(1, 0, 0): (2, 2.0), (2, 1, 1): (1, 2.0), (3, 2, 1): (1, 4.25)}
    for ix in range(4):
        for iy in range(3):
            for iz in range(2):
                c, m = expected.get((ix, iy, iz), (0, 0.0))
                assert counts[ix, iy, iz] == c
                assert abs(mass[ix, iy, iz] - m) < 1e-6


def test_gridding_output_heights_pick_first_level_above(orc, fp):
    # concentration_gridding.wgsl:81-93: first k with z <= outheights[k], else top
    oh = (100.0, 500.0, 1000.0, 3000.0)
    zs = [0.0, 100.0, 100.01, 499.0, 2999.0, 3000.0, 8000.0]
    want = [0, 0, 1, 1, 3, 3, 3]
    parts = np.concatenate([particle_for_cell(fp, 0, 0, z, 1.0) for z in zs])
    for p, k in zip(parts, want):
        counts, _ = orc.grid_accumulate(p.reshape(1), (2, 2, 4), outheights=oh)
        assert counts.reshape(2, 2, 4)[0, 0, k] == 1


def test_gridding_quantisation_rounds_saturates_and_wraps(orc, fp):
    # quantize_mass_to_u32 (concentration_gridding.wgsl:95-104): round half up,
    # non-positive -> 0, saturate at 2^32-1; the per-cell sum wraps mod 2^32
    for m, scale, q in [(1.2345e-3, 1e3, 1), (1.5e-3, 1e3, 2), (-1.0, 1e3, 0), (0.0, 1e3, 0),
                        (5.0, 1e9, 0xFFFFFFFF)]:
        _, mass = orc.grid_accumulate(particle_for_cell(fp, 0, 0, 0.0, m), (1, 1, 1), mass_scale=scale)
        assert int(mass[0]) == q
    two = np.concatenate([particle_for_cell(fp, 0, 0, 0.0, 3.0)] * 2)
    _, mass = orc.grid_accumulate(two, (1, 1, 1), mass_scale=1e9)
    assert int(mass[0]) == (2 * 3_000_000_000) % (1 << 32)


def test_gridding_refined_output_grid_maps_fractional_position(orc, fp):
    # refine r: output cell = cell * r + floor(frac * r); r = 1 is the met cell
    p = particle_for_cell(fp, 3, 5, 0.0, 1.0)
    p["pos_x"], p["pos_y"] = 0.75, 0.25
    counts, _ = orc.grid_accumulate(p, (8, 12, 1), refine=(2, 2))
    assert np.nonzero(counts)[0].tolist() == [(7 * 12 + 10) * 1 + 0]
    counts, _ = orc.grid_accumulate(p, (8, 12, 1))
    assert np.nonzero(counts)[0].tolist() == [3 * 12 + 5]
    # clamped at the upper edge (met cell n-1 maps past the output grid)
    q = particle_for_cell(fp, 3, 5, 0.0, 1.0)
    counts, _ = orc.grid_accumulate(q, (7, 11, 1), refine=(2, 2))
    assert np.nonzero(counts)[0].tolist() == [6 * 11 + 10]


def _parts(fp, n, inactive):
    p = np.zeros(n, fp.PARTICLE_DTYPE)
    p["flags"] = [0 if inactive(i) else 1 for i in range(n)]
    p["pos_x"] = np.arange(n, dtype=np.float32) * 1e-3
    p["time"] = np.arange(n)
    return p


@pytest.mark.parametrize("n,inactive", [
    (1000, lambda i: i % 3 == 0),     # every third inactive
    (512, lambda i: False),           # all active
    (256, lambda i: True),            # all inactive
    (100, lambda i: i != 42),         # single active
    (300, lambda i: i % 5 == 0),      # non power of two
    (1000, lambda i: i % 2 == 0),     # multi-workgroup coverage
])
def test_compaction_patterns(orc, fp, n, inactive):
    # gpu/compaction.rs:625-740
    p = _parts(fp, n, inactive)
    a, out, smap = orc.compact(p)
    expected = [i for i in range(n) if not inactive(i)]
    assert a == len(expected)
    assert smap.tolist() == expected
    assert np.array_equal(out[:a].view(np.uint32), p[expected].view(np.uint32))
    # tail: deactivated copy of the slot itself (compaction.wgsl:212-216)
    tail = p[a:].copy()
    tail["flags"] &= ~np.uint32(1)
    assert np.array_equal(out[a:].view(np.uint32), tail.view(np.uint32))


def test_morton_encode_interleaves_bits(orc):
    rng = np.random.default_rng(0)
    for _ in range(200):
        x, y, z = (int(v) for v in rng.integers(0, 1 << 21, 3))
        want = 0
        for b in range(21):
            want |= ((x >> b) & 1) << (3 * b) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b + 2)
        assert orc.morton_encode(x, y, z) == want


def test_morton_key_increases_for_x_cells(orc, fp):
    # particles/mod.rs:919-935
    p = np.zeros(3, fp.PARTICLE_DTYPE)
    p["cell_x"] = [0, 1, 2]
    p["flags"] = 1
    k = orc.morton_keys(p, 2, bounds=(0, 3, 0, 1, 0, 1))
    assert k[0] < k[1] < k[2]


def test_spatial_sort_orders_active_and_keeps_inactive_stable(orc, fp):
    # particles/mod.rs:960-992: new_to_old == [2, 0, 1, 3, 4, 5]
    p = np.zeros(6, fp.PARTICLE_DTYPE)
    p["cell_x"][:3] = [3, 1, 2]
    p["flags"][:3] = [1, 0, 1]
    p["mass"][:3, 0] = 1.0
    out, n2o = orc.sort_morton(p, 4, bounds=(0, 4, 0, 1, 0, 1))
    assert n2o.tolist() == [2, 0, 1, 3, 4, 5]
    assert out["cell_x"][0] == 2 and out["cell_x"][1] == 3
    assert out["flags"][2] == 0 and out["flags"][3] == 0
    keys = orc.morton_keys(p, 4, bounds=(0, 4, 0, 1, 0, 1))
    assert keys[1] == np.uint64(0xFFFFFFFFFFFFFFFF) and keys[3] == np.uint64(0xFFFFFFFFFFFFFFFF)


def test_spatial_sort_rejects_invalid_bit_counts(orc, fp):
    p = np.zeros(4, fp.PARTICLE_DTYPE)
    for bits in (0, 22):
        with pytest.raises(ValueError):
            orc.sort_morton(p, bits)


def test_spatial_sort_matches_brute_force_stable_sort(orc, fp, workloads):
    w = workloads.small_random(5000)
    p = w.particles
    for bits in (3, 10, 21):
        keys = orc.morton_keys(p, bits)
        out, n2o = orc.sort_morton(p, bits)
        assert n2o.tolist() == sorted(range(len(p)), key=lambda i: int(keys[i]))
        assert np.array_equal(out.view(np.uint32), p[n2o].view(np.uint32))


def test_active_extent_bounds_and_identity_when_nothing_active(orc, fp):
    p = np.zeros(5, fp.PARTICLE_DTYPE)
    out, n2o = orc.sort_morton(p, 10)
    assert n2o.tolist() == list(range(5))
    p["flags"][2] = 1
    p["cell_x"][2] = 7
    keys = orc.morton_keys(p, 10)
    assert keys[2] == 0  # a single active particle: zero span -> all axes quantise to 0
