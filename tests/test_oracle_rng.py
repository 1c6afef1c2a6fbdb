"""Philox4x32-10, counter arithmetic and Box-Muller of the oracle.

Pinned by the reference's own known-answer tests (physics/rng.rs:167-215) and
by the published Random123 philox4x32-10 known-answer vectors, which fix
every round constant, the round function and the key schedule.
"""
import numpy as np
import pytest


def hexes(block):
    return [int(v) for v in block]


def test_philox_zero_vector_matches_reference_block(orc):
    # physics/rng.rs:171-177
    assert hexes(orc.philox4x32_10([0, 0, 0, 0], [0, 0])) == [
        0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]


def test_philox_all_ones_vector_matches_random123(orc):
    # Random123 kat_vectors: philox4x32 10 rounds, counter and key all ones
    assert hexes(orc.philox4x32_10([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)) == [
        0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]


def test_philox_pi_digits_vector_matches_random123(orc):
    # Random123 kat_vectors: counter/key from the hex digits of pi
    out = orc.philox4x32_10([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                            [0xA4093822, 0x299F31D0])
    assert hexes(out) == [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_philox_libm_build_is_the_same_integer_function(orc):
    c, k = [1, 2, 3, 4], [5, 6]
    assert hexes(orc.philox4x32_10(c, k, "libm")) == hexes(orc.philox4x32_10(c, k))


def test_philox_counter_add_handles_carry(orc):
    # physics/rng.rs:210-215
    assert hexes(orc.counter_add([0xFFFFFFFF, 0xFFFFFFFF, 7, 9], 1)) == [0, 0, 8, 9]


@pytest.mark.parametrize("base,offset", [
    ([0, 0, 0, 0], 5),
    ([0xFFFFFFF0, 0, 0, 0], 0x20),
    ([0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF], 1),
    ([123, 456, 789, 0], 0x1_0000_0001),
    ([0x89ABCDEF, 0xFFFFFFFF, 0x00000001, 0x7], (1 << 63) + 12345),
])
def test_counter_add_is_128_bit_modular_addition(orc, base, offset):
    v = sum(int(b) << (32 * i) for i, b in enumerate(base))
    r = (v + offset) % (1 << 128)
    want = [(r >> (32 * i)) & 0xFFFFFFFF for i in range(4)]
    assert hexes(orc.counter_add(base, offset)) == want


def test_philox_stream_uses_consecutive_counters(orc):
    key = [0xDECAFBAD, 0x12345678]
    base = [0xFFFFFFFE, 0, 2, 3]
    s = orc.philox_stream(base, key, 4)
    for i in range(4):
        assert hexes(s[i]) == hexes(orc.philox4x32_10(orc.counter_add(base, i), key))


def test_u32_to_uniform01_uses_high_24_bits_exactly(orc):
    # physics/rng.rs:89-96
    assert orc.u32_to_uniform01(0) == 0.0
    assert orc.u32_to_uniform01(0xFF) == 0.0
    assert orc.u32_to_uniform01(0x100) == np.float32(2.0 ** -24)
    assert orc.u32_to_uniform01(0xFFFFFFFF) == np.float32(1.0 - 2.0 ** -24)


def test_philox_uniform_distribution_mean_is_near_half(orc):
    # physics/rng.rs:191-208 (stream of 200k uniforms)
    s = orc.philox_stream([0x243F6A88, 0x85A308D3, 0, 0], [0xA4093822, 0x299F31D0], 50_000)
    u = (s.reshape(-1) >> 8).astype(np.float64) * 2.0 ** -24
    assert np.all((u >= 0) & (u < 1))
    assert abs(u.mean() - 0.5) < 0.01


def test_box_muller_closed_form_special_points(orc):
    # u1 = 0 -> theta = 0: (r, 0); u1 = 0.25 -> theta = pi/2 (f32): (~0, r)
    r = np.sqrt(-2.0 * np.log(np.float64(0.5)))
    g = orc.box_muller(0.5, 0.0)
    assert g[1] == 0.0 and abs(g[0] - r) < 2e-7
    g = orc.box_muller(0.5, 0.25)
    assert abs(g[0]) < 1e-6 and abs(g[1] - r) < 2e-7
    # u0 is clamped to the smallest normal float: radius sqrt(-2 ln 2^-126)
    g = orc.box_muller(0.0, 0.0)
    assert abs(g[0] - np.sqrt(2 * 126 * np.log(2.0))) < 1e-5


def test_box_muller_moments_on_philox_stream(orc):
    # mean 0, variance 1 of the Gaussian pairs drawn as in the step kernel
    s = orc.philox_stream([7, 0, 0, 0], [42, 0], 200_000)
    u = ((s >> 8).astype(np.float32) * np.float32(2.0 ** -24)).astype(np.float32)
    ru = np.clip(u[:, 0], np.float32(1.1754944e-38), np.float32(0.99999994))
    th = np.float32(6.2831855) * np.clip(u[:, 1], 0, np.float32(0.99999994))
    rad = np.sqrt(-2.0 * np.log(ru.astype(np.float64)))
    g = np.concatenate([rad * np.cos(th), rad * np.sin(th)])
    assert abs(g.mean()) < 0.01
    assert abs(g.var() - 1.0) < 0.01
    # and the oracle's f32 pairs agree with this float64 evaluation
    for i in range(0, 200_000, 20_011):
        o = orc.box_muller(float(u[i, 0]), float(u[i, 1]))
        assert abs(o[0] - rad[i] * np.cos(th[i])) < 2e-6 * max(1.0, rad[i])
        assert abs(o[1] - rad[i] * np.sin(th[i])) < 2e-6 * max(1.0, rad[i])
