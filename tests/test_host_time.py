"""Host-side temporal weight (fp_interpolation_alpha) against
interpolation_alpha (simulation/timeloop.rs:2086-2120): the closed form
(target - t0) / (t1 - t0) in binary64 rounded once to float, clamping or
strict rejection outside the bracket, and rejection of an empty bracket.
Host code only: no device is touched."""
import numpy as np
import pytest


def test_alpha_is_the_binary64_ratio_rounded_once(fp):
    rng = np.random.default_rng(11)
    for _ in range(2000):
        t0 = int(rng.integers(-10**12, 10**12))
        w = int(rng.integers(1, 10**9))
        t = t0 + int(rng.integers(0, w + 1))
        want = np.float32(np.float64(t - t0) / np.float64(w))
        assert fp.interpolation_alpha(t0, t0 + w, t) == float(want)
        assert fp.interpolation_alpha(t0, t0 + w, t, strict=True) == float(want)


def test_alpha_end_points_and_midpoint_are_exact(fp):
    assert fp.interpolation_alpha(0, 21600, 0) == 0.0
    assert fp.interpolation_alpha(0, 21600, 21600) == 1.0
    assert fp.interpolation_alpha(100, 300, 200) == 0.5
    # 600 s steps in a 6 h bracket: k / 36 rounded to float
    for k in range(37):
        assert fp.interpolation_alpha(0, 21600, 600 * k) == float(np.float32(k / 36.0))


def test_alpha_outside_the_bracket_clamps_or_fails(fp):
    assert fp.interpolation_alpha(0, 3600, -5) == 0.0
    assert fp.interpolation_alpha(0, 3600, 7200) == 1.0
    for bad in (-1, 3601):
        with pytest.raises(fp.FpError) as e:
            fp.interpolation_alpha(0, 3600, bad, strict=True)
        assert e.value.code == fp.FP_E_INVALID


def test_empty_or_reversed_bracket_is_rejected(fp):
    for t0, t1 in ((0, 0), (10, 5)):
        with pytest.raises(fp.FpError) as e:
            fp.interpolation_alpha(t0, t1, t0)
        assert e.value.code == fp.FP_E_INVALID


def test_ensemble_and_adjoint_descriptor_layouts(fp):
    import ctypes as C
    assert C.sizeof(fp.EnsDesc) == 40
    assert C.sizeof(fp.AdjDesc) == 56
