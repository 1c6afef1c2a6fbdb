"""Accuracy of the deterministic f32 transcendentals (fp_math.h).

The oracle and the HIP kernel both evaluate exp/log/sin/cos/pow with these
functions, so HIP-vs-oracle parity is bitwise; these sweeps pin them to
glibc (what the reference's Rust CPU path calls) and to the correctly
rounded value (float64 evaluation rounded once).
"""
import numpy as np
import pytest

N = 400_000


def ulps(a, b):
    ai = a.view(np.int32).astype(np.int64)
    bi = b.view(np.int32).astype(np.int64)
    ai = np.where(ai < 0, -(ai & 0x7FFFFFFF), ai)
    bi = np.where(bi < 0, -(bi & 0x7FFFFFFF), bi)
    return np.abs(ai - bi)


def _sweep(name, rng):
    if name == "exp":
        return [rng.uniform(-103.9, 88.7, N), rng.uniform(-1, 1, N), rng.uniform(-60, 0, N)], None
    if name == "log":
        return [np.exp(rng.uniform(-87, 88, N)), rng.uniform(0.5, 2.0, N),
                rng.uniform(1.1754944e-38, 1.0, N), rng.uniform(1e-45, 1e-38, N // 10)], None
    if name in ("sin", "cos"):
        return [rng.uniform(0.0, 6.2831855, N), rng.uniform(-6.3, 6.3, N)], None
    raise KeyError(name)


@pytest.mark.parametrize("name", ["exp", "log", "sin", "cos"])
def test_f32_functions_within_one_ulp_of_glibc_and_correct_rounding(orc, name):
    rng = np.random.default_rng(1234)
    xs, _ = _sweep(name, rng)
    for x in xs:
        x = x.astype(np.float32)
        ours = orc.math_eval(name, x)
        libm = orc.math_eval(name + "_libm", x)
        cr = orc.math_eval(name + "_d", x)
        assert ulps(ours, libm).max() <= 1, name
        assert ulps(ours, cr).max() <= 1, name


@pytest.mark.parametrize("y", [np.float32(1.0) / np.float32(3.0), np.float32(2.0) / np.float32(3.0),
                               -np.float32(1.0) / np.float32(3.0), np.float32(0.8)])
def test_powf_exponents_of_the_hanna_profiles_are_correctly_rounded(orc, y):
    rng = np.random.default_rng(99)
    for x in (rng.uniform(0.0, 1.0, N), rng.uniform(1e-3, 1.0, N), rng.uniform(12.0, 500.0, N)):
        x = x.astype(np.float32)
        yy = np.full_like(x, y)
        ours = orc.math_eval("pow", x, yy)
        assert ulps(ours, orc.math_eval("pow_d", x, yy)).max() == 0
        assert ulps(ours, orc.math_eval("pow_libm", x, yy)).max() <= 1


def test_powf_general_arguments_and_special_cases(orc):
    rng = np.random.default_rng(5)
    x = np.exp(rng.uniform(-20, 20, N)).astype(np.float32)
    y = rng.uniform(-3, 3, N).astype(np.float32)
    ok = np.isfinite(orc.math_eval("pow_d", x, y))
    ours = orc.math_eval("pow", x, y)
    assert ulps(ours[ok], orc.math_eval("pow_d", x, y)[ok]).max() <= 1
    sx = np.array([0.0, 0.0, 0.0, 1.0, 2.0, np.inf, 4.0], np.float32)
    sy = np.array([0.8, 0.0, -1.0, 7.0, 0.0, 0.5, 0.5], np.float32)
    got = orc.math_eval("pow", sx, sy)
    assert got.tolist() == [0.0, 1.0, np.inf, 1.0, 1.0, np.inf, 2.0]


def test_exp_log_special_values(orc):
    x = np.array([0.0, -np.inf, np.inf, 1.0, 89.5, -105.0], np.float32)
    e = orc.math_eval("exp", x)
    assert e[0] == 1.0 and e[1] == 0.0 and e[2] == np.inf and e[4] == np.inf and e[5] == 0.0
    assert e[3] == np.float32(np.e)
    # every argument at or below -104 (the clamp) gives exactly 0, NaN stays NaN
    low = orc.math_eval("exp", np.array([-104.0, -104.5, -150.0, -1e30], np.float32))
    assert (low.view(np.uint32) == 0).all()
    assert np.isnan(orc.math_eval("exp", np.array([np.nan], np.float32)))[0]
    lg = orc.math_eval("log", np.array([1.0, 0.0, np.inf, 2.0], np.float32))
    assert lg[0] == 0.0 and lg[1] == -np.inf and lg[2] == np.inf
    assert lg[3] == np.float32(np.log(2.0))


def test_box_muller_radius_argument_range(orc):
    """Every Box-Muller radius argument -2 log u of the step, u the 24-bit
    uniform of a Philox word floored at 2^-126 (langevin_fused.wgsl:186-192),
    lies in [2^-96, FLT_MAX]: the range where the step's square root needs no
    range test.  All 2^24 values of u."""
    u = (np.arange(1 << 24, dtype=np.float64) / float(1 << 24)).astype(np.float32)
    u = np.maximum(u, np.float32(1.1754944e-38))
    arg = (np.float32(-2.0) * orc.math_eval("log", u)).astype(np.float32)
    assert np.all(np.isfinite(arg))
    assert float(arg.min()) >= 2.0 ** -96 and float(arg.min()) > 1.1e-7
    assert float(arg.max()) <= 175.0
