"""The device evaluates the step's arithmetic primitives exactly as the CPU.

fp_math.h is compiled for both sides; this checks, over millions of inputs
that cover the ranges the step uses (and subnormal/overflow edges), that the
gfx950 code rounds identically: exp (with v_ldexp scaling into subnormals),
log, the restricted log, sin/cos, pow, and the correctly rounded sqrt and
division the compiler emits under -fhip-fp32-correctly-rounded-divide-sqrt.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 1_000_000


def bits_equal(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    same = (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))
    return bool(same.all()), int((~same).sum())


@pytest.fixture(scope="module")
def ctx(fp):
    with fp.GpuTimestep(16, (4, 4, 4), n_species=1) as c:
        yield c


def _inputs(rng):
    return {
        "exp": np.concatenate([rng.uniform(-105, 90, N), rng.uniform(-104, -87, N // 4),
                               rng.uniform(-1, 1, N // 4), [0.0, -0.0, np.inf, -np.inf, np.nan, 89.0,
                                                            -104.0, 88.72284]]),
        "log": np.concatenate([np.exp(rng.uniform(-103, 88, N)),
                               rng.uniform(1e-45, 1.2e-38, N // 4), [0.0, -1.0, np.inf, np.nan, 1.0]]),
        "log_unit": rng.uniform(1.1754944e-38, 1.0, N),
        "sin": rng.uniform(0.0, 6.2831855, N),
        "cos": rng.uniform(0.0, 6.2831855, N),
    }


@pytest.mark.parametrize("fn", ["exp", "log", "log_unit", "sin", "cos"])
def test_unary_primitives_match_cpu_bitwise(ctx, orc, fn):
    x = _inputs(np.random.default_rng(11))[fn].astype(np.float32)
    gpu = ctx.math_probe(fn, x)
    cpu = orc.math_eval("log" if fn == "log_unit" else fn, x)
    ok, bad = bits_equal(gpu, cpu)
    assert ok, f"{fn}: {bad} of {x.size} differ"


def test_log_reduction_division_exhaustive(ctx, orc):
    """Every float in [0.5, 1) (and its negative-exponent copy in the unit
    log): together they reach every reduced mantissa m in [sqrt(1/2), sqrt(2)),
    hence every operand pair of the log's f / (2 + f), which the device
    evaluates without the division's range scaling."""
    x = np.arange(0x3F000000, 0x3F800000, dtype=np.uint32).view(np.float32)
    for fn in ("log", "log_unit"):
        ok, bad = bits_equal(ctx.math_probe(fn, x), orc.math_eval("log", x))
        assert ok, f"{fn}: {bad} of {x.size} differ"
    y = (x * np.float32(2.0 ** -100)).astype(np.float32)
    ok, bad = bits_equal(ctx.math_probe("log_unit", y), orc.math_eval("log", y))
    assert ok, f"log_unit at 2^-100 scale: {bad} differ"


def test_pow_matches_cpu_bitwise(ctx, orc):
    rng = np.random.default_rng(12)
    xs, ys = [], []
    for y in (np.float32(2) / np.float32(3), -np.float32(1) / np.float32(3), np.float32(0.8),
              np.float32(1) / np.float32(3)):
        xs.append(rng.uniform(0, 1, N // 4))
        ys.append(np.full(N // 4, y))
    xs.append(np.exp(rng.uniform(-30, 30, N // 4)))
    ys.append(rng.uniform(-3, 3, N // 4))
    x = np.concatenate(xs).astype(np.float32)
    y = np.concatenate(ys).astype(np.float32)
    ok, bad = bits_equal(ctx.math_probe("pow", x, y), orc.math_eval("pow", x, y))
    assert ok, f"pow: {bad} differ"


def test_sqrt_and_division_are_correctly_rounded(ctx):
    rng = np.random.default_rng(13)
    u = rng.integers(0, 2 ** 32, N, dtype=np.uint64).astype(np.uint32)
    x = u.view(np.float32)
    x = x[np.isfinite(x)]
    ok, bad = bits_equal(ctx.math_probe("sqrt", np.abs(x)), np.sqrt(np.abs(x)))
    assert ok, f"sqrt: {bad} differ"
    v = rng.integers(0, 2 ** 32, x.size, dtype=np.uint64).astype(np.uint32).view(np.float32)
    keep = np.isfinite(v) & (v != 0)
    a, b = x[keep], v[keep]
    with np.errstate(all="ignore"):
        want = (a / b).astype(np.float32)
    ok, bad = bits_equal(ctx.math_probe("div", a, b), want)
    assert ok, f"div: {bad} differ"


def test_step_sqrt_exhaustive(ctx):
    """The step's square root (hardware sqrt and rsqrt, one fused correction)
    is correctly rounded for every float in [2^-96, FLT_MAX]: each result is
    checked on the device against the exact binary64 squares of the midpoints
    to its neighbours."""
    fails, first = ctx.math_sweep("sqrt_step", 0x0F800000, 0x7F7FFFFF)
    assert fails == 0, f"{fails} incorrectly rounded, first at bits {first:#010x}"


def test_step_sqrt_sweep_detects_misrounding(ctx):
    """Below 2^-96 the residual of the same sequence underflows: the sweep
    must report failures there (the exact test has teeth)."""
    fails, first = ctx.math_sweep("sqrt_step", 0x00800000, 0x01000000)
    assert fails > 0 and 0x00800000 <= first <= 0x01000000


def _level_tables():
    k = np.arange(60, dtype=np.float64)
    workload = (15000.0 * (k / 59) ** 2 + 10.0).astype(np.float32)
    # levels crowded near the ground: the first bands hold up to 40 levels
    crowded = np.concatenate([0.25 + 0.05 * np.arange(40), 40.0 + 250.0 * np.arange(1, 21)]).astype(np.float32)
    # repeated heights (the scan stops at the first of a run)
    ties = np.array([10, 10, 20, 20, 20, 50, 50, 400, 400, 401, 5000, 5000], np.float32)
    ground = workload.copy()
    ground[0] = 0.0       # a first level at 0 m: the scan lookup
    return {"workload": workload, "crowded": crowded, "ties": ties, "ground": ground,
            "two": np.array([5.0, 3000.0], np.float32), "index": np.zeros(8, np.float32)}


def _level_queries(lh, rng):
    top = float(max(lh[-1], 1.0))
    edges = np.asarray(lh, np.float32)
    near = np.concatenate([edges, np.nextafter(edges, np.float32(np.inf)), np.nextafter(edges, np.float32(-np.inf))])
    inv_w = np.float32(512.0) / np.float32(top)
    bands = (np.arange(513, dtype=np.float32) / inv_w).astype(np.float32)
    near_bands = np.concatenate([bands, np.nextafter(bands, np.float32(np.inf)),
                                 np.nextafter(bands, np.float32(-np.inf))])
    special = np.array([0.0, -0.0, 1e-30, -1e-30, 1e-45, -5.0, np.nan, np.inf, -np.inf, 3e38, -3e38,
                        top, top * 1.5], np.float32)
    return np.concatenate([rng.uniform(-0.1 * top, 1.2 * top, 200_000).astype(np.float32),
                           rng.uniform(0.0, float(edges[min(3, len(edges) - 1)]) + 1.0, 50_000).astype(np.float32),
                           near, near_bands, special]).astype(np.float32)


@pytest.mark.parametrize("table", ["workload", "crowded", "ties", "ground", "two", "index"])
def test_level_lookup_matches_oracle(fp, orc, table):
    """The step's height -> fractional level lookup equals the reference's
    linear scan (meters_to_grid_level) bit for bit: at, just above and just
    below every level and every band edge of the lookup's band index, below
    the first level, above the top, for NaN, infinities and signed zeros, on
    tables that take the two-comparison path (workload heights, crowded low
    levels that overflow it into the scan, repeated heights, two levels), the
    plain scan (first level at 0 m) and level-index mode."""
    lh = _level_tables()[table]
    nz = lh.size
    z = _level_queries(lh, np.random.default_rng(21))
    with fp.GpuTimestep(16, (4, 4, nz), n_species=1) as c:
        met = np.zeros(16 * nz, np.float32)
        c.upload_met_bracket(met, met, met, met, met, met, lh)
        gpu = c.math_probe("level", z)
    cpu = orc.level_eval(lh, z)
    ok, bad = bits_equal(gpu, cpu)
    assert ok, f"{table}: {bad} of {z.size} differ"
