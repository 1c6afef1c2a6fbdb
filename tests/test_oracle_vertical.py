"""The oracle's hybrid sigma-pressure to height transform
(orc_hybrid_to_height; io/vertical_transform.rs:106-334), pinned by the
reference's own tests (vertical_transform.rs:360-510) and by the hypsometric
closed forms: isothermal layers dz = R/g T ln(p1/p2), layers with a linear
change of virtual temperature dz = R/g ln(p1/p2) (Tv2 - Tv1) / ln(Tv2/Tv1)
(the log-mean temperature), and the horizontal mean of the column heights."""
import math

import numpy as np
import pytest

SCALE = 287.05 / 9.81


def coords(bk, ak=(0.0, 0.0, 0.0)):
    return np.float32(ak), np.float32(bk)


def test_hybrid_pressure_is_a_plus_b_ps(orc):
    # vertical_transform.rs:360-377
    ak, bk = coords((1.0, 0.8, 0.6))
    code, _, p, _, _ = orc.hybrid_to_height(np.float32([[100_000.0], [90_000.0]]),
                                            np.full((2, 1, 3), 290.0, np.float32), ak, bk)
    assert code == 0
    assert p[0, 0].tolist() == pytest.approx([100_000.0, 80_000.0, 60_000.0], abs=1e-2)
    assert p[1, 0].tolist() == pytest.approx([90_000.0, 72_000.0, 54_000.0], abs=1e-2)
    ak2 = np.float32([500.0, 300.0, 100.0])
    _, _, p2, _, _ = orc.hybrid_to_height(np.float32([[100_000.0]]), np.full((1, 1, 3), 290.0, np.float32), ak2, bk)
    assert p2[0, 0].tolist() == pytest.approx([100_500.0, 80_300.0, 60_100.0], abs=1e-2)


def test_isothermal_profile_matches_the_hypsometric_solution(orc):
    # vertical_transform.rs:379-417
    ak, bk = coords((1.0, 0.8, 0.6))
    code, _, _, z, mean = orc.hybrid_to_height(np.float32([[100_000.0]]), np.full((1, 1, 3), 300.0, np.float32),
                                               ak, bk, q=np.zeros((1, 1, 3), np.float32))
    z1 = SCALE * 300.0 * math.log(100_000.0 / 80_000.0)
    z2 = z1 + SCALE * 300.0 * math.log(80_000.0 / 60_000.0)
    assert code == 0
    assert z[0, 0].tolist() == pytest.approx([0.0, z1, z2], abs=1e-2)
    assert mean.tolist() == pytest.approx([0.0, z1, z2], abs=1e-2)


def test_linear_virtual_temperature_uses_the_log_mean_temperature(orc):
    ak, bk = coords((1.0, 0.9, 0.75))
    t = np.float32([[[300.0, 290.0, 270.0]]])
    code, _, p, z, _ = orc.hybrid_to_height(np.float32([[101_000.0]]), t, ak, bk)
    assert code == 0
    want = [0.0]
    for k in (1, 2):
        t1, t2 = float(t[0, 0, k - 1]), float(t[0, 0, k])
        p1, p2 = float(p[0, 0, k - 1]), float(p[0, 0, k])
        want.append(want[-1] + SCALE * math.log(p1 / p2) * (t2 - t1) / math.log(t2 / t1))
    assert z[0, 0].tolist() == pytest.approx(want, rel=2e-6)
    # below the 0.2 K switch the layer uses the upper level's virtual temperature
    t_small = np.float32([[[300.0, 299.9, 299.85]]])
    _, _, p, z, _ = orc.hybrid_to_height(np.float32([[101_000.0]]), t_small, ak, bk)
    dz1 = SCALE * math.log(float(p[0, 0, 0]) / float(p[0, 0, 1])) * 299.9
    assert float(z[0, 0, 1]) == pytest.approx(dz1, rel=2e-6)


def test_moisture_raises_the_layers(orc):
    # vertical_transform.rs:419-449
    ak, bk = coords((1.0, 0.8, 0.6))
    args = (np.float32([[100_000.0]]), np.full((1, 1, 3), 290.0, np.float32), ak, bk)
    _, _, _, dry, _ = orc.hybrid_to_height(*args, q=np.zeros((1, 1, 3), np.float32))
    _, _, _, moist, _ = orc.hybrid_to_height(*args, q=np.full((1, 1, 3), 0.015, np.float32))
    assert moist[0, 0, 2] > dry[0, 0, 2]
    assert float(moist[0, 0, 2]) / float(dry[0, 0, 2]) == pytest.approx(1.0 + 0.608 * 0.015, rel=1e-5)
    # negative humidity counts as dry
    _, _, _, neg, _ = orc.hybrid_to_height(*args, q=np.full((1, 1, 3), -0.01, np.float32))
    assert np.array_equal(neg, dry)


def test_synthetic_profile_is_monotonic_and_the_mean_is_the_column_average(orc):
    # vertical_transform.rs:450-487
    ak, bk = coords((1.0, 0.85, 0.7))
    t = np.zeros((2, 2, 3), np.float32)
    t[..., 0], t[..., 1], t[..., 2] = 292.0, 286.0, 280.0
    ps = np.float32([[101_000.0, 99_000.0], [97_000.0, 95_000.0]])
    code, _, _, z, mean = orc.hybrid_to_height(ps, t, ak, bk)
    assert code == 0
    assert (np.diff(z, axis=2) > 0).all()
    assert mean.tolist() == pytest.approx(z.reshape(4, 3).astype(np.float64).mean(axis=0).tolist(), rel=1e-6)


def test_errors_and_their_precedence(orc):
    # vertical_transform.rs:489-509: a pressure inversion
    ak, bk = coords((1.0, 1.1, 0.8))
    code, ijk, *_ = orc.hybrid_to_height(np.float32([[100_000.0]]), np.full((1, 1, 3), 290.0, np.float32), ak, bk)
    assert code == 2 and ijk == (0, 0, 1)
    ak, bk = coords((1.0, 0.8, 0.6))
    t = np.full((3, 2, 3), 290.0, np.float32)
    ps = np.full((3, 2), 100_000.0, np.float32)
    t[1, 1, 2] = np.nan
    t[2, 0, 1] = -5.0
    code, ijk, *_ = orc.hybrid_to_height(ps, t, ak, bk)
    assert code == 3 and ijk == (1, 1, 2)                # the first column in (i, j) order
    ps[2, 1] = -1.0                                      # a pressure error anywhere comes first
    code, ijk, *_ = orc.hybrid_to_height(ps, t, ak, bk)
    assert code == 1 and ijk == (2, 1, 0)
    ps[2, 1] = 100_000.0
    t2 = np.full((1, 1, 3), 290.0, np.float32)
    t2[0, 0, 0] = 0.0                                    # level 0 is checked before the loop
    code, ijk, *_ = orc.hybrid_to_height(np.float32([[1e5]]), t2, ak, bk)
    assert code == 3 and ijk == (0, 0, 0)
    q = np.zeros((1, 1, 3), np.float32)
    q[0, 0, 1] = np.inf
    code, ijk, *_ = orc.hybrid_to_height(np.float32([[1e5]]), np.full((1, 1, 3), 290.0, np.float32), ak, bk, q=q)
    assert code == 3 and ijk == (0, 0, 1)


def test_deterministic_log_agrees_with_libm(orc):
    rng = np.random.default_rng(2)
    nx, ny, nz = 6, 5, 20
    bk = np.linspace(1.0, 0.1, nz).astype(np.float32)
    ak = np.linspace(0.0, 2000.0, nz).astype(np.float32)
    ps = rng.uniform(90_000, 103_000, (nx, ny)).astype(np.float32)
    t = (290.0 - 6.5e-3 * 1000 * np.arange(nz) * 0.6 + rng.normal(0, 1.0, (nx, ny, nz))).astype(np.float32)
    q = np.clip(rng.normal(0.005, 0.003, (nx, ny, nz)), -0.001, None).astype(np.float32)
    a = orc.hybrid_to_height(ps, t, ak, bk, q=q)
    b = orc.hybrid_to_height(ps, t, ak, bk, q=q, variant="libm")
    assert a[0] == 0 and b[0] == 0
    np.testing.assert_allclose(a[3], b[3], rtol=2e-6, atol=1e-3)
    np.testing.assert_allclose(a[4], b[4], rtol=2e-6, atol=1e-3)
