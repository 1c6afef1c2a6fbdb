"""Is the reference's vertical Langevin model well mixed?

A well-mixed model keeps particles that start uniformly distributed over the
boundary layer, with velocities drawn from the local sigma_w, uniformly
distributed (Thomson's criterion).  Run on the oracle (the reference's
production step, langevin_fused.wgsl) in a horizontally uniform neutral
boundary layer 3 km deep with no mean wind, the distribution does not stay
uniform: the particles leave the lowest layers and the mean height rises by
tens of metres per hour, far beyond the Monte Carlo error.  The cause is the
neutral profile's height derivative: the reference sets
dsigma_w/dz = -2e-4 sigma_w (langevin_fused.wgsl:239-240,310-311; physics/hanna.rs:73),
which is the derivative of sigma_w(z) = 1.3 u* exp(-2e-4 z/u*) only for
u* = 1 m/s; with the consistent -2e-4 sigma_w / u* (the binary64 model with
that one change) the rise is less than a quarter as large.

The adjoint's emission gradient is compared with the reference's backward
run in tests/test_gpu_adjoint_backward.py; this drift is why sources above
and below the receptor see forward and backward runs differ there."""
import numpy as np
import pytest

import forward64 as F

N, STEPS, H = 100_000, 12, 3000.0
NX, NY, NZ = 8, 8, 16


@pytest.fixture(scope="module")
def layer(workloads):
    lh = workloads.level_heights(NZ, 12_000.0)
    zero = np.zeros(NX * NY * NZ, np.float32)
    met = dict(shape=(NX, NY, NZ), u0=zero, v0=zero, w0=zero, u1=zero, v1=zero, w1=zero, level_heights=lh)
    m = NX * NY
    pbl = dict(shape=(NX, NY), ustar=np.full(m, 0.3, np.float32), wstar=np.zeros(m, np.float32),
               hmix=np.full(m, H, np.float32), oli=np.zeros(m, np.float32))
    rng = np.random.default_rng(1)
    z = rng.uniform(0.01, H, N)
    sigw = F.hanna(np.full(N, 0.3), np.zeros(N), np.full(N, np.inf), np.full(N, H), z)["sigw"]
    w = sigw * rng.normal(size=N)
    return lh, met, pbl, z, w


def profile(z):
    h, _ = np.histogram(z, np.linspace(0.0, H, 11))
    return h / (len(z) / 10.0)


def oracle_run(orc, workloads, layer):
    lh, met, pbl, z, w = layer
    p = workloads.make_particles(N, 4.0, 4.0, z.astype(np.float32), 1.0 / N)
    p["turb_w"] = w.astype(np.float32)
    ctr = (0, 0, 0, 0)
    means = []
    for _ in range(STEPS):
        ctr = tuple(int(v) for v in orc.step(p, met, pbl, dict(dt_seconds=600.0, alpha=0.0, x_scale=4e-5,
                                                                y_scale=4e-5, key=(3, 9), counter=ctr, n_threads=8)))
        means.append(float(p["pos_z"].astype(np.float64).mean()))
    return p["pos_z"].astype(np.float64), means


def binary64_run(layer, consistent_derivative):
    lh, met, pbl, z, w = layer
    m64 = F.Met64(met["shape"], *[met[k] for k in ("u0", "v0", "w0", "u1", "v1", "w1")], lh)
    p64 = {k: np.asarray(pbl[k], np.float64) for k in ("ustar", "wstar", "hmix", "oli")}
    st = dict(x=np.full(N, 4.0), y=np.full(N, 4.0), z=z.copy(), tu=np.zeros(N), tv=np.zeros(N), tw=w.copy(),
              cbt=np.ones(N), mass=np.ones(N), active=np.ones(N, bool))
    vertical, hanna = F.vertical, F.hanna
    try:
        if consistent_derivative:
            def vertical_u(ust, wst, ol, h, zin, neu, uns):
                s, d, t = vertical(ust, wst, ol, h, zin, neu, uns)
                return s, np.where(neu, d / ust, d), t

            def hanna_u(*a):
                r = hanna(*a)
                r["dsw"] = np.where(r["neu"], r["dsw"] / r["ust"], r["dsw"])
                return r
            F.vertical, F.hanna = vertical_u, hanna_u
        ctr = [0, 0, 0, 0]
        import oracle
        for _ in range(STEPS):
            F.step(st, m64, p64, dict(dt=600.0, alpha=0.0, x_scale=4e-5, y_scale=4e-5, key=(3, 9), counter=ctr,
                                      n_slots=N, slots=np.arange(N)))
            ctr = [int(v) for v in oracle.counter_add(ctr, 2 * N)]
    finally:
        F.vertical, F.hanna = vertical, hanna
    return st["z"]


def test_uniform_initial_layer_does_not_stay_uniform(orc, workloads, layer):
    z0 = layer[3]
    z, means = oracle_run(orc, workloads, layer)
    mc = np.std(z0) / np.sqrt(N)
    rise = means[-1] - z0.mean()
    prof = profile(z)
    # a well-mixed model keeps the mean within a few Monte Carlo errors and
    # every 300 m bin within a few percent; the reference's model does neither
    assert rise > 20 * mc and rise > 40.0, rise
    assert prof[0] < 0.85 and prof[-1] > 1.08, prof
    # the drift accumulates: the mean rises monotonically, about 5-8 m per step
    assert all(b > a for a, b in zip(means[2::3], means[5::3]))


def test_the_drift_comes_from_the_neutral_sigma_w_derivative(orc, workloads, layer):
    z0 = layer[3]
    z32, _ = oracle_run(orc, workloads, layer)
    ref = binary64_run(layer, False)
    fix = binary64_run(layer, True)
    # the binary64 model reproduces the oracle's drift ...
    assert abs(ref.mean() - z32.mean()) < 3.0
    # ... and with the dimensionally consistent derivative most of it is gone
    assert (fix.mean() - z0.mean()) < 0.25 * (ref.mean() - z0.mean())
    assert profile(fix)[0] > 0.93
