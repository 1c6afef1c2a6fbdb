"""The binary64 restatement of the step (oracle/forward64.py), the adjoint
tests' finite-difference reference, against the float32 oracle: the same
particles, met, PBL, temporal weights and Philox counters give the same
trajectories to float32 rounding (the models differ only in the precision of
each operation).  A mistake in the restatement (a dropped term, a wrong
branch or sub-step order) moves the trajectories by far more."""
import numpy as np
import pytest

import forward64 as F

from conftest import pbl_dict


def smooth_case(workloads):
    import dataclasses
    w = workloads.small_random(10, shape=(48, 40, 24), n_levels=3)
    w.met = dataclasses.replace(w.met, seconds_between_levels=1800.0)
    counts = [600, 257, 64, 700, 1, 378]
    c = workloads.adjoint_case(sum(counts), len(counts), 4, met=w.met, pbl=w.pbl, x_scale=w.x_scale)
    c.counts[:] = counts
    c.coeff[:] = [0.3, -0.2, 0.5, 0.1]
    return c, w


def f32_run(orc, fp, c, pbl, n_steps, counter):
    from test_gpu_adjoint import initial_particles, perturbed_levels
    parts = initial_particles(fp, c)
    lv = perturbed_levels(c, c.coeff)
    T = int(c.met.seconds_between_levels)
    ctr = tuple(counter)
    for k in range(n_steps):
        t = int(k * c.dt_seconds)
        b = min(t // T, len(lv) - 2)
        met = dict(shape=c.met.shape, u0=lv[b][0], v0=lv[b][1], w0=lv[b][2], u1=lv[b + 1][0], v1=lv[b + 1][1],
                   w1=lv[b + 1][2], level_heights=c.met.level_heights)
        ctr = tuple(int(v) for v in orc.step(parts, met, pbl, dict(
            dt_seconds=c.dt_seconds, alpha=fp.interpolation_alpha(b * T, (b + 1) * T, t), x_scale=c.x_scale,
            y_scale=c.y_scale, key=c.key, counter=ctr)))
    return parts


def test_binary64_forward_tracks_the_oracle_in_a_smooth_neutral_layer(orc, fp, workloads):
    c, _ = smooth_case(workloads)
    n = c.pbl.ustar.size
    c.pbl = workloads.Pbl(c.pbl.shape, np.full(n, 0.3, np.float32), np.zeros(n, np.float32),
                          np.full(n, 3000.0, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32))
    c.position[:, 2] = np.random.default_rng(31).uniform(1000.0, 1800.0, len(c.position)).astype(np.float32)
    got = f32_run(orc, fp, c, pbl_dict(c.pbl), 6, (5, 0, 0, 0))
    _, _, st = F.adjoint_forward(c, c.emission, c.position, c.coeff, np.zeros((1, 3)), None, 6, 7, 9, 300.0,
                                 (5, 0, 0, 0))
    x = got["cell_x"] + got["pos_x"].astype(np.float64)
    y = got["cell_y"] + got["pos_y"].astype(np.float64)
    assert np.max(np.abs(x - st["x"])) < 5e-5 and np.max(np.abs(y - st["y"])) < 5e-5
    assert np.max(np.abs(got["pos_z"] - st["z"])) < 5e-3
    assert np.max(np.abs(got["turb_w"] - st["tw"])) < 5e-6
    # the run moves the particles: the agreement is not trivial
    assert np.median(np.abs(got["pos_z"] - np.repeat(c.position[:, 2], c.counts))) > 10.0


def test_binary64_forward_tracks_the_oracle_across_regimes_and_reflections(orc, fp, workloads):
    c, w = smooth_case(workloads)
    c.position[:, 2] = np.random.default_rng(3).uniform(5.0, 600.0, len(c.position)).astype(np.float32)
    got = f32_run(orc, fp, c, pbl_dict(w.pbl), 3, (9, 0, 0, 0))
    _, _, st = F.adjoint_forward(c, c.emission, c.position, c.coeff, np.zeros((1, 3)), None, 3, 7, 9, 300.0,
                                 (9, 0, 0, 0))
    z = got["pos_z"].astype(np.float64)
    dz = np.abs(z - st["z"]) / np.maximum(np.abs(st["z"]), 1.0)
    assert np.quantile(dz, 0.99) < 1e-4
    assert np.mean(got["cbt"] == st["cbt"]) > 0.99  # reflections agree
    x = got["cell_x"] + got["pos_x"].astype(np.float64)
    assert np.quantile(np.abs(x - st["x"]), 0.99) < 5e-5


@pytest.mark.parametrize("mutation", ["no_drift", "swap_normals", "no_reflection"])
def test_the_comparison_detects_a_planted_mistake(orc, fp, workloads, monkeypatch, mutation):
    """The tolerance above separates a correct restatement from a wrong one."""
    c, w = smooth_case(workloads)
    c.position[:, 2] = np.random.default_rng(3).uniform(5.0, 600.0, len(c.position)).astype(np.float32)
    got = f32_run(orc, fp, c, pbl_dict(w.pbl), 3, (9, 0, 0, 0))
    if mutation == "no_drift":
        real = F.ou_v
        monkeypatch.setattr(F, "ou_v", lambda p, s, d, t, dt, rho, e: real(p, s, 0.0 * d, t, dt, rho, e))
    elif mutation == "swap_normals":
        real = F.normals
        monkeypatch.setattr(F, "normals", lambda *a: (lambda o: [o[1], o[0]] + o[2:])(real(*a)))
    else:
        monkeypatch.setattr(F, "reflect", lambda z, tw, cbt, zmin, h: (z, tw, cbt))
    _, _, st = F.adjoint_forward(c, c.emission, c.position, c.coeff, np.zeros((1, 3)), None, 3, 7, 9, 300.0,
                                 (9, 0, 0, 0))
    z = got["pos_z"].astype(np.float64)
    dz = np.abs(z - st["z"]) / np.maximum(np.abs(st["z"]), 1.0)
    x = got["cell_x"] + got["pos_x"].astype(np.float64)
    assert np.quantile(dz, 0.99) >= 1e-4 or np.quantile(np.abs(x - st["x"]), 0.99) >= 5e-5
