"""The adjoint's emission gradient against a backward (receptor-oriented) run.

For a receptor observing the cloud-in-cell concentration c_R and sources of
unit emission, dc_R/dq_s = E_s[w_R(x_T)]: the mean receptor kernel weight of
source s's particles after T (the pathwise gradient the adjoint returns).
The reference's backward mode (advection with -dt, turbulence and deposition
with +dt; timeloop.rs:20-23, 1724-1746, tests/integration/
source_receptor_consistency.rs) releases particles at the receptor and
reads them at the sources: B_s = E_R[w_s(x_{-T})] with the same kernel
centred on source s.  In a horizontally uniform wind the horizontal
transport is mirror-symmetric, so for sources at the receptor's level G_s
and B_s estimate the same number and agree within their Monte Carlo errors
(independent Philox streams).  Vertically, the neutral Hanna profile
(sigma_w falling with height) with the drift of update_vertical_component
is not well mixed (tests/test_oracle_well_mixed.py: a uniform layer loses
its lowest particles and rises): forward and backward particles both rise
on average (about 50 m in 2 h here), so a source 300 m below the receptor
is seen more by the forward run and one 300 m above less.  To first order a
common rise D translates both clouds, and the backward weight read at the
source node raised by 2 D is the forward weight; the off-level differences
must have the sign and size of that prediction (within the translation
model's own error), and the below + above sums agree.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NX, NY, NZ = 24, 20, 16
OBS_NZ, OBS_DZ = 12, 300.0
N_STEPS, DT = 12, 600.0
SCALE = 1.0 / 25_000.0          # grid units per metre (25 km cells)
RECEPTOR = (12, 10, 5)          # node (i, j, l): z = 1500 m
SOURCES = [(x, y, l) for x in (10, 11) for y in (9, 10, 11) for l in (4, 5, 6)]
N_FWD, N_BWD = 20_000, 40_000


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def tent(d):
    return np.clip(1.0 - np.abs(d), 0.0, None)


def kernel_weight(p, node, dz_m=0.0):
    """Cloud-in-cell weight of particles p for lattice node (i, j, l)
    (the adjoint's observation operator: met columns x levels OBS_DZ apart),
    with the node raised by dz_m metres."""
    x = p["cell_x"].astype(np.float64) + p["pos_x"]
    y = p["cell_y"].astype(np.float64) + p["pos_y"]
    z = (p["pos_z"].astype(np.float64) - dz_m) / OBS_DZ
    return tent(x - node[0]) * tent(y - node[1]) * tent(z - node[2])


@pytest.fixture(scope="module")
def setup(workloads):
    lh = workloads.level_heights(NZ, 12_000.0)
    n = NX * NY * NZ
    u = np.full(n, 5.0, np.float32)
    v = np.full(n, 1.0, np.float32)
    w = np.zeros(n, np.float32)
    m = NX * NY
    pbl = dict(ustar=np.full(m, 0.3, np.float32), wstar=np.zeros(m, np.float32),
               hmix=np.full(m, 3000.0, np.float32), oli=np.zeros(m, np.float32))
    return lh, (u, v, w), pbl


def test_emission_gradient_matches_the_backward_run(fp, workloads, setup):
    say("adjoint vs backward run: 12 sources, 12 steps, 20k forward / 40k backward particles per release")
    lh, (u, v, w), pbl = setup
    cfg = fp.StepConfig(dt_seconds=DT, x_scale=SCALE, y_scale=SCALE, key=(3, 9), counter=(0, 0, 0, 0))
    # forward + adjoint
    counts = [N_FWD] * len(SOURCES)
    adj = fp.GpuAdjoint(counts, (NX, NY, NZ), n_met_levels=2, n_modes=0, obs_nz=OBS_NZ, obs_dz_m=OBS_DZ,
                        n_steps=N_STEPS, obs_first_step=N_STEPS, checkpoint_every=4)
    adj.upload_met([(u, v, w), (u, v, w)], lh, 6 * 3600.0)
    adj.upload_pbl(pbl["ustar"], pbl["wstar"], pbl["hmix"], pbl["oli"], np.zeros(NX * NY, np.float32))
    q = np.ones(len(SOURCES), np.float32)
    pos = np.array([(x, y, l * OBS_DZ) for x, y, l in SOURCES], np.float32)
    adj.set_receptors([RECEPTOR])
    c_r, _ = adj.forward(cfg, q, pos)
    adj.set_receptors([RECEPTOR], [float(c_r[0]) - 0.5])     # loss (c - obs)^2 with c - obs = 0.5
    adj.forward(cfg, q, pos)
    g, _, _ = adj.backward()                                   # dL/dq_s = 2 * 0.5 * dc/dq_s
    state = adj.download_state()
    adj.close()
    # forward weights per source from the final state (cross-check and variance)
    wf = kernel_weight(state, RECEPTOR)
    wf = wf.reshape(len(SOURCES), N_FWD)
    gf = wf.mean(axis=1)
    np.testing.assert_allclose(g, gf, rtol=1e-4, atol=1e-7)
    # backward run from the receptor point
    p = workloads.make_particles(N_BWD, float(RECEPTOR[0]), float(RECEPTOR[1]), RECEPTOR[2] * OBS_DZ,
                                 np.float32(1.0 / N_BWD))
    with fp.GpuTimestep(N_BWD, (NX, NY, NZ), n_species=1) as ctx:
        ctx.upload_met_bracket(u, v, w, u, v, w, lh)
        ctx.upload_pbl(pbl["ustar"], pbl["wstar"], pbl["hmix"], pbl["oli"])
        ctx.upload_store(p)
        bcfg = fp.StepConfig(dt_seconds=DT, x_scale=SCALE, y_scale=SCALE, key=(11, 5), counter=(0, 0, 0, 0),
                             direction=-1)
        for _ in range(N_STEPS):
            bcfg.counter = ctx.step(bcfg)
        back = ctx.download_particles()
    mid = SOURCES.index((11, 10, 5))
    zf = state["pos_z"].reshape(len(SOURCES), N_FWD)[mid].astype(np.float64)
    zb = back["pos_z"].astype(np.float64)
    rise = 0.5 * ((zf.mean() - 1500.0) + (zb.mean() - 1500.0))
    say(f"  mean z after {N_STEPS} steps from 1500 m: forward {zf.mean():.1f} m, backward {zb.mean():.1f} m")
    assert 20.0 < rise < 100.0
    rows = []
    for s, node in enumerate(SOURCES):
        wb = kernel_weight(back, node)
        b = wb.mean()
        pred = kernel_weight(back, node, 2.0 * rise).mean() - b
        sigma = np.sqrt(wf[s].var() / N_FWD + wb.var() / N_BWD)
        rows.append((node, float(g[s]), float(b), float(sigma), float(pred)))
        say(f"  source {node}: adjoint {g[s]:.5f}  backward {b:.5f}  mc sigma {sigma:.5f}  "
            f"drift prediction {pred:+.5f}")
    by = {node: (gs, b, sigma) for node, gs, b, sigma, _ in rows}
    for node, gs, b, sigma, pred in rows:
        if node[2] != 5 and max(gs, b) >= 0.01:
            # the forward - backward difference is the drift's translation
            # (same sign, size within 70 % of the first-order prediction)
            assert abs((gs - b) - pred) <= 0.7 * abs(pred) + 5.0 * sigma, (node, gs, b, pred, sigma)
    seen = 0
    for (x, y, l), (gs, b, sigma) in by.items():
        if l == 5:  # same level: mirror-symmetric transport
            assert abs(gs - b) <= 5.0 * sigma + 0.02 * max(gs, b), ((x, y, l), gs, b, sigma)
            seen += max(gs, b) > 0.03
    assert seen >= 3
    for x in (10, 11):
        for y in (9, 10, 11):
            lo, hi = by[(x, y, 4)], by[(x, y, 6)]
            if max(lo[0], lo[1], hi[0], hi[1]) < 0.01:
                continue
            # the rising particles: below is seen more forward, above more backward
            assert lo[0] > lo[1] and hi[0] < hi[1], ((x, y), lo, hi)
            gsum, bsum = lo[0] + hi[0], lo[1] + hi[1]
            sig = np.hypot(lo[2], hi[2])
            assert abs(gsum - bsum) <= 5.0 * sig + 0.03 * max(gsum, bsum), ((x, y), gsum, bsum, sig)
