"""Differentiable runs (fp_adj_*) on the GPU, through the C ABI.

* The forward run is the production step: the final particle state equals
  the oracle's step sequence on the same initial particles, perturbed met,
  temporal weights and counters, bit for bit, and the receptor values equal
  a host cloud-in-cell sum over the oracle's states.
* The emission gradient is exact: receptor values are linear in the
  emissions, so dL/dq_s = sum_r 2 (c_r - obs_r) c_r(e_s), with c(e_s) from
  forward runs with one unit source.
* Release-position and wind-coefficient gradients agree with central finite
  differences of the forward run (pathwise derivative under common random
  numbers).
"""
import sys

import numpy as np
import pytest

from conftest import pbl_dict

pytestmark = pytest.mark.gpu

COUNTS = [600, 257, 64, 700, 1, 378]
COEFF = [0.3, -0.2, 0.5, 0.1]
OBS_NZ, OBS_DZ = 9, 300.0


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def case(workloads):
    import dataclasses
    w = workloads.small_random(10, shape=(48, 40, 24), n_levels=3)
    # levels 30 min apart: the 6-step runs cross a bracket boundary at step 3
    w.met = dataclasses.replace(w.met, seconds_between_levels=1800.0)
    c = workloads.adjoint_case(sum(COUNTS), len(COUNTS), len(COEFF), met=w.met, pbl=w.pbl, x_scale=w.x_scale)
    c.counts[:] = COUNTS
    c.coeff[:] = COEFF
    return c


@pytest.fixture(scope="module")
def smooth_case(workloads, case):
    """The same sources and met with a horizontally uniform neutral boundary
    layer 3 km deep and releases at 1000-1800 m: no particle meets a PBL-cell
    boundary, a regime change or a reflection within the run, so the loss is
    smooth in the parameters and central differences resolve the pathwise
    derivative."""
    import copy
    c = copy.deepcopy(case)
    n = c.pbl.ustar.size
    c.pbl = workloads.Pbl(c.pbl.shape, np.full(n, 0.3, np.float32), np.zeros(n, np.float32),
                          np.full(n, 3000.0, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32))
    rng = np.random.default_rng(31)
    c.position[:, 2] = rng.uniform(1000.0, 1800.0, len(c.position)).astype(np.float32)
    return c


def make_adj(fp, c, n_steps=6, obs_first=4, ckpt=4, compact=False, emission_only=False):
    adj = fp.GpuAdjoint(c.counts, c.met.shape, n_met_levels=len(c.met.levels), n_modes=len(c.coeff),
                        obs_nz=OBS_NZ, obs_dz_m=OBS_DZ, n_steps=n_steps, obs_first_step=obs_first,
                        checkpoint_every=ckpt, compact=compact, emission_only=emission_only)
    adj.upload_met(c.met.levels, c.met.level_heights, c.met.seconds_between_levels)
    adj.upload_basis(c.basis_u, c.basis_v)
    adj.upload_pbl(c.pbl.ustar, c.pbl.wstar, c.pbl.hmix, c.pbl.oli, c.pbl.precip)
    return adj


def cfg_of(fp, c, **kw):
    cfg = fp.StepConfig(dt_seconds=c.dt_seconds, x_scale=c.x_scale, y_scale=c.y_scale, key=c.key,
                        counter=(5, 0, 0, 0), dry_mode=1, dry_vd=2e-3)
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


def all_nodes(c):
    nx, ny, _ = c.met.shape
    i, j, l = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(OBS_NZ), indexing="ij")
    return np.stack([i.ravel(), j.ravel(), l.ravel()], axis=1)


def split_cell(v, n):
    cl = np.clip(np.float32(v), np.float32(0.0), np.float32(n - 1))
    ci = int(np.floor(cl))
    fr = np.float32(cl - np.float32(ci))
    if ci >= n - 1:
        return n - 1, np.float32(0.0)
    return ci, fr


def initial_particles(fp, c):
    p = np.zeros(int(np.sum(c.counts)), fp.PARTICLE_DTYPE)
    nx, ny, _ = c.met.shape
    o = 0
    for s, n in enumerate(c.counts):
        cx, px = split_cell(c.position[s, 0], nx)
        cy, py = split_cell(c.position[s, 1], ny)
        sl = slice(o, o + int(n))
        p["cell_x"][sl], p["pos_x"][sl] = cx, px
        p["cell_y"][sl], p["pos_y"][sl] = cy, py
        p["pos_z"][sl] = c.position[s, 2]
        p["mass"][sl, 0] = np.float32(c.emission[s]) / np.float32(n)
        p["flags"][sl] = 1
        p["cbt"][sl] = 1
        o += int(n)
    return p


def perturbed_levels(c, coeff):
    nx, ny, nz = c.met.shape
    pu = np.zeros(nx * ny, np.float32)
    pv = np.zeros(nx * ny, np.float32)
    for k, ck in enumerate(np.asarray(coeff, np.float32)):
        pu = (pu + ck * c.basis_u[k]).astype(np.float32)
        pv = (pv + ck * c.basis_v[k]).astype(np.float32)
    out = []
    for (u, v, w) in c.met.levels:
        out.append(((u.reshape(nx * ny, nz) + pu[:, None]).astype(np.float32).reshape(-1),
                    (v.reshape(nx * ny, nz) + pv[:, None]).astype(np.float32).reshape(-1), w))
    return out


def cic_values(c, parts, table, acc):
    """Host restatement of the observation deposit (f32 weights, f64 sums)."""
    nx, ny, _ = c.met.shape
    f32 = np.float32
    act = (parts["flags"] & 1) != 0
    x = (parts["cell_x"].astype(f32) + parts["pos_x"])[act]
    y = (parts["cell_y"].astype(f32) + parts["pos_y"])[act]
    z = parts["pos_z"][act]
    m = parts["mass"][act, 0]
    xc = np.clip(x, f32(0), f32(nx - 1))
    yc = np.clip(y, f32(0), f32(ny - 1))
    zc = (np.clip(z, f32(0), f32((OBS_NZ - 1) * OBS_DZ)) / f32(OBS_DZ)).astype(f32)
    i0 = np.minimum(np.floor(xc).astype(np.int64), nx - 2)
    j0 = np.minimum(np.floor(yc).astype(np.int64), ny - 2)
    l0 = np.minimum(np.floor(zc).astype(np.int64), OBS_NZ - 2)
    fx = (xc - i0.astype(f32)).astype(f32)
    fy = (yc - j0.astype(f32)).astype(f32)
    fz = (zc - l0.astype(f32)).astype(f32)
    for k in range(8):
        wx = fx if k & 1 else (f32(1) - fx)
        wy = fy if k & 2 else (f32(1) - fy)
        wz = fz if k & 4 else (f32(1) - fz)
        w = ((wx * wy).astype(f32) * wz).astype(f32)
        node = ((i0 + (k & 1)) * ny + j0 + ((k >> 1) & 1)) * OBS_NZ + l0 + (k >> 2)
        rec = table[node]
        ok = rec >= 0
        np.add.at(acc, rec[ok], (m[ok] * w[ok]).astype(f32).astype(np.float64))


@pytest.mark.parametrize("ckpt", [4, 1, 8])
def test_forward_is_the_production_step(fp, orc, case, ckpt):
    """Checkpoints every 4 steps, every step (each step writes a new
    checkpoint) and beyond the run (only the last step writes one)."""
    from test_gpu_parity import oracle_params
    c = case
    n_steps, obs_first = 6, 4
    say(f"adjoint forward vs oracle: {len(COUNTS)} sources, {sum(COUNTS)} particles, {n_steps} steps, "
        f"checkpoint every {ckpt}")
    nodes = all_nodes(c)
    with make_adj(fp, c, n_steps, obs_first, ckpt=ckpt) as adj:
        adj.set_receptors(nodes)
        cfg = cfg_of(fp, c)
        vals, loss = adj.forward(cfg, c.emission, c.position, c.coeff)
        got = adj.download_state()
    parts = initial_particles(fp, c)
    lv = perturbed_levels(c, c.coeff)
    T = int(c.met.seconds_between_levels)
    table = np.arange(len(nodes))
    want_vals = np.zeros(len(nodes))
    counter = tuple(cfg.counter)
    for k in range(n_steps):
        t = int(k * c.dt_seconds)
        b = min(t // T, len(lv) - 2)
        met = dict(shape=c.met.shape, u0=lv[b][0], v0=lv[b][1], w0=lv[b][2], u1=lv[b + 1][0],
                   v1=lv[b + 1][1], w1=lv[b + 1][2], level_heights=c.met.level_heights)
        cfg.alpha = fp.interpolation_alpha(b * T, (b + 1) * T, t)
        cfg.counter = counter
        counter = tuple(int(v) for v in orc.step(parts, met, pbl_dict(c.pbl), oracle_params(c, cfg)))
        if k + 1 >= obs_first:
            cic_values(c, parts, table, want_vals)
    assert got.tobytes() == parts.tobytes(), "final state differs from the oracle's step sequence"
    assert np.count_nonzero(want_vals) > 0
    np.testing.assert_allclose(vals, want_vals, rtol=1e-9, atol=1e-12)
    assert loss == pytest.approx(float(np.sum(want_vals ** 2)), rel=1e-9)


def receptors_for(fp, c):
    """Receptors on the lattice nodes the run reaches, observations from
    perturbed-truth parameters."""
    nodes = all_nodes(c)
    with make_adj(fp, c) as adj:
        adj.set_receptors(nodes)
        rng = np.random.default_rng(9)
        q_true = (c.emission * np.exp(0.3 * rng.normal(size=len(COUNTS)))).astype(np.float32)
        r_true = c.position.copy()
        r_true[:, :2] += rng.normal(0.0, 0.3, (len(COUNTS), 2)).astype(np.float32)
        c_true = (c.coeff + rng.normal(0.0, 0.3, len(COEFF))).astype(np.float32)
        truth, _ = adj.forward(cfg_of(fp, c), q_true, r_true, c_true)
        base, _ = adj.forward(cfg_of(fp, c), c.emission, c.position, c.coeff)
    hit = np.nonzero((truth > 0) | (base > 0))[0]
    return nodes[hit], truth[hit]


@pytest.fixture(scope="module")
def receptors(fp, case):
    return receptors_for(fp, case)


@pytest.fixture(scope="module")
def smooth_receptors(fp, smooth_case):
    return receptors_for(fp, smooth_case)


def test_emission_gradient_is_exact(fp, case, receptors):
    c = case
    nodes, obs = receptors
    say(f"adjoint emission gradient: {len(nodes)} receptors")
    with make_adj(fp, c) as adj:
        adj.set_receptors(nodes, obs)
        cfg = cfg_of(fp, c)
        vals, loss = adj.forward(cfg, c.emission, c.position, c.coeff)
        gq, gr, gc = adj.backward()
        want = np.zeros(len(COUNTS))
        for s in range(len(COUNTS)):
            unit = np.zeros(len(COUNTS), np.float32)
            unit[s] = 1.0
            col, _ = adj.forward(cfg, unit, c.position, c.coeff)
            want[s] = np.sum(2.0 * (vals - obs) * col)
    say(f"  loss {loss:.6e}  adjoint {gq}  linear {want}")
    np.testing.assert_allclose(gq, want, rtol=2e-5, atol=1e-6 * np.max(np.abs(want)))


def central_difference(adj, cfg, c, which, idx, h):
    q, r, k = c.emission.copy(), c.position.copy(), c.coeff.copy()
    out = []
    for sgn in (1.0, -1.0):
        qq, rr, kk = q.copy(), r.copy(), k.copy()
        if which == "r":
            rr[idx] = np.float32(rr[idx] + sgn * h)
        else:
            kk[idx] = np.float32(kk[idx] + sgn * h)
        _, loss = adj.forward(cfg, qq, rr, kk)
        out.append(loss)
    hr = (np.float32(r[idx] + h) - np.float32(r[idx] - h)) if which == "r" else \
        (np.float32(k[idx] + h) - np.float32(k[idx] - h))
    return (out[0] - out[1]) / float(hr)


def test_position_and_wind_gradients_match_central_differences(fp, smooth_case, smooth_receptors):
    """Central differences of the f32 forward run (position steps of 1e-2
    cells and 1 m, coefficient steps of 1e-2 m/s) carry noise of about 1e-3
    of each parameter group's largest gradient and, for the positions of
    sources with a few hundred particles, about 1e-2 of the value (they move
    by that much between step sizes); the adjoint must agree within 1e-2 of
    each value plus 2e-3 of its group's scale."""
    c = smooth_case
    nodes, obs = smooth_receptors
    say("adjoint vs central differences: positions and wind coefficients")
    with make_adj(fp, c) as adj:
        adj.set_receptors(nodes, obs)
        cfg = cfg_of(fp, c)
        adj.forward(cfg, c.emission, c.position, c.coeff)
        gq, gr, gc = adj.backward()
        rows = []
        for s in (0, 1, 3, 5):
            for d, h in ((0, 1e-2), (1, 1e-2), (2, 1.0)):
                rows.append(("r", (s, d), gr[s, d], central_difference(adj, cfg, c, "r", (s, d), h)))
        for k in range(len(COEFF)):
            rows.append(("c", k, gc[k], central_difference(adj, cfg, c, "c", k, 1e-2)))
    scale = {w: max(abs(fd) for ww, _, _, fd in rows if ww == w) for w in ("r", "c")}
    bad = []
    for w, i, g, fd in rows:
        tol = 1e-2 * abs(fd) + 2e-3 * scale[w]
        say(f"  {w}{i}: adjoint {g: .6e}  fd {fd: .6e}  |diff| {abs(g - fd):.2e}  tol {tol:.2e}")
        if abs(g - fd) > tol:
            bad.append((w, i, g, fd))
    assert not bad, bad


def test_smooth_gradients_match_binary64_central_differences(fp, smooth_case, smooth_receptors):
    """The adjoint against central differences of the binary64 restatement of
    the same run (oracle/forward64.py, pinned against the float32 oracle):
    rounding noise no longer limits the step, so steps of 1e-4 cells, 1e-3 m
    and 1e-5 m/s resolve the derivative of the smooth configuration, and
    positions, wind coefficients and emissions must agree within 1e-3 of the
    value (entries below 1 % of their group's largest gradient: within 1e-5
    of that scale)."""
    import forward64 as F
    c = smooth_case
    nodes, obs = smooth_receptors
    say("adjoint vs binary64 central differences: positions, wind coefficients, emissions")
    with make_adj(fp, c) as adj:
        adj.set_receptors(nodes, obs)
        cfg = cfg_of(fp, c)
        adj.forward(cfg, c.emission, c.position, c.coeff)
        gq, gr, gc = adj.backward()

    def loss(q, r, k):
        return F.adjoint_forward(c, q, r, k, nodes, obs, 6, 4, OBS_NZ, OBS_DZ, tuple(cfg.counter),
                                 dry_vd=cfg.dry_vd)[1]

    q0, r0, k0 = (c.emission.astype(np.float64), c.position.astype(np.float64), c.coeff.astype(np.float64))

    def fd(which, idx, h):
        out = []
        for sgn in (1.0, -1.0):
            q, r, k = q0.copy(), r0.copy(), k0.copy()
            {"q": q, "r": r, "c": k}[which][idx] += sgn * h
            out.append(loss(q, r, k))
        return (out[0] - out[1]) / (2.0 * h)

    rows = []
    for s in (0, 1, 3, 5):
        for d, h in ((0, 1e-4), (1, 1e-4), (2, 1e-3)):
            rows.append(("r", (s, d), float(gr[s, d]), fd("r", (s, d), h)))
    for k in range(len(COEFF)):
        rows.append(("c", k, float(gc[k]), fd("c", k, 1e-5)))
    for s in range(len(COUNTS)):
        rows.append(("q", s, float(gq[s]), fd("q", s, 1e-4 * float(q0[s]))))
    scale = {w: max(abs(f) for ww, _, _, f in rows if ww == w) for w in ("r", "c", "q")}
    bad = []
    for w, i, g, f in rows:
        tol = 1e-3 * abs(f) if abs(f) >= 1e-2 * scale[w] else 1e-5 * scale[w]
        say(f"  {w}{i}: adjoint {g: .8e}  fd64 {f: .8e}  rel {abs(g - f) / max(abs(f), 1e-300):.2e}")
        if abs(g - f) > tol:
            bad.append((w, i, g, f))
    assert not bad, bad


def test_autograd_function_returns_the_backward_gradients(fp, case, receptors):
    import torch
    c = case
    nodes, obs = receptors
    Misfit = fp.receptor_misfit_function()
    with make_adj(fp, c) as adj:
        adj.set_receptors(nodes, obs)
        q = torch.tensor(c.emission, requires_grad=True)
        r = torch.tensor(c.position, requires_grad=True)
        k = torch.tensor(c.coeff, requires_grad=True)
        loss = Misfit.apply(adj, cfg_of(fp, c), q, r, k)
        (2.0 * loss).backward()
        gq, gr, gc = adj.backward()
    assert np.allclose(q.grad.numpy(), 2.0 * gq) and np.allclose(r.grad.numpy(), 2.0 * gr)
    assert np.allclose(k.grad.numpy(), 2.0 * gc)


def test_call_order_and_arguments_are_checked(fp, case):
    c = case
    with pytest.raises(fp.FpError):
        fp.GpuAdjoint([], c.met.shape)
    with pytest.raises(fp.FpError):
        fp.GpuAdjoint([0, 5], c.met.shape)
    with pytest.raises(fp.FpError):
        fp.GpuAdjoint([5], c.met.shape, obs_nz=1)
    adj = fp.GpuAdjoint(c.counts, c.met.shape, n_met_levels=3, n_modes=len(COEFF), obs_nz=OBS_NZ,
                        n_steps=4, obs_first_step=2, checkpoint_every=3)
    with adj:
        with pytest.raises(fp.FpError) as e:
            adj.forward(cfg_of(fp, c), c.emission, c.position, c.coeff)
        assert e.value.code == fp.FP_E_STATE
        with pytest.raises(fp.FpError) as e:
            adj.backward()
        assert e.value.code == fp.FP_E_STATE
        adj.upload_met(c.met.levels, c.met.level_heights, c.met.seconds_between_levels)
        adj.upload_basis(c.basis_u, c.basis_v)
        adj.upload_pbl(c.pbl.ustar, c.pbl.wstar, c.pbl.hmix, c.pbl.oli)
        with pytest.raises(fp.FpError):
            adj.set_receptors([[c.met.shape[0], 0, 0]])
        with pytest.raises(fp.FpError):
            adj.set_receptors([[1, 2, 3], [1, 2, 3]])
        with pytest.raises(fp.FpError) as e:
            adj.forward(cfg_of(fp, c, dry_mode=2), c.emission, c.position, c.coeff)
        assert e.value.code == fp.FP_E_INVALID
        adj.set_receptors([[1, 2, 3]], [1.0])
        _, loss = adj.forward(cfg_of(fp, c), c.emission, c.position, c.coeff)
        adj.backward()


# ---------------------------------------------------------------- one-step VJP

VJP_CASES = {
    "production": dict(dry_mode=1, dry_vd=2e-3),
    "no_substeps": dict(n_substeps=0),
    "two_substeps_wet_precip": dict(n_substeps=2, wet_mode=3, wet_a=1e-4, wet_b=0.8, wet_fraction=1.0),
    "backward_density_gradient": dict(direction=-1, rho_grad_over_rho=-1.2e-4),
}
# direction scales: x, y [cells], z [m], turb_u, turb_v, turb_w [m/s], mass (relative)
D_SCALE = np.array([1.0, 1.0, 200.0, 1.0, 1.0, 1.0, 1.0])


def state7(parts):
    s = np.zeros((7, len(parts)))
    s[0] = parts["cell_x"].astype(np.float64) + parts["pos_x"].astype(np.float64)
    s[1] = parts["cell_y"].astype(np.float64) + parts["pos_y"].astype(np.float64)
    s[2] = parts["pos_z"]
    s[3], s[4], s[5] = parts["turb_u"], parts["turb_v"], parts["turb_w"]
    s[6] = parts["mass"][:, 0]
    return s


def shifted(parts, delta):
    """parts moved by delta (7, n) in f32, the horizontal shift applied to the
    cell fraction (carrying into the cell index)."""
    p = parts.copy()
    f32 = np.float32
    for cell, pos, dd in (("cell_x", "pos_x", delta[0]), ("cell_y", "pos_y", delta[1])):
        fr = (p[pos] + dd.astype(f32)).astype(f32)
        lo, hi = fr < 0, fr >= 1
        p[cell] = p[cell] - lo.astype(np.int32) + hi.astype(np.int32)
        fr = np.where(lo, fr + f32(1), np.where(hi, fr - f32(1), fr)).astype(f32)
        p[pos] = fr
    p["pos_z"] = (p["pos_z"] + delta[2].astype(f32)).astype(f32)
    for k, name in ((3, "turb_u"), (4, "turb_v"), (5, "turb_w")):
        p[name] = (p[name] + delta[k].astype(f32)).astype(f32)
    p["mass"][:, 0] = (p["mass"][:, 0] + delta[6].astype(f32)).astype(f32)
    return p


@pytest.fixture(scope="module")
def vjp_work(workloads):
    return workloads.small_random(20_000, shape=(48, 40, 24))


@pytest.mark.parametrize("name", sorted(VJP_CASES))
def test_step_vjp_matches_directional_differences(fp, vjp_work, name):
    """<J^T lam, d> == <lam, (F(X + h d) - F(X - h d)) / 2h> per particle."""
    from test_gpu_parity import base_cfg, make_ctx
    w = vjp_work
    say(f"step VJP vs directional differences: case={name} n={len(w.particles)}")
    rng = np.random.default_rng(21)
    n = len(w.particles)
    lam_out = rng.normal(size=(7, n)).astype(np.float32)
    d = rng.normal(size=(7, n)) * D_SCALE[:, None]
    d[6] *= w.particles["mass"][:, 0]
    h = 1e-3
    with make_ctx(fp, w, n_species=1, out=False) as ctx:
        cfg = base_cfg(fp, w, **VJP_CASES[name])
        lam_in = ctx.step_vjp(cfg, lam_out)
        outs, ins = [], []
        for sgn in (1.0, -1.0):
            p = shifted(w.particles, sgn * h * d)
            ins.append(state7(p))
            ctx.upload_store(p)
            ctx.step(cfg)
            outs.append(state7(ctx.download_particles()))
    d_eff = (ins[0] - ins[1]) / (2 * h)
    fd = np.sum(lam_out * (outs[0] - outs[1]) / (2 * h), axis=0)
    vjp = np.sum(lam_in.astype(np.float64) * d_eff, axis=0)
    act = (w.particles["flags"] & 1) != 0
    assert np.array_equal(lam_in[:, ~act], lam_out[:, ~act]), "inactive slots must pass lambda through"
    rel = np.abs(fd - vjp)[act] / (np.abs(fd) + np.abs(vjp) + 1e-6)[act]
    q50, q90, q99 = np.quantile(rel, [0.5, 0.9, 0.99])
    say(f"  relative error quantiles: 50% {q50:.2e}  90% {q90:.2e}  99% {q99:.2e}")
    assert q50 < 2e-3 and q90 < 2e-2, (q50, q90, q99)


def test_step_vjp_matches_binary64_directional_differences(fp, vjp_work):
    """The one-step VJP of the production step (4 sub-steps, dry deposition)
    against central differences of the binary64 restatement of that step
    (oracle/forward64.py), per particle: without float32 rounding in the
    differences, a step of 1e-6 along each direction resolves the derivative.
    Measured on MI355X: relative error 6.9e-8 at the median, 5.9e-7 at the 90th
    and 1.6e-5 at the 99th percentile; asserted: 1e-6, 1e-5 and 1e-3 (the
    float32 differences of the test above resolve only 1e-4 and 1e-3)."""
    import forward64 as F
    from test_gpu_parity import base_cfg, make_ctx
    w = vjp_work
    n = len(w.particles)
    say(f"step VJP vs binary64 directional differences: production step, n={n}")
    rng = np.random.default_rng(21)
    lam_out = rng.normal(size=(7, n)).astype(np.float32)
    d = rng.normal(size=(7, n)) * D_SCALE[:, None]
    d[6] *= w.particles["mass"][:, 0]
    cfg = base_cfg(fp, w, **VJP_CASES["production"])
    with make_ctx(fp, w, n_species=1, out=False) as ctx:
        lam_in = ctx.step_vjp(cfg, lam_out)
    (u0, v0, w0), (u1, v1, w1) = w.met.levels
    met = F.Met64(w.met.shape, u0, v0, w0, u1, v1, w1, w.met.level_heights)
    pbl = {k: np.asarray(getattr(w.pbl, k), np.float64) for k in ("ustar", "wstar", "hmix", "oli")}
    prm = dict(dt=float(cfg.dt_seconds), alpha=float(np.float32(cfg.alpha)), x_scale=float(np.float32(cfg.x_scale)),
               y_scale=float(np.float32(cfg.y_scale)), key=tuple(cfg.key), counter=list(cfg.counter), n_slots=n,
               slots=np.arange(n), dry_vd=float(np.float32(cfg.dry_vd)), dry_href=float(np.float32(cfg.dry_href)))
    p = w.particles
    act = (p["flags"] & 1) != 0
    base = state7(p)
    h = 1e-6
    outs = []
    for sgn in (1.0, -1.0):
        x = base + sgn * h * d
        st = dict(x=x[0], y=x[1], z=x[2], tu=x[3], tv=x[4], tw=x[5], cbt=p["cbt"].astype(np.float64),
                  mass=x[6], active=act.copy())
        F.step(st, met, pbl, prm)
        outs.append(np.stack([st["x"], st["y"], st["z"], st["tu"], st["tv"], st["tw"], st["mass"]]))
    fd = np.sum(lam_out * (outs[0] - outs[1]) / (2 * h), axis=0)
    vjp = np.sum(lam_in.astype(np.float64) * d, axis=0)
    rel = np.abs(fd - vjp)[act] / (np.abs(fd) + np.abs(vjp) + 1e-6)[act]
    q50, q90, q99 = np.quantile(rel, [0.5, 0.9, 0.99])
    say(f"  relative error quantiles: 50% {q50:.2e}  90% {q90:.2e}  99% {q99:.2e}")
    assert q50 < 1e-6 and q90 < 1e-5 and q99 < 1e-3, (q50, q90, q99)


def test_step_vjp_wind_adjoint_matches_directional_difference(fp, vjp_work, workloads):
    """d/de <lam, F(X; wind + e dW)> == <wind adjoint, dW> for a random
    horizontal-wind offset dW added to every level and time level.  The
    boundary layer is uniform and neutral and the particles sit well inside it,
    so no particle crosses a PBL-cell boundary or reflects within the
    difference and the sum over particles is smooth in e."""
    import copy
    from test_gpu_parity import base_cfg
    w = copy.deepcopy(vjp_work)
    n2 = w.pbl.ustar.size
    w.pbl = workloads.Pbl(w.pbl.shape, np.full(n2, 0.3, np.float32), np.zeros(n2, np.float32),
                          np.full(n2, 3000.0, np.float32), np.zeros(n2, np.float32), None)
    w.particles["pos_z"] = np.random.default_rng(23).uniform(1000.0, 1800.0, len(w.particles))
    say("step VJP wind adjoint vs directional difference")
    nx, ny, nz = w.met.shape
    rng = np.random.default_rng(22)
    # unit weights on the horizontal positions (the quantities a horizontal
    # wind offset moves directly) and positive offsets: every particle adds to
    # the derivative with the same sign, so f32 differences resolve it
    lam_out = np.zeros((7, len(w.particles)), np.float32)
    lam_out[:2] = 1.0
    du = rng.uniform(0.5, 1.5, nx * ny).astype(np.float32)
    dv = rng.uniform(0.5, 1.5, nx * ny).astype(np.float32)
    h = 1e-2
    cfg = base_cfg(fp, w)
    vals = []
    with fp.GpuTimestep(len(w.particles), w.met.shape, n_species=1) as ctx:
        ctx.upload_pbl(w.pbl.ustar, w.pbl.wstar, w.pbl.hmix, w.pbl.oli)
        (u0, v0, w0), (u1, v1, w1) = w.met.levels[0], w.met.levels[1]
        ctx.upload_met_bracket(u0, v0, w0, u1, v1, w1, w.met.level_heights)
        ctx.upload_store(w.particles)
        _, wu, wv = ctx.step_vjp(cfg, lam_out, want_wind=True)
        for sgn in (1.0, -1.0):
            off_u = np.repeat(np.float32(sgn * h) * du, nz)
            off_v = np.repeat(np.float32(sgn * h) * dv, nz)
            ctx.upload_met_bracket(u0 + off_u, v0 + off_v, w0, u1 + off_u, v1 + off_v, w1,
                                   w.met.level_heights)
            ctx.upload_store(w.particles)
            ctx.step(cfg)
            vals.append(np.sum(lam_out * state7(ctx.download_particles())))
    fd = (vals[0] - vals[1]) / (2 * h)
    adj = float(np.sum(wu * du) + np.sum(wv * dv))
    say(f"  adjoint {adj:.8e}  directional difference {fd:.8e}")
    assert abs(adj - fd) <= 2e-3 * abs(fd)
    # the same derivative from central differences of the binary64 restatement
    # (oracle/forward64.py), whose step of 1e-4 m/s has no float32 noise
    import forward64 as F
    pbl = {k: np.asarray(getattr(w.pbl, k), np.float64) for k in ("ustar", "wstar", "hmix", "oli")}
    p = w.particles
    act = (p["flags"] & 1) != 0
    base = state7(p)
    prm = dict(dt=float(cfg.dt_seconds), alpha=float(np.float32(cfg.alpha)), x_scale=float(np.float32(cfg.x_scale)),
               y_scale=float(np.float32(cfg.y_scale)), key=tuple(cfg.key), counter=list(cfg.counter),
               n_slots=len(p), slots=np.arange(len(p)))
    h64, vals64 = 1e-4, []
    for sgn in (1.0, -1.0):
        off_u = np.repeat(sgn * h64 * du.astype(np.float64), nz)
        off_v = np.repeat(sgn * h64 * dv.astype(np.float64), nz)
        met = F.Met64(w.met.shape, u0 + off_u, v0 + off_v, w0, u1 + off_u, v1 + off_v, w1, w.met.level_heights)
        st = dict(x=base[0].copy(), y=base[1].copy(), z=base[2].copy(), tu=base[3].copy(), tv=base[4].copy(),
                  tw=base[5].copy(), cbt=p["cbt"].astype(np.float64), mass=base[6].copy(), active=act.copy())
        F.step(st, met, pbl, prm)
        vals64.append(np.sum(lam_out[0] * st["x"]) + np.sum(lam_out[1] * st["y"]))
    fd64 = (vals64[0] - vals64[1]) / (2 * h64)
    say(f"  binary64 directional difference {fd64:.8e}  relative error {abs(adj - fd64) / abs(fd64):.2e}")
    assert abs(adj - fd64) <= 1e-6 * abs(fd64)  # measured 3.5e-8


# ---------------------------------------------------------------- compaction

def lossy_cfg(fp, c):
    """Strong dry and precip-driven wet deposition with the loss rule: the
    particles of a source die once they keep less than 80 % of the smallest
    initial particle mass."""
    m0 = float(np.min(c.emission / c.counts.astype(np.float32)))
    return cfg_of(fp, c, dry_vd=5e-2, wet_mode=3, wet_a=1e-3, wet_b=0.8, wet_fraction=1.0,
                  loss_mode=1, loss_min_mass=0.8 * m0)


def test_compacting_forward_is_the_oracle_sequence(fp, orc, case):
    """Step, observe, compact every step: states, active counts and receptor
    values equal the oracle's step + compaction sequence."""
    from test_gpu_parity import oracle_params
    c = case
    n_steps, obs_first = 6, 4
    nodes = all_nodes(c)
    cfg = lossy_cfg(fp, c)
    say(f"compacting adjoint forward vs oracle: {sum(COUNTS)} particles, {n_steps} steps")
    with make_adj(fp, c, n_steps, obs_first, compact=True) as adj:
        adj.set_receptors(nodes)
        vals, _ = adj.forward(cfg, c.emission, c.position, c.coeff)
        counts = adj.active_counts()
        got = adj.download_state(int(counts[-1]))
    parts = initial_particles(fp, c)
    lv = perturbed_levels(c, c.coeff)
    T = int(c.met.seconds_between_levels)
    table = np.arange(len(nodes))
    want_vals = np.zeros(len(nodes))
    counter = tuple(cfg.counter)
    n_cur = len(parts)
    want_counts = [n_cur]
    for k in range(n_steps):
        t = int(k * c.dt_seconds)
        b = min(t // T, len(lv) - 2)
        met = dict(shape=c.met.shape, u0=lv[b][0], v0=lv[b][1], w0=lv[b][2], u1=lv[b + 1][0],
                   v1=lv[b + 1][1], w1=lv[b + 1][2], level_heights=c.met.level_heights)
        cfg.alpha = fp.interpolation_alpha(b * T, (b + 1) * T, t)
        cfg.counter = counter
        live = parts[:n_cur].copy()
        counter = tuple(int(v) for v in orc.step(live, met, pbl_dict(c.pbl), oracle_params(c, cfg)))
        if k + 1 >= obs_first:
            cic_values(c, live, table, want_vals)
        n_cur, out, _ = orc.compact(live)
        parts[: len(out)] = out
        want_counts.append(n_cur)
    assert list(counts) == want_counts
    assert want_counts[-1] < want_counts[0], "the case must lose particles"
    assert got.tobytes() == parts[: want_counts[-1]].tobytes()
    np.testing.assert_allclose(vals, want_vals, rtol=1e-9, atol=1e-12)


def test_identity_compaction_leaves_the_gradients_unchanged(fp, case, receptors):
    """Without losses every compaction is the identity: the saved-map replay
    and the adjoint's un-compaction give the gradients of the run without
    compaction; the emission-only backward gives its emission gradient."""
    c = case
    nodes, obs = receptors
    say("identity compaction: gradients with and without the map path")
    grads = []
    for compact, emission_only in ((False, False), (True, False), (True, True)):
        with make_adj(fp, c, compact=compact, emission_only=emission_only) as adj:
            adj.set_receptors(nodes, obs)
            adj.forward(cfg_of(fp, c), c.emission, c.position, c.coeff)
            assert list(adj.active_counts()) == [sum(COUNTS)] * 7
            grads.append(adj.backward())
    (q0, r0, k0), (q1, r1, k1), (q2, r2, k2) = grads
    np.testing.assert_allclose(q1, q0, rtol=1e-6)
    np.testing.assert_allclose(r1, r0, rtol=1e-5, atol=1e-6 * np.max(np.abs(r0)))
    np.testing.assert_allclose(k1, k0, rtol=1e-5)
    np.testing.assert_allclose(q2, q0, rtol=1e-5)
    assert not np.any(r2) and not np.any(k2)


def test_emission_only_backward_matches_the_full_adjoint_through_losses(fp, case):
    """With particles lost and compacted every step, the emission-only
    backward (survival factors through the saved survivor maps) and the full
    adjoint give the same emission gradient."""
    c = case
    cfg = lossy_cfg(fp, c)
    nodes = all_nodes(c)
    with make_adj(fp, c, compact=True) as adj:
        adj.set_receptors(nodes)
        base, _ = adj.forward(cfg, c.emission, c.position, c.coeff)
    hit = np.nonzero(base > 0)[0]
    obs = base[hit] * np.random.default_rng(41).uniform(0.5, 1.5, len(hit))
    out = []
    for emission_only in (False, True):
        with make_adj(fp, c, compact=True, emission_only=emission_only) as adj:
            adj.set_receptors(nodes[hit], obs)
            adj.forward(cfg, c.emission, c.position, c.coeff)
            counts = adj.active_counts()
            out.append(adj.backward())
    say(f"  active counts {list(counts)}; emission gradients {out[0][0]} vs {out[1][0]}")
    assert counts[-1] < counts[0]
    np.testing.assert_allclose(out[1][0], out[0][0], rtol=1e-5, atol=1e-7 * np.max(np.abs(out[0][0])))
    assert np.all(np.isfinite(out[0][1])) and np.all(np.isfinite(out[0][2]))


def test_compacting_backward_matches_central_differences(fp, case):
    """The backward recomputes each segment by stepping only the survivors
    of every step, each at its own slot, through the saved maps.  With
    deposition off and a loss threshold between the sources' particle masses,
    whole sources die in the first step (the later survivors move down) and
    the loss is exactly quadratic in the emissions, so the central difference
    at relative step 1e-3 resolves the gradient; a recomputed state that
    differed from the forward's would change it."""
    c = case
    masses = np.sort(np.unique((c.emission / c.counts.astype(np.float32)).astype(np.float32)))
    assert len(masses) >= 4
    thr = float(np.sqrt(masses[1] * masses[2]))
    dead = (c.emission / c.counts.astype(np.float32)) < thr
    first_dead = int(np.argmax(dead))
    assert dead.sum() == 2 and np.any(~dead[first_dead:]), "later survivors must move"
    cfg = cfg_of(fp, c, dry_mode=0, loss_mode=1, loss_min_mass=thr)
    nodes = all_nodes(c)
    with make_adj(fp, c, compact=True) as adj:
        adj.set_receptors(nodes)
        base, _ = adj.forward(cfg, c.emission, c.position, c.coeff)
    hit = np.nonzero(base > 0)[0]
    obs = base[hit] * np.random.default_rng(43).uniform(0.5, 1.5, len(hit))
    say(f"compacting backward vs central differences: {len(hit)} receptors, sources lost {np.nonzero(dead)[0]}")
    for emission_only in (True, False):
        with make_adj(fp, c, compact=True, emission_only=emission_only) as adj:
            adj.set_receptors(nodes[hit], obs)
            adj.forward(cfg, c.emission, c.position, c.coeff)
            counts = adj.active_counts()
            gq, _, _ = adj.backward()
            fd = np.zeros(len(COUNTS))
            for s in range(len(COUNTS)):
                h = np.float32(1e-3) * c.emission[s]
                out = []
                for sgn in (1.0, -1.0):
                    q = c.emission.copy()
                    q[s] = np.float32(q[s] + sgn * h)
                    out.append(adj.forward(cfg, q, c.position, c.coeff)[1])
                fd[s] = (out[0] - out[1]) / float(np.float32(c.emission[s] + h) - np.float32(c.emission[s] - h))
        say(f"  emission_only={emission_only}: counts {list(counts)} adjoint {gq} fd {fd}")
        assert counts[1] == sum(COUNTS) - int(np.sum(c.counts[dead])) == counts[-1]
        assert np.all(gq[dead] == 0.0) and np.all(np.abs(fd[dead]) <= 1e-6 * np.max(np.abs(fd)))
        np.testing.assert_allclose(gq, fd, rtol=1e-3, atol=1e-4 * np.max(np.abs(fd)))


@pytest.mark.parametrize("compact", [False, True])
def test_gradients_do_not_depend_on_the_checkpoint_interval(fp, case, receptors, compact):
    """The backward recomputes every state between checkpoints exactly, so the
    gradients are those of any other checkpoint interval (up to the order of
    the f64 atomic sums)."""
    c = case
    nodes, obs = receptors
    cfg = lossy_cfg(fp, c) if compact else cfg_of(fp, c)
    say(f"gradients vs checkpoint interval (compact={compact})")
    grads = []
    for ckpt in (1, 3, 4, 8):
        with make_adj(fp, c, ckpt=ckpt, compact=compact) as adj:
            adj.set_receptors(nodes, obs)
            adj.forward(cfg, c.emission, c.position, c.coeff)
            grads.append(adj.backward())
    for g in grads[1:]:
        for a, b in zip(g, grads[0]):
            np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12 * max(1.0, float(np.max(np.abs(b)))))


def test_compacting_24h_stress_run_matches_the_oracle_sequence(fp, orc, workloads):
    """The deposition stress configuration at 200k particles on the 1 deg
    met: 144 steps with dry deposition, precipitation scavenging in patches
    and the loss rule, compacted after every step through the gathered steps
    (records never moved until the last compaction).  Active counts, the
    final compacted records and receptor sums equal the oracle's step +
    compaction sequence."""
    import os
    from test_gpu_parity import oracle_params
    c = workloads.stress_case(200_000, 512, deg=1.0)
    n_steps, obs_first = 144, 133
    m0 = float(np.median(c.emission / c.counts.astype(np.float32)))
    cfg = fp.StepConfig(dt_seconds=c.dt_seconds, x_scale=c.x_scale, y_scale=c.y_scale, key=c.key,
                        dry_mode=1, dry_vd=1e-2, wet_mode=3, wet_a=1e-4, wet_b=0.8, wet_fraction=1.0,
                        loss_mode=1, loss_min_mass=0.3 * m0)
    nx, ny, _ = c.met.shape
    # receptors on the lattice nodes around the first 300 sources, three levels
    src = np.floor(c.position[:300, :2]).astype(np.int64)
    nodes = np.unique(np.array([(i + di, j + dj, l) for i, j in src for di in (0, 1) for dj in (0, 1)
                                for l in (0, 1, 2) if i + di < nx and j + dj < ny]), axis=0)
    say(f"compacting stress run vs oracle: {int(c.counts.sum())} particles, {n_steps} steps")
    adj = fp.GpuAdjoint(c.counts, c.met.shape, n_met_levels=2, n_modes=0, obs_nz=11, obs_dz_m=500.0,
                        n_steps=n_steps, obs_first_step=obs_first, checkpoint_every=8, compact=True,
                        emission_only=True)
    with adj:
        adj.upload_met(c.met.levels, c.met.level_heights, c.met.seconds_between_levels)
        adj.upload_pbl(c.pbl.ustar, c.pbl.wstar, c.pbl.hmix, c.pbl.oli, c.pbl.precip)
        adj.set_receptors(nodes)
        vals, _ = adj.forward(cfg, c.emission, c.position)
        counts = adj.active_counts()
        got = adj.download_state(int(counts[-1]))
    parts = initial_particles(fp, c)
    (u0, v0, w0), (u1, v1, w1) = c.met.levels
    met = dict(shape=c.met.shape, u0=u0, v0=v0, w0=w0, u1=u1, v1=v1, w1=w1, level_heights=c.met.level_heights)
    T = int(c.met.seconds_between_levels)
    threads = max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))))
    table = np.full(nx * ny * 11, -1, np.int64)
    table[(nodes[:, 0] * ny + nodes[:, 1]) * 11 + nodes[:, 2]] = np.arange(len(nodes))
    want_vals = np.zeros(len(nodes))
    counter = tuple(cfg.counter)
    n_cur = len(parts)
    want_counts = [n_cur]
    for k in range(n_steps):
        t = int(k * c.dt_seconds)
        cfg.alpha = fp.interpolation_alpha(0, T, t)
        cfg.counter = counter
        live = parts[:n_cur].copy()
        p = oracle_params(c, cfg)
        p["n_threads"] = threads
        counter = tuple(int(v) for v in orc.step(live, met, pbl_dict(c.pbl), p))
        if k + 1 >= obs_first:
            cic_values_lattice(c, live, table, want_vals, 11, 500.0)
        n_cur, out, _ = orc.compact(live)
        parts[: len(out)] = out
        want_counts.append(n_cur)
    say(f"  active {want_counts[0]} -> {want_counts[-1]}")
    assert list(counts) == want_counts
    assert want_counts[-1] < 0.7 * want_counts[0], "the run must lose particles"
    assert got.tobytes() == parts[: want_counts[-1]].tobytes()
    assert np.count_nonzero(want_vals) > 0
    np.testing.assert_allclose(vals, want_vals, rtol=1e-9, atol=1e-12 * max(1.0, float(np.max(want_vals))))


def cic_values_lattice(c, parts, table, acc, obs_nz, obs_dz):
    """cic_values for an observation lattice of obs_nz levels obs_dz apart."""
    nx, ny, _ = c.met.shape
    f32 = np.float32
    act = (parts["flags"] & 1) != 0
    x = (parts["cell_x"].astype(f32) + parts["pos_x"])[act]
    y = (parts["cell_y"].astype(f32) + parts["pos_y"])[act]
    z = parts["pos_z"][act]
    m = parts["mass"][act, 0]
    xc = np.clip(x, f32(0), f32(nx - 1))
    yc = np.clip(y, f32(0), f32(ny - 1))
    zc = (np.clip(z, f32(0), f32((obs_nz - 1) * obs_dz)) / f32(obs_dz)).astype(f32)
    i0 = np.minimum(np.floor(xc).astype(np.int64), nx - 2)
    j0 = np.minimum(np.floor(yc).astype(np.int64), ny - 2)
    l0 = np.minimum(np.floor(zc).astype(np.int64), obs_nz - 2)
    fx = (xc - i0.astype(f32)).astype(f32)
    fy = (yc - j0.astype(f32)).astype(f32)
    fz = (zc - l0.astype(f32)).astype(f32)
    for k in range(8):
        wx = fx if k & 1 else (f32(1) - fx)
        wy = fy if k & 2 else (f32(1) - fy)
        wz = fz if k & 4 else (f32(1) - fz)
        w = ((wx * wy).astype(f32) * wz).astype(f32)
        node = ((i0 + (k & 1)) * ny + j0 + ((k >> 1) & 1)) * obs_nz + l0 + (k >> 2)
        rec = table[node]
        ok = rec >= 0
        np.add.at(acc, rec[ok], (m[ok] * w[ok]).astype(f32).astype(np.float64))
