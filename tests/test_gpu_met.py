"""Met ingest on the device: the hybrid sigma-pressure to height transform
(fp_met_hybrid_to_height; io/vertical_transform.rs:106-334) bit-exact
against the oracle, including the sequential f32 mean profile on a full
360 x 181 x 60 grid and the reference's error precedence, and the
asynchronous bracket prefetch (fp_met_prefetch_bracket /
fp_met_commit_prefetch) giving the same particle states as synchronous
uploads."""
import sys

import numpy as np
import pytest

from conftest import assert_particles_equal

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def atmosphere(nx, ny, nz, seed):
    """Hybrid coefficients and a standard-atmosphere-like state with noise."""
    rng = np.random.default_rng(seed)
    s = np.linspace(0.0, 1.0, nz)
    bk = ((1.0 - s) ** 2.5).astype(np.float32)
    ak = (8000.0 * np.sin(np.pi * s) ** 2 + 10.0 * s).astype(np.float32)
    ps = rng.uniform(85_000.0, 104_000.0, (nx, ny)).astype(np.float32)
    p = ak[None, None, :] + bk[None, None, :] * ps[:, :, None]
    t = (288.0 * (p / 101_325.0) ** 0.19 + rng.normal(0, 0.8, (nx, ny, nz))).astype(np.float32)
    t = np.maximum(t, 190.0).astype(np.float32)
    q = np.clip(0.012 * (p / 101_325.0) ** 3 + rng.normal(0, 5e-4, (nx, ny, nz)), -1e-4, None).astype(np.float32)
    return ps, t, q, ak, bk


@pytest.mark.parametrize("shape,moist", [((48, 40, 24), True), ((48, 40, 24), False), ((360, 181, 60), True)])
def test_vertical_transform_matches_oracle_bitwise(fp, orc, shape, moist):
    say(f"hybrid to height: shape={shape} moist={moist}")
    ps, t, q, ak, bk = atmosphere(*shape, seed=sum(shape))
    qq = q if moist else None
    code, _, p_want, z_want, mean_want = orc.hybrid_to_height(ps, t, ak, bk, q=qq)
    assert code == 0
    with fp.GpuTimestep(16, shape, n_species=1) as ctx:
        got = ctx.hybrid_to_height(ps, t, ak, bk, humidity=qq, fields=True)
    assert np.array_equal(got["pressure"].view(np.uint32), p_want.view(np.uint32))
    assert np.array_equal(got["height"].view(np.uint32), z_want.view(np.uint32))
    assert np.array_equal(got["mean_heights"].view(np.uint32), mean_want.view(np.uint32))


def test_vertical_transform_errors_follow_the_reference_order(fp, orc):
    say("hybrid to height: error kinds and locations")
    shape = (16, 12, 10)
    ps, t, q, ak, bk = atmosphere(*shape, seed=5)
    cases = []
    t1 = t.copy(); t1[9, 3, 7] = np.nan; t1[4, 11, 2] = -1.0            # two thermo errors: the earlier column
    cases.append((ps, t1, q))
    ps2 = ps.copy(); ps2[15, 11] = -2.0e5                               # pressure error beats the thermo errors
    cases.append((ps2, t1, q))
    q3 = q.copy(); q3[2, 5, 0] = np.inf                                 # level 0 of a column
    cases.append((ps, t, q3))
    ps4 = ps.copy(); ps4[7, 7] = 1.0                                    # tiny surface pressure: inversion
    cases.append((ps4, t, q))
    with fp.GpuTimestep(16, shape, n_species=1) as ctx:
        for k, (a, b, c) in enumerate(cases):
            code, ijk, *_ = orc.hybrid_to_height(a, b, ak, bk, q=c)
            assert code != 0, k
            with pytest.raises(fp.FpError) as e:
                ctx.hybrid_to_height(a, b, ak, bk, humidity=c)
            assert (e.value.err.code, e.value.err.i, e.value.err.j, e.value.err.k) == (code, *ijk), k
        ctx.hybrid_to_height(ps, t, ak, bk, humidity=q)                  # usable after errors


def test_prefetched_brackets_equal_synchronous_uploads(fp, workloads):
    say("met prefetch: 3 brackets, prefetch overlapping steps vs synchronous uploads")
    w = workloads.small_random(20_000, shape=(48, 40, 24), n_levels=4)
    n = len(w.particles)
    lv = w.met.levels
    lh2 = (w.met.level_heights * np.float32(1.02)).astype(np.float32)
    brackets = [(lv[0], lv[1], w.met.level_heights), (lv[1], lv[2], w.met.level_heights), (lv[2], lv[3], lh2)]
    from test_gpu_parity import base_cfg
    states = []
    for mode in ("sync", "prefetch"):
        with fp.GpuTimestep(n, w.met.shape, n_species=4) as ctx:
            ctx.upload_pbl(w.pbl.ustar, w.pbl.wstar, w.pbl.hmix, w.pbl.oli, w.pbl.precip)
            ctx.upload_store(w.particles)
            a, b, lh = brackets[0]
            ctx.upload_met_bracket(*a, *b, lh)
            counter = (11, 0, 0, 0)
            for nb in range(1, 3):
                a, b, lh = brackets[nb]
                if mode == "prefetch":
                    pinned = [fp.pinned_array(x.shape) for x in (*a, *b)]
                    for dst, src in zip(pinned, (*a, *b)):
                        dst[...] = src
                    ctx.prefetch_bracket(*pinned, lh)
                for k in range(3):
                    counter = ctx.step(base_cfg(fp, w, counter=counter, alpha=k / 3.0))
                if mode == "prefetch":
                    ctx.commit_prefetch()
                else:
                    ctx.upload_met_bracket(*a, *b, lh)
                counter = ctx.step(base_cfg(fp, w, counter=counter, alpha=0.0))
            states.append(ctx.download_particles())
    assert_particles_equal(states[0], states[1], "prefetched vs synchronous brackets")


def test_prefetch_call_order(fp, workloads):
    say("met prefetch: call order and validation")
    w = workloads.small_random(1000, shape=(16, 12, 8))
    a, b = w.met.levels[0], w.met.levels[1]
    with fp.GpuTimestep(1000, w.met.shape, n_species=1) as ctx:
        with pytest.raises(fp.FpError):
            ctx.commit_prefetch()                                        # nothing pending
        bad = w.met.level_heights.copy()
        bad[3] = np.nan
        with pytest.raises(fp.FpError):
            ctx.prefetch_bracket(*a, *b, bad)
        ctx.prefetch_bracket(*a, *b, w.met.level_heights)
        with pytest.raises(fp.FpError):
            ctx.prefetch_bracket(*a, *b, w.met.level_heights)            # one pending at a time
        ctx.commit_prefetch()
        ctx.prefetch_bracket(*b, *a, w.met.level_heights)
        ctx.commit_prefetch()
        ctx.sync()
