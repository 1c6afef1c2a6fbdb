"""Golden vectors (tests/golden/*.npz, written by make_golden.py).

CPU: the oracle reproduces them bit for bit.  GPU: the HIP path does too.
"""
import os

import numpy as np
import pytest

from conftest import ROOT, met_dict, pbl_dict

GOLD = os.path.join(ROOT, "tests", "golden")


def load(name):
    return dict(np.load(os.path.join(GOLD, name)))


def as_particles(fp, words):
    return np.ascontiguousarray(words).view(fp.PARTICLE_DTYPE).reshape(-1)


@pytest.fixture(scope="module")
def pr(workloads):
    return workloads.point_release()


def _params(w, k, counter):
    return dict(dt_seconds=w.dt_seconds, alpha=(k * w.dt_seconds) / w.met.seconds_between_levels,
                x_scale=w.x_scale, y_scale=w.y_scale, key=w.key, counter=counter, dry_mode=1,
                dry_vd=w.dry_vd, dry_href=15.0)


def test_oracle_reproduces_point_release_golden(fp, orc, pr):
    g = load("point_release_64.npz")
    p = as_particles(fp, g["init"]).copy()
    counter = (0, 0, 0, 0)
    for k in range(10):
        counter = tuple(int(v) for v in orc.step(p, met_dict(pr.met), pbl_dict(pr.pbl),
                                                 _params(pr, k, counter)))
        assert list(counter) == g["counters"][k].tolist()
        if k == 0:
            assert np.array_equal(p.view(np.uint32), g["step1"])
    assert np.array_equal(p.view(np.uint32), g["step10"])
    c, m = orc.grid_accumulate(p, pr.out_shape, pr.outheights, pr.refine, 0, 1.0e6)
    assert np.array_equal(c, g["grid_counts"]) and np.array_equal(m, g["grid_mass"])
    assert g["grid_counts"].sum() == np.count_nonzero(p["flags"] & 1)


def test_oracle_reproduces_compaction_and_sort_golden(fp, orc):
    g = load("compaction_sort_512.npz")
    p = as_particles(fp, g["particles"])
    a, out, smap = orc.compact(p)
    assert a == int(g["active"][0])
    assert np.array_equal(smap, g["survivor_map"])
    assert np.array_equal(out.view(np.uint32), g["compacted"])
    s, n2o = orc.sort_morton(p, 10)
    assert np.array_equal(n2o, g["n2o10"]) and np.array_equal(s.view(np.uint32), g["sorted10"])
    s, n2o = orc.sort_morton(p, 3, bounds=(0.0, 16.0, 0.0, 12.0, 0.0, 2500.0))
    assert np.array_equal(n2o, g["n2o3"]) and np.array_equal(s.view(np.uint32), g["sorted3"])


def test_oracle_reproduces_philox_golden(orc):
    g = load("philox_64.npz")
    assert np.array_equal(orc.philox_stream(g["base"], g["key"], 64), g["blocks"])


@pytest.mark.gpu
def test_gpu_reproduces_point_release_golden(fp, pr):
    g = load("point_release_64.npz")
    init = as_particles(fp, g["init"])
    u0, v0, w0 = pr.met.levels[0]
    u1, v1, w1 = pr.met.levels[1]
    with fp.GpuTimestep(len(init), pr.met.shape, out_shape=pr.out_shape) as ctx:
        ctx.upload_met_bracket(u0, v0, w0, u1, v1, w1, pr.met.level_heights)
        ctx.upload_pbl(pr.pbl.ustar, pr.pbl.wstar, pr.pbl.hmix, pr.pbl.oli)
        ctx.upload_store(init)
        ctx.configure_grid(pr.outheights, pr.refine, 0, 1.0e6)
        ctx.clear_grid()
        counter = (0, 0, 0, 0)
        for k in range(10):
            prm = _params(pr, k, counter)
            cfg = fp.StepConfig(**{k2: v for k2, v in prm.items()}, grid_fused=(k == 9))
            counter = ctx.step(cfg)
            assert list(counter) == g["counters"][k].tolist()
            if k == 0:
                assert np.array_equal(ctx.download_particles().view(np.uint32), g["step1"])
        assert np.array_equal(ctx.download_particles().view(np.uint32), g["step10"])
        c, m = ctx.download_grid()
    assert np.array_equal(c, g["grid_counts"]) and np.array_equal(m, g["grid_mass"])


@pytest.mark.gpu
def test_gpu_reproduces_compaction_sort_and_philox_golden(fp):
    g = load("compaction_sort_512.npz")
    p = as_particles(fp, g["particles"])
    with fp.GpuTimestep(len(p), (16, 12, 8)) as ctx:
        ctx.upload_store(p)
        a, smap = ctx.compact()
        assert a == int(g["active"][0]) and np.array_equal(smap, g["survivor_map"])
        assert np.array_equal(ctx.download_particles().view(np.uint32), g["compacted"])
        ctx.upload_store(p)
        ctx.set_dispatch_count(len(p))
        n2o = ctx.sort_morton(10)
        assert np.array_equal(n2o, g["n2o10"])
        assert np.array_equal(ctx.download_particles().view(np.uint32), g["sorted10"])
        ph = load("philox_64.npz")
        assert np.array_equal(ctx.philox_blocks(ph["base"], ph["key"], 64), ph["blocks"])
