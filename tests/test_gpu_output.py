"""Concentration output intervals on the device (fp_output_sample /
fp_output_finalize) against the oracle: 64-bit interval sums of the
fixed-point grids, the reference's kg conversion (gpu/gridding.rs:210-213)
for one sample, interval means and the cell-volume concentration
(docs/science/concentration-gridding.md:35-41) bit-exact, and the NetCDF
snapshot of the result."""
import sys

import numpy as np
import pytest

from cdf_reader import read as cdf_read
from conftest import met_dict, pbl_dict

pytestmark = pytest.mark.gpu

GEO = dict(xlon0=-10.0, ylat0=35.0, dx=0.25, dy=0.2)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def small(workloads):
    return workloads.small_random(20_000, shape=(48, 40, 24))


def make_ctx(fp, w):
    ctx = fp.GpuTimestep(len(w.particles), w.met.shape, n_species=4, out_shape=w.out_shape)
    u0, v0, w0 = w.met.levels[0]
    u1, v1, w1 = w.met.levels[1]
    ctx.upload_met_bracket(u0, v0, w0, u1, v1, w1, w.met.level_heights)
    ctx.upload_pbl(w.pbl.ustar, w.pbl.wstar, w.pbl.hmix, w.pbl.oli, w.pbl.precip)
    ctx.upload_store(w.particles)
    return ctx


def test_interval_means_and_concentrations_match_oracle(fp, orc, small):
    say("output: 4 sampled steps, interval mean and concentration vs oracle")
    from test_gpu_parity import base_cfg, oracle_params
    nx, ny, nz = small.out_shape
    want = small.particles.copy()
    acc_c = np.zeros(nx * ny * nz, np.uint64)
    acc_m = np.zeros(nx * ny * nz, np.uint64)
    with make_ctx(fp, small) as ctx:
        ctx.configure_grid(small.outheights, small.refine, 0, small.mass_scale)
        ctx.output_reset()
        counter = (11, 0, 0, 0)
        for k in range(4):
            cfg = base_cfg(fp, small, counter=counter, grid_fused=True, dry_mode=1, dry_vd=3e-3)
            ctx.clear_grid()
            counter = ctx.step(cfg)
            ctx.output_sample()
            orc.step(want, met_dict(small.met), pbl_dict(small.pbl), oracle_params(small, cfg))
            oc, om = orc.grid_accumulate(want, small.out_shape, small.outheights, small.refine, 0, small.mass_scale)
            acc_c += oc.reshape(-1).astype(np.uint64)
            acc_m += om.reshape(-1).astype(np.uint64)
        geo = fp.OutputGeometry(**GEO)
        got = ctx.output_finalize(geo, concentration=True)
    assert got["n_samples"] == 4
    vol = orc.output_volumes(GEO["ylat0"], GEO["dx"], GEO["dy"], small.outheights, ny)
    ref = orc.output_finalize(acc_c, acc_m, 4, small.mass_scale, small.out_shape, vol)
    assert np.array_equal(got["count"], ref["count"])
    assert np.array_equal(got["mass_kg"].view(np.uint32), ref["mass_kg"].view(np.uint32))
    assert np.array_equal(got["conc_kg_m3"].view(np.uint32), ref["conc_kg_m3"].view(np.uint32))
    assert int(got["count"].sum()) == 4 * np.count_nonzero(want["flags"] & 1)


def test_one_sample_is_the_reference_download_and_sums_persist(fp, small, tmp_path):
    say("output: one sample vs the downloaded grid; reset; NetCDF of the interval")
    with make_ctx(fp, small) as ctx:
        ctx.configure_grid(small.outheights, small.refine, 0, small.mass_scale)
        ctx.clear_grid()
        ctx.accumulate_grid()
        counts, mass = ctx.download_grid()
        ctx.output_reset()
        ctx.output_sample()
        one = ctx.output_finalize()
        assert np.array_equal(one["count"], counts.reshape(-1))
        want = mass.reshape(-1).astype(np.float32) / np.float32(small.mass_scale)
        assert np.array_equal(one["mass_kg"].view(np.uint32), want.view(np.uint32))
        again = ctx.output_finalize()                       # the sums are kept
        assert np.array_equal(again["mass_kg"], one["mass_kg"])
        ctx.output_sample()                                 # same grid twice: same mean, double counts
        two = ctx.output_finalize()
        assert np.array_equal(two["mass_kg"], one["mass_kg"]) and two["n_samples"] == 2
        assert np.array_equal(two["count"], 2 * one["count"])
        ctx.output_reset()
        with pytest.raises(fp.FpError):
            ctx.output_finalize()                           # no samples after a reset
        ctx.output_sample()
        fin = ctx.output_finalize(fp.OutputGeometry(**GEO), concentration=True)
    path = tmp_path / "interval.nc"
    fp.write_netcdf(str(path), small.out_shape, fin["count"], fin["mass_kg"], 1_706_000_000,
                    conc_kg_m3=fin["conc_kg_m3"])
    r = cdf_read(str(path))
    assert r["version"] == 5
    assert np.array_equal(r["vars"]["concentration_particle_count"]["data"].reshape(-1), fin["count"])
    assert np.array_equal(r["vars"]["concentration_mass_kg"]["data"].reshape(-1), fin["mass_kg"])
    assert np.array_equal(r["vars"]["concentration_kg_m3"]["data"].reshape(-1), fin["conc_kg_m3"])


def test_output_errors(fp, small):
    say("output: errors")
    with make_ctx(fp, small) as ctx:
        with pytest.raises(fp.FpError):
            ctx.output_finalize()                           # nothing sampled
        ctx.configure_grid(None, small.refine, 0, small.mass_scale)   # level-index output heights
        ctx.clear_grid()
        ctx.accumulate_grid()
        ctx.output_sample()
        ctx.output_finalize()                               # counts and mass need no geometry
        with pytest.raises(fp.FpError):
            ctx.output_finalize(concentration=True)         # no geometry
        with pytest.raises(fp.FpError):
            ctx.output_finalize(fp.OutputGeometry(**GEO), concentration=True)  # no metre heights
        ctx.configure_grid(small.outheights, small.refine, 0, small.mass_scale)
        with pytest.raises(fp.FpError):
            ctx.output_finalize(fp.OutputGeometry(xlon0=0.0, ylat0=0.0, dx=0.0, dy=1.0), concentration=True)
    with fp.GpuTimestep(16, (8, 8, 4), n_species=1) as ctx:
        with pytest.raises(fp.FpError):
            ctx.output_sample()                             # no output grid
