"""Concentration output on the host side: the NetCDF snapshot writer
(fp_write_netcdf; write_gridded_output_netcdf, io/netcdf_output.rs:141-229)
checked with scipy's netCDF-3 reader (CDF-2) and a format-grammar reader
(CDF-5, tests/cdf_reader.py), and the oracle's output restatement pinned by
the reference's own values: the download conversion (value as f32) /
mass_scale (gpu/gridding.rs:210-213), dx_meters_at_latitude
(coords/mod.rs:156-158 and its tests :377-398), and the cell volume
V = dx_m dy_m dz_m of docs/science/concentration-gridding.md:35-41."""
import math

import numpy as np
import pytest
import scipy.io

from cdf_reader import read as cdf_read

# io/netcdf_output.rs:248-318: the reference's smoke snapshot
SNAP = dict(shape=(2, 3, 2), time_s=1_706_000_000, species_index=0,
            count=[1, 0, 2, 3, 0, 0, 4, 1, 0, 0, 0, 2],
            mass_kg=[0.1, 0.0, 0.2, 0.3, 0.0, 0.0, 0.4, 0.1, 0.0, 0.0, 0.0, 0.2],
            dry_kg_m2=[0.01, 0.02, 0.03, 0.0, 0.0, 0.1], wet_kg_m2=[0.2, 0.1, 0.0, 0.3, 0.2, 0.0])
GLOBAL_ATTRS = {"title": "FLEXPART-GPU concentration/deposition MVP output",
                "source": "flexpart-gpu (I-03 NetCDF concentration/deposition writer)",
                "conventions": "CF-1.10", "deterministic_fields": "true"}


def write(fp, path, fmt, **kw):
    s = dict(SNAP)
    s.update(kw)
    fp.write_netcdf(str(path), s["shape"], s["count"], s["mass_kg"], s["time_s"], s["species_index"],
                    s["dry_kg_m2"], s["wet_kg_m2"], s.get("conc_kg_m3"), fmt=fmt)


def test_cdf2_snapshot_reads_back_with_scipy(fp, tmp_path):
    p = tmp_path / "snap2.nc"
    write(fp, p, 2)
    with scipy.io.netcdf_file(str(p), "r", mmap=False) as f:
        assert f.version_byte == 2
        assert {k: v for k, v in f.dimensions.items()} == {"time": 1, "x": 2, "y": 3, "z": 2}
        for k, v in GLOBAL_ATTRS.items():
            assert getattr(f, k).decode() == v
        assert f.species_index == 0
        c = f.variables["concentration_particle_count"]
        assert c.dimensions == ("time", "x", "y", "z") and c.data.shape == (1, 2, 3, 2)
        flat = c.data.reshape(-1)
        assert flat[0] == 1 and flat[2] == 2 and flat[11] == 2
        m = f.variables["concentration_mass_kg"].data.reshape(-1)
        assert abs(m[0] - 0.1) < 1e-6 and abs(m[11] - 0.2) < 1e-6
        assert f.variables["concentration_mass_kg"].units == b"kg"
        d = f.variables["dry_deposition_kg_m2"]
        assert d.dimensions == ("time", "x", "y")
        assert abs(d.data.reshape(-1)[0] - 0.01) < 1e-6 and abs(d.data.reshape(-1)[5] - 0.1) < 1e-6
        w = f.variables["wet_deposition_kg_m2"].data.reshape(-1)
        assert np.array_equal(w, np.float32(SNAP["wet_kg_m2"]))
        assert f.variables["time"].data[0] == 1_706_000_000
        assert f.variables["time"].units == b"seconds since 1970-01-01 00:00:00 UTC"
        for name, n, axis in (("x", 2, b"X"), ("y", 3, b"Y"), ("z", 2, b"Z")):
            assert np.array_equal(f.variables[name].data, np.arange(n, dtype=np.float32))
            assert f.variables[name].axis == axis


def test_format_reader_agrees_with_scipy_on_cdf2(fp, tmp_path):
    p = tmp_path / "snap2.nc"
    write(fp, p, 2)
    mine = cdf_read(str(p))
    with scipy.io.netcdf_file(str(p), "r", mmap=False) as f:
        assert mine["dims"] == dict(f.dimensions)
        for name, v in f.variables.items():
            assert np.array_equal(mine["vars"][name]["data"], v.data), name
            assert tuple(mine["vars"][name]["dims"]) == v.dimensions


def test_cdf5_snapshot_has_the_reference_types_and_the_same_data(fp, tmp_path):
    p5, p2 = tmp_path / "snap5.nc", tmp_path / "snap2.nc"
    write(fp, p5, 5)
    write(fp, p2, 2)
    a, b = cdf_read(str(p5)), cdf_read(str(p2))
    assert a["version"] == 5 and a["dims"] == {"time": 1, "x": 2, "y": 3, "z": 2}
    assert {k: v[1] for k, v in a["attrs"].items() if v[0] == 2} == GLOBAL_ATTRS
    assert a["attrs"]["species_index"][0] == 4 and a["attrs"]["species_index"][1][0] == 0
    assert a["vars"]["concentration_particle_count"]["type"] == 9            # NC_UINT
    assert a["vars"]["time"]["type"] == 10                                   # NC_INT64
    assert a["vars"]["time"]["data"][0] == 1_706_000_000
    assert np.array_equal(a["vars"]["concentration_particle_count"]["data"].reshape(-1),
                          np.uint32(SNAP["count"]))
    for name in ("x", "y", "z", "concentration_mass_kg", "dry_deposition_kg_m2", "wet_deposition_kg_m2"):
        assert np.array_equal(a["vars"][name]["data"], b["vars"][name]["data"]), name
        assert a["vars"][name]["attrs"] == b["vars"][name]["attrs"], name
    # data sections: variables laid out in order, each 4-byte aligned, file ends with the last one
    begins = [v["begin"] for v in a["vars"].values()]
    assert begins == sorted(begins) and all(x % 4 == 0 for x in begins)
    last = list(a["vars"].values())[-1]
    assert a["size"] == last["begin"] + last["vsize"]


def test_large_counts_and_volume_concentration_variable(fp, tmp_path):
    p = tmp_path / "big.nc"
    cnt = np.array(SNAP["count"], np.uint32)
    cnt[3] = 3_000_000_000
    conc = np.linspace(0, 1e-9, 12, dtype=np.float32)
    write(fp, p, 5, count=cnt, conc_kg_m3=conc)
    r = cdf_read(str(p))
    assert r["vars"]["concentration_particle_count"]["data"].reshape(-1)[3] == 3_000_000_000
    assert np.array_equal(r["vars"]["concentration_kg_m3"]["data"].reshape(-1), conc)
    assert r["vars"]["concentration_kg_m3"]["attrs"]["units"][1] == "kg m-3"
    with pytest.raises(fp.FpError):
        write(fp, tmp_path / "big2.nc", 2, count=cnt)                        # not representable in CDF-2


def test_writer_rejects_bad_arguments(fp, tmp_path):
    with pytest.raises(fp.FpError):
        write(fp, tmp_path / "a.nc", 3)
    with pytest.raises(fp.FpError):
        write(fp, tmp_path / "no_such_dir" / "a.nc", 5)
    with pytest.raises(fp.FpError):
        fp.write_netcdf(str(tmp_path / "z.nc"), (0, 3, 2), [], [], 0)


def test_dx_meters_at_latitude_reference_values(orc):
    # coords/mod.rs:377-398: ~111195 m per degree at the equator, half at 60N, ~0 at the pole.
    # A 1 x 1 degree cell of 1 m depth has volume dx_m(lat_c) * dy_m with dy_m = dx_m(0).
    dy_m = 6.371e6 * math.pi / 180.0
    eq = float(orc.output_volumes(-0.5, 1.0, 1.0, [1.0], 1)[0, 0]) / dy_m
    assert abs(eq - 111_195.0) < 100.0
    n60 = float(orc.output_volumes(59.5, 1.0, 1.0, [1.0], 1)[0, 0]) / dy_m
    assert abs(n60 / eq - 0.5) < 0.01
    pole = float(orc.output_volumes(89.5, 1.0, 1.0, [1.0], 1)[0, 0]) / dy_m
    assert abs(pole) < 1.0


def test_cell_volumes_match_the_closed_form(orc):
    """V = (dx * R cos(lat_c) PI180) * (dy * R PI180) * dz with the latitude and
    the angle rounded to f32 as the reference's f32 helper has them."""
    R, P = 6.371e6, np.float32(math.pi / 180.0)
    oh = [100.0, 500.0, 1000.0, 3000.0]
    v = orc.output_volumes(-90.0, 0.5, 0.5, oh, 360)
    for iy in (0, 1, 45, 180, 300, 359):
        lat = np.float32(-90.0 + (iy + 0.5) * 0.5)
        ang = float(np.float32(lat * P))
        area = (0.5 * R * math.cos(ang) * float(P)) * (0.5 * R * float(P))
        for k, dz in enumerate(np.diff([0.0] + oh)):
            assert float(v[iy, k]) == pytest.approx(area * dz, rel=1e-6), (iy, k)


def test_finalize_conversion_is_the_reference_download_for_one_sample(orc):
    rng = np.random.default_rng(5)
    n = 4096
    mass = rng.integers(0, 2 ** 32, n, dtype=np.uint64)
    mass[:4] = [0, 1, 2 ** 24 + 1, 2 ** 32 - 1]
    cnt = rng.integers(0, 1000, n, dtype=np.uint64)
    out = orc.output_finalize(cnt, mass, 1, 1e6, (16, 16, 16))
    want = mass.astype(np.uint32).astype(np.float32) / np.float32(1e6)     # (value as f32) / mass_scale
    assert np.array_equal(out["mass_kg"].view(np.uint32), want.view(np.uint32))
    assert np.array_equal(out["count"], cnt.astype(np.uint32))


def test_finalize_means_saturation_and_concentration(orc):
    acc = np.array([0, 10, 7, 2 ** 40, 5 * (2 ** 32 - 1)], np.uint64)
    out = orc.output_finalize(acc, acc, 5, 1e5, (1, 1, 5), vol=np.arange(1, 6, dtype=np.float32))
    assert out["count"].tolist() == [0, 10, 7, 0xFFFFFFFF, 0xFFFFFFFF]
    means = acc.astype(np.float64) / 5.0
    want = means.astype(np.float32) / np.float32(1e5)
    assert np.array_equal(out["mass_kg"], want)
    assert np.array_equal(out["conc_kg_m3"], want / np.arange(1, 6, dtype=np.float32))
