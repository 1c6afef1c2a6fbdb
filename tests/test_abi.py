"""The C ABI library loads and exports every entry point the header declares.

No compute call is made here (no GPU in the CPU suite).
"""
import os
import re

import numpy as np

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "flexpart_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fp_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_drop_in_surface():
    names = declared_functions()
    for required in ("fp_ctx_create", "fp_ctx_destroy", "fp_particles_upload_slots",
                     "fp_particles_download", "fp_met_upload_bracket", "fp_pbl_upload",
                     "fp_step", "fp_grid_accumulate", "fp_grid_download", "fp_compact",
                     "fp_sort_morton", "fp_sync", "fp_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol(fp):
    lib = fp.library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared_functions()) == set(fp.ABI), "ctypes table and header differ"


def test_particle_record_layout_matches_reference_struct(fp):
    # particles/mod.rs:76-168: 96 bytes, field offsets in declaration order
    dt = fp.PARTICLE_DTYPE
    assert dt.itemsize == 96
    offs = {n: dt.fields[n][1] for n in dt.names}
    assert offs["pos_x"] == 0 and offs["cell_x"] == 12 and offs["flags"] == 20
    assert offs["mass"] == 24 and offs["vel_u"] == 40 and offs["turb_u"] == 52
    assert offs["time"] == 64 and offs["cbt"] == 88 and offs["pad0"] == 92


def test_step_params_struct_size_is_stable(fp):
    import ctypes as C
    assert C.sizeof(fp.StepParams) == 132
    assert C.sizeof(fp.CtxDesc) == 64
    assert C.sizeof(fp.GridParams) == 80
    assert C.sizeof(fp.SortParams) == 56


def test_version_string_is_available_without_a_device(fp):
    assert b"gfx950" in fp.library().fp_version()


def test_step_config_serialises_every_field(fp):
    cfg = fp.StepConfig(dt_seconds=900.0, key=(1, 2), counter=(3, 4, 5, 6), n_substeps=2,
                        dry_mode=fp.DRY_UNIFORM, dry_vd=2e-3, grid_fused=True)
    c = cfg.to_c()
    assert c.dt_seconds == 900.0 and list(c.key) == [1, 2] and list(c.counter) == [3, 4, 5, 6]
    assert c.n_substeps == 2 and c.dry_mode == 1 and abs(c.dry_vd - 2e-3) < 1e-9 and c.grid_fused == 1
    assert np.isclose(c.min_height_m, 0.01)


def test_met_grids_beyond_two_to_the_27_points_are_rejected(fp):
    """Contexts, ensembles and adjoint runs address the packed met bracket (32 B
    per point) with 32-bit byte offsets: a grid of 2^27 + a few points is
    refused at creation, before any device call, with the limit in the
    message."""
    import ctypes as C
    lib = fp.library()
    big = dict(met_nx=2049, met_ny=1024, met_nz=64)   # 134,283,264 points > 2^27
    h = C.c_void_p()
    d = fp.CtxDesc(capacity=1, n_species=1, **big)
    assert lib.fp_ctx_create(0, C.byref(d), C.byref(h)) == fp.FP_E_INVALID
    assert b"2^27" in lib.fp_last_error(None)
    caps = (C.c_uint64 * 2)(1, 1)
    e = fp.EnsDesc(n_members=2, n_species=1, capacities=caps, **big)
    assert lib.fp_ens_create(0, C.byref(e), C.byref(h)) == fp.FP_E_INVALID
    assert b"2^27" in lib.fp_ens_last_error(None)
    per = (C.c_uint32 * 1)(4)
    a = fp.AdjDesc(n_sources=1, n_met_levels=2, obs_nz=2, obs_dz_m=100.0, n_steps=2, obs_first_step=1,
                   checkpoint_every=1, particles_per_source=per, **big)
    assert lib.fp_adj_create(0, C.byref(a), C.byref(h)) == fp.FP_E_INVALID
    assert b"2^27" in lib.fp_adj_last_error(None)
