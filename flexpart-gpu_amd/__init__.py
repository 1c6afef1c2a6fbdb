"""flexpart-gpu_amd: MI355X (gfx950) per-particle timestep behind a C ABI.

This module is the Python host side of ``include/flexpart_hip.h``: a ctypes
binding of ``lib/libflexpart_hip.so`` plus ``GpuTimestep``, a thin object that
mirrors how the reference time loop drives its GPU backend
(simulation/timeloop.rs:915-1316 over the ``crate::gpu`` API): particle
uploads and downloads, met-bracket and PBL uploads, the fused step, the
concentration grid, compaction and the Morton re-sort.

There is no CPU fallback: importing works anywhere, but every call that needs
the library raises ``LibraryMissing`` when it is not built and ``FpError``
when the HIP runtime reports a failure.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libflexpart_hip.so")

FP_OK = 0
FP_E_INVALID = -1
FP_E_DEVICE = -2
FP_E_NOMEM = -3
FP_E_STATE = -4
# largest n fp_survivor_map_static accepts (its static scratch: one count per 2048 slots)
SURVIVOR_MAP_STATIC_MAX = 134_217_728

FLAG_ACTIVE = 1
MAX_SPECIES = 4
MAX_OUTPUT_LEVELS = 16

DRY_OFF, DRY_UNIFORM, DRY_PER_SLOT = 0, 1, 2
WET_OFF, WET_UNIFORM, WET_PER_SLOT, WET_PRECIP_FIELD = 0, 1, 2, 3
LOSS_NONE, LOSS_MIN_MASS = 0, 1
KERNEL_STEP, KERNEL_GRID, KERNEL_COMPACT, KERNEL_SORT, KERNEL_PBL, KERNEL_OUTPUT = 0, 1, 2, 3, 4, 5

# The reference's 96-byte Particle (particles/mod.rs:76-168) == fp_particle.
PARTICLE_DTYPE = np.dtype(
    [
        ("pos_x", "<f4"), ("pos_y", "<f4"), ("pos_z", "<f4"),
        ("cell_x", "<i4"), ("cell_y", "<i4"), ("flags", "<u4"),
        ("mass", "<f4", (MAX_SPECIES,)),
        ("vel_u", "<f4"), ("vel_v", "<f4"), ("vel_w", "<f4"),
        ("turb_u", "<f4"), ("turb_v", "<f4"), ("turb_w", "<f4"),
        ("time", "<i4"), ("timestep", "<i4"), ("time_mem", "<i4"), ("time_split", "<i4"),
        ("release_point", "<i4"), ("class_id", "<i4"), ("cbt", "<i4"), ("pad0", "<u4"),
    ]
)
assert PARTICLE_DTYPE.itemsize == 96

SOA_FIELDS = ("pos_x", "pos_y", "pos_z", "cell_x", "cell_y", "flags", "mass0", "mass1",
              "mass2", "mass3", "vel_u", "vel_v", "vel_w", "turb_u", "turb_v", "turb_w",
              "time", "timestep", "time_mem", "time_split", "release_point", "class_id",
              "cbt", "pad0")


class LibraryMissing(RuntimeError):
    """libflexpart_hip.so is not built (run ``python flexpart-gpu_amd/build.py``)."""


class FpError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class CtxDesc(C.Structure):
    _fields_ = [("capacity", C.c_uint64), ("met_nx", C.c_uint32), ("met_ny", C.c_uint32),
                ("met_nz", C.c_uint32), ("n_species", C.c_uint32), ("out_nx", C.c_uint32),
                ("out_ny", C.c_uint32), ("out_nz", C.c_uint32), ("reserved", C.c_uint32 * 6)]


class StepParams(C.Structure):
    _fields_ = [
        ("dt_seconds", C.c_float), ("alpha", C.c_float),
        ("x_scale", C.c_float), ("y_scale", C.c_float), ("z_scale", C.c_float),
        ("direction", C.c_int32), ("key", C.c_uint32 * 2), ("counter", C.c_uint32 * 4),
        ("n_substeps", C.c_uint32), ("min_height_m", C.c_float), ("rho_grad_over_rho", C.c_float),
        ("dry_mode", C.c_uint32), ("dry_vd", C.c_float), ("dry_href", C.c_float),
        ("wet_mode", C.c_uint32), ("wet_lambda", C.c_float), ("wet_fraction", C.c_float),
        ("wet_a", C.c_float), ("wet_b", C.c_float),
        ("loss_mode", C.c_uint32), ("loss_min_mass", C.c_float),
        ("grid_fused", C.c_uint32), ("reserved", C.c_uint32 * 7),
    ]


class GridParams(C.Structure):
    _fields_ = [("refine_x", C.c_uint32), ("refine_y", C.c_uint32), ("species", C.c_uint32),
                ("mass_scale", C.c_float), ("outheights", C.c_float * MAX_OUTPUT_LEVELS)]


class SortParams(C.Structure):
    _fields_ = [("bits_per_axis", C.c_uint32), ("explicit_bounds", C.c_uint32),
                ("bounds", C.c_double * 6)]


class AutotuneParams(C.Structure):
    _fields_ = [("warmup_steps", C.c_uint32), ("measured_steps", C.c_uint32),
                ("sort_bits", C.c_uint32), ("n_sort_intervals", C.c_uint32),
                ("sort_intervals", C.c_uint32 * 8), ("sort_trial_steps", C.c_uint32),
                ("cache_path", C.c_char_p)]


class AutotuneResult(C.Structure):
    _fields_ = [("block_size", C.c_uint32), ("particles_per_thread", C.c_uint32),
                ("sort_interval", C.c_uint32), ("from_cache", C.c_uint32),
                ("step_ms", C.c_double), ("step_ms_with_sort", C.c_double)]


class EnsDesc(C.Structure):
    _fields_ = [("n_members", C.c_uint32), ("met_nx", C.c_uint32), ("met_ny", C.c_uint32),
                ("met_nz", C.c_uint32), ("n_species", C.c_uint32), ("out_nx", C.c_uint32),
                ("out_ny", C.c_uint32), ("out_nz", C.c_uint32),
                ("capacities", C.POINTER(C.c_uint64))]


class GridDomain(C.Structure):
    _fields_ = [("xlon0", C.c_double), ("ylat0", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
                ("nx", C.c_uint32), ("ny", C.c_uint32)]


class Release(C.Structure):
    _fields_ = [("lon_min", C.c_double), ("lon_max", C.c_double), ("lat_min", C.c_double),
                ("lat_max", C.c_double), ("z_min", C.c_double), ("z_max", C.c_double),
                ("mass_kg", C.c_double), ("particle_count", C.c_uint64),
                ("start_s", C.c_int64), ("end_s", C.c_int64),
                ("release_point", C.c_int32), ("class_id", C.c_int32)]


class AdjDesc(C.Structure):
    _fields_ = [("n_sources", C.c_uint32), ("met_nx", C.c_uint32), ("met_ny", C.c_uint32),
                ("met_nz", C.c_uint32), ("n_met_levels", C.c_uint32), ("n_modes", C.c_uint32),
                ("obs_nz", C.c_uint32), ("obs_dz_m", C.c_float), ("n_steps", C.c_uint32),
                ("obs_first_step", C.c_uint32), ("checkpoint_every", C.c_uint32),
                ("flags", C.c_uint32), ("particles_per_source", C.POINTER(C.c_uint32))]


ADJ_COMPACT, ADJ_EMISSION_ONLY = 1, 2


class PblOptions(C.Structure):
    """fp_pbl_options: PblComputationOptions (io/pbl_params.rs:36-71); the
    defaults are the reference's."""
    _fields_ = [("roughness_length_m", C.c_float), ("wind_reference_height_m", C.c_float),
                ("heat_flux_neutral_threshold_w_m2", C.c_float), ("fallback_mixing_height_m", C.c_float),
                ("hmix_min_m", C.c_float), ("hmix_max_m", C.c_float)]

    def __init__(self, **kw):
        d = dict(roughness_length_m=0.1, wind_reference_height_m=10.0, heat_flux_neutral_threshold_w_m2=1.0,
                 fallback_mixing_height_m=800.0, hmix_min_m=100.0, hmix_max_m=4500.0)
        d.update(kw)
        super().__init__(**d)


class VtError(C.Structure):
    _fields_ = [("code", C.c_uint32), ("i", C.c_uint32), ("j", C.c_uint32), ("k", C.c_uint32)]


VT_OK, VT_NONPOSITIVE_PRESSURE, VT_NONDECREASING_PRESSURE, VT_INVALID_THERMODYNAMICS, VT_NONMONOTONIC_HEIGHT = range(5)


class OutputGeometry(C.Structure):
    """fp_output_geometry: south-west corner and cell size of the output grid [deg]."""
    _fields_ = [("xlon0", C.c_double), ("ylat0", C.c_double), ("dx", C.c_double), ("dy", C.c_double)]


class NetcdfSnapshot(C.Structure):
    _fields_ = [("time_s", C.c_int64), ("species_index", C.c_int32), ("nx", C.c_uint32), ("ny", C.c_uint32),
                ("nz", C.c_uint32), ("count", C.c_void_p), ("mass_kg", C.c_void_p), ("dry_kg_m2", C.c_void_p),
                ("wet_kg_m2", C.c_void_p), ("conc_kg_m3", C.c_void_p)]


# the ten surface fields of fp_surface_upload, in FP_SURF_* order
SURFACE_FIELDS = ("surface_pressure_pa", "temperature_2m_k", "u10_ms", "v10_ms", "surface_stress_n_m2",
                  "sensible_heat_flux_w_m2", "solar_radiation_w_m2", "mixing_height_m", "friction_velocity_ms",
                  "inv_obukhov_length_per_m")


_f = C.POINTER(C.c_float)
_d = C.POINTER(C.c_double)
_u32 = C.POINTER(C.c_uint32)
_u64 = C.POINTER(C.c_uint64)
_vp = C.c_void_p

# name -> (restype, argtypes); every symbol declared in include/flexpart_hip.h
ABI = {
    "fp_version": (C.c_char_p, []),
    "fp_last_error": (C.c_char_p, [_vp]),
    "fp_ctx_create": (C.c_int, [C.c_int, C.POINTER(CtxDesc), C.POINTER(_vp)]),
    "fp_ctx_destroy": (C.c_int, [_vp]),
    "fp_sync": (C.c_int, [_vp]),
    "fp_particles_upload_slots": (C.c_int, [_vp, _u32, _vp, C.c_uint64]),
    "fp_particles_upload": (C.c_int, [_vp, _vp, C.c_uint64]),
    "fp_particles_download": (C.c_int, [_vp, _vp, C.c_uint64]),
    "fp_particles_set_dispatch_count": (C.c_int, [_vp, C.c_uint64]),
    "fp_particles_dispatch_count": (C.c_uint64, [_vp]),
    "fp_particles_soa_pointers": (C.c_int, [_vp, C.POINTER(_vp), C.c_uint32]),
    "fp_met_upload_bracket": (C.c_int, [_vp, _f, _f, _f, _f, _f, _f, _f]),
    "fp_met_upload_bracket_dev": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _f]),
    "fp_pbl_upload": (C.c_int, [_vp, _f, _f, _f, _f, _f]),
    "fp_forcing_upload": (C.c_int, [_vp, _f, _f, _f]),
    "fp_surface_upload": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), C.c_int64]),
    "fp_surface_upload_dev": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), C.c_int64]),
    "fp_pbl_from_surface": (C.c_int, [_vp, C.POINTER(PblOptions), C.c_int64, C.c_int, _vp]),
    "fp_pbl_download": (C.c_int, [_vp, _f, _f, _f, _f]),
    "fp_host_alloc": (C.c_void_p, [C.c_uint64]),
    "fp_host_free": (C.c_int, [_vp]),
    "fp_met_prefetch_bracket": (C.c_int, [_vp, _f, _f, _f, _f, _f, _f, _f]),
    "fp_met_commit_prefetch": (C.c_int, [_vp]),
    "fp_met_hybrid_to_height": (C.c_int, [_vp, _f, _f, _f, _f, _f, _f, _f, _f, C.POINTER(VtError)]),
    "fp_output_reset": (C.c_int, [_vp]),
    "fp_output_sample": (C.c_int, [_vp]),
    "fp_output_finalize": (C.c_int, [_vp, C.POINTER(OutputGeometry), _u32, _f, _f, _u32]),
    "fp_write_netcdf": (C.c_int, [C.c_char_p, C.POINTER(NetcdfSnapshot), C.c_int]),
    "fp_step": (C.c_int, [_vp, C.POINTER(StepParams), _u32]),
    "fp_step_vjp": (C.c_int, [_vp, C.POINTER(StepParams), _f, _f, _d, _d]),
    "fp_grid_clear": (C.c_int, [_vp]),
    "fp_grid_configure": (C.c_int, [_vp, C.POINTER(GridParams)]),
    "fp_grid_accumulate": (C.c_int, [_vp]),
    "fp_grid_download": (C.c_int, [_vp, _u32, _u32]),
    "fp_compact": (C.c_int, [_vp, C.POINTER(C.c_uint64), _u32]),
    "fp_survivor_map_static": (C.c_int, [C.c_int, _vp, C.c_uint64, _vp, _vp, _vp]),
    "fp_sort_morton": (C.c_int, [_vp, C.POINTER(SortParams), _u32]),
    "fp_philox_blocks": (C.c_int, [_vp, _u32, _u32, C.c_uint64, _u32]),
    "fp_math_probe": (C.c_int, [_vp, C.c_int, _f, _f, _f, C.c_uint64]),
    "fp_math_sweep": (C.c_int, [_vp, C.c_int, C.c_uint32, C.c_uint32, _u64, _u32]),
    "fp_profile_enable": (C.c_int, [_vp, C.c_int]),
    "fp_profile_read": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "fp_profile_reset": (C.c_int, [_vp]),
    "fp_set_launch": (C.c_int, [_vp, C.c_uint32, C.c_uint32]),
    "fp_autotune": (C.c_int, [_vp, C.POINTER(StepParams), C.POINTER(AutotuneParams),
                              C.POINTER(AutotuneResult)]),
    "fp_release_inject": (C.c_int, [_vp, C.POINTER(GridDomain), C.POINTER(Release), C.c_uint32, _u64,
                                    C.c_int64, _u32, C.c_uint64, C.c_uint64, _u64]),
    "fp_interpolation_alpha": (C.c_int, [C.c_int64, C.c_int64, C.c_int64, C.c_int,
                                         C.POINTER(C.c_float)]),
    "fp_adj_create": (C.c_int, [C.c_int, C.POINTER(AdjDesc), C.POINTER(_vp)]),
    "fp_adj_destroy": (C.c_int, [_vp]),
    "fp_adj_last_error": (C.c_char_p, [_vp]),
    "fp_adj_met_upload": (C.c_int, [_vp, _f, _f, _f, _f, C.c_double]),
    "fp_adj_basis_upload": (C.c_int, [_vp, _f, _f]),
    "fp_adj_pbl_upload": (C.c_int, [_vp, _f, _f, _f, _f, _f]),
    "fp_adj_receptors": (C.c_int, [_vp, C.c_uint32, _u32, _d]),
    "fp_adj_forward": (C.c_int, [_vp, C.POINTER(StepParams), _f, _f, _f, _d, _d]),
    "fp_adj_backward": (C.c_int, [_vp, _d, _d, _d]),
    "fp_adj_download_state": (C.c_int, [_vp, _vp, C.c_uint64]),
    "fp_adj_active_counts": (C.c_int, [_vp, _u64]),
    "fp_adj_last_ms": (C.c_int, [_vp, _d, _d]),
    "fp_ens_create": (C.c_int, [C.c_int, C.POINTER(EnsDesc), C.POINTER(_vp)]),
    "fp_ens_destroy": (C.c_int, [_vp]),
    "fp_ens_last_error": (C.c_char_p, [_vp]),
    "fp_ens_sync": (C.c_int, [_vp]),
    "fp_ens_upload_member": (C.c_int, [_vp, C.c_uint32, _vp, C.c_uint64]),
    "fp_ens_download_member": (C.c_int, [_vp, C.c_uint32, _vp, C.c_uint64]),
    "fp_ens_met_upload_member": (C.c_int, [_vp, C.c_uint32, _f, _f, _f, _f, _f, _f, _f]),
    "fp_ens_met_upload_member_dev": (C.c_int, [_vp, C.c_uint32, _vp, _vp, _vp, _vp, _vp, _vp, _f]),
    "fp_ens_pbl_upload": (C.c_int, [_vp, _f, _f, _f, _f, _f]),
    "fp_ens_grid_configure": (C.c_int, [_vp, C.POINTER(GridParams)]),
    "fp_ens_grid_clear": (C.c_int, [_vp]),
    "fp_ens_grid_download_member": (C.c_int, [_vp, C.c_uint32, _u32, _u32]),
    "fp_ens_configure": (C.c_int, [_vp, C.POINTER(StepParams), _u32, _u32, _u64, C.c_double,
                                   C.c_double]),
    "fp_ens_step": (C.c_int, [_vp, C.c_uint32, C.c_int, _u32]),
    "fp_ens_set_launch": (C.c_int, [_vp, C.c_uint32, C.c_uint32]),
    "fp_ens_last_step_ms": (C.c_int, [_vp, C.POINTER(C.c_double)]),
    "fp_ens_sort_morton": (C.c_int, [_vp, C.POINTER(SortParams)]),
    "fp_ens_autotune": (C.c_int, [_vp, C.POINTER(AutotuneParams), C.POINTER(AutotuneResult)]),
}

_LIB: C.CDLL | None = None


def library() -> C.CDLL:
    """Load libflexpart_hip.so (raises LibraryMissing when it is not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(f"{LIB_PATH} not found; run python flexpart-gpu_amd/build.py")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in ABI.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


def _fp(a):
    return None if a is None else a.ctypes.data_as(_f)


def _f32(a, n=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(-1))
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


@dataclass
class StepConfig:
    """Per-step parameters (fp_step_params)."""
    dt_seconds: float = 600.0
    alpha: float = 0.0
    x_scale: float = 1.0
    y_scale: float = 1.0
    z_scale: float = 1.0
    direction: int = 1
    key: tuple = (0xDECAFBAD, 0x12345678)
    counter: tuple = (0, 0, 0, 0)
    n_substeps: int = 4
    min_height_m: float = 0.01
    rho_grad_over_rho: float = 0.0
    dry_mode: int = DRY_OFF
    dry_vd: float = 0.0
    dry_href: float = 15.0
    wet_mode: int = WET_OFF
    wet_lambda: float = 0.0
    wet_fraction: float = 0.0
    wet_a: float = 0.0
    wet_b: float = 0.0
    loss_mode: int = LOSS_NONE
    loss_min_mass: float = 0.0
    grid_fused: bool = False

    def to_c(self) -> StepParams:
        p = StepParams()
        for name, _ in StepParams._fields_:
            if name == "reserved":
                continue
            v = getattr(self, name)
            if name in ("key", "counter"):
                arr = getattr(p, name)
                for i, x in enumerate(v):
                    arr[i] = int(x) & 0xFFFFFFFF
            elif name == "grid_fused":
                p.grid_fused = 1 if v else 0
            else:
                setattr(p, name, v)
        return p


class GpuTimestep:
    """One device context: particle store, met bracket, PBL, output grid.

    Method names follow the reference's ``crate::gpu`` surface
    (ParticleBuffers, DualWindBuffers, PblBuffers, the encode_* passes) so
    that a caller reads like the reference time loop.
    """

    def __init__(self, capacity: int, met_shape, n_species: int = 4, out_shape=None,
                 device: int = 0):
        lib = library()
        nx, ny, nz = met_shape
        ox, oy, oz = out_shape if out_shape is not None else (0, 0, 0)
        d = CtxDesc(capacity, nx, ny, nz, n_species, ox, oy, oz)
        h = _vp()
        rc = lib.fp_ctx_create(device, C.byref(d), C.byref(h))
        if rc != FP_OK:
            raise FpError(rc, lib.fp_last_error(None).decode())
        self._lib = lib
        self._h = h
        self.capacity = int(capacity)
        self.met_shape = (nx, ny, nz)
        self.out_shape = (ox, oy, oz)
        self.n_species = n_species

    # -- plumbing
    def _chk(self, rc):
        if rc != FP_OK:
            raise FpError(rc, self._lib.fp_last_error(self._h).decode())

    def close(self):
        if self._h:
            self._lib.fp_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        self._chk(self._lib.fp_sync(self._h))

    # -- particles (ParticleBuffers)
    def upload_store(self, particles: np.ndarray):
        p = np.ascontiguousarray(particles, dtype=PARTICLE_DTYPE)
        self._chk(self._lib.fp_particles_upload(self._h, p.ctypes.data, len(p)))

    def upload_particle_slots(self, slots, particles: np.ndarray):
        s = np.ascontiguousarray(np.asarray(slots, dtype=np.uint32))
        p = np.ascontiguousarray(particles, dtype=PARTICLE_DTYPE)
        if len(s) != len(p):
            raise ValueError("slots and particles differ in length")
        self._chk(self._lib.fp_particles_upload_slots(self._h, s.ctypes.data_as(_u32),
                                                      p.ctypes.data, len(p)))

    def download_particles(self, n=None) -> np.ndarray:
        n = self.capacity if n is None else int(n)
        out = np.zeros(n, PARTICLE_DTYPE)
        self._chk(self._lib.fp_particles_download(self._h, out.ctypes.data, n))
        return out

    def set_dispatch_count(self, n: int):
        self._chk(self._lib.fp_particles_set_dispatch_count(self._h, int(n)))

    def particle_count(self) -> int:
        return int(self._lib.fp_particles_dispatch_count(self._h))

    def soa_pointers(self) -> dict:
        ptrs = (_vp * len(SOA_FIELDS))()
        self._chk(self._lib.fp_particles_soa_pointers(self._h, ptrs, len(SOA_FIELDS)))
        return dict(zip(SOA_FIELDS, [int(p or 0) for p in ptrs]))

    # -- met / PBL / forcing
    def upload_met_bracket(self, u0, v0, w0, u1, v1, w1, level_heights):
        n = int(np.prod(self.met_shape))
        arrs = [_f32(a, n) for a in (u0, v0, w0, u1, v1, w1)]
        lh = _f32(level_heights, self.met_shape[2])
        self._chk(self._lib.fp_met_upload_bracket(self._h, *[_fp(a) for a in arrs], _fp(lh)))

    def upload_met_bracket_dev(self, ptrs, level_heights):
        lh = _f32(level_heights, self.met_shape[2])
        self._chk(self._lib.fp_met_upload_bracket_dev(self._h, *[_vp(int(p)) for p in ptrs],
                                                      _fp(lh)))

    def upload_pbl(self, ustar, wstar, hmix, oli, precip=None):
        n = self.met_shape[0] * self.met_shape[1]
        arrs = [_f32(a, n) for a in (ustar, wstar, hmix, oli)]
        pr = None if precip is None else _f32(precip, n)
        self._chk(self._lib.fp_pbl_upload(self._h, *[_fp(a) for a in arrs], _fp(pr)))

    def upload_surface(self, level: int, fields, time_s: int):
        """Surface fields of bracket end `level` (0: t0, 1: t1) valid at
        time_s (fp_surface_upload).  fields: mapping with the SURFACE_FIELDS
        names, or a sequence in that order, of met_nx*met_ny arrays; torch
        CUDA tensors are copied device to device."""
        seq = [fields[k] for k in SURFACE_FIELDS] if hasattr(fields, "keys") else list(fields)
        if len(seq) != len(SURFACE_FIELDS):
            raise ValueError(f"{len(SURFACE_FIELDS)} surface fields required")
        n = self.met_shape[0] * self.met_shape[1]
        ptrs = (_vp * len(seq))()
        if all(hasattr(a, "data_ptr") and getattr(a, "is_cuda", False) for a in seq):
            for k, a in enumerate(seq):
                if a.numel() != n or not a.is_contiguous() or str(a.dtype) != "torch.float32":
                    raise ValueError("surface fields must be contiguous float32 tensors of met_nx*met_ny")
                ptrs[k] = a.data_ptr()
            self._chk(self._lib.fp_surface_upload_dev(self._h, level, ptrs, int(time_s)))
            return
        keep = [_f32(a, n) for a in seq]
        for k, a in enumerate(keep):
            ptrs[k] = a.ctypes.data
        self._chk(self._lib.fp_surface_upload(self._h, level, ptrs, int(time_s)))

    def pbl_from_surface(self, options: "PblOptions | None" = None, now_s: int = 0, strict: bool = False,
                         passthrough=None):
        """PBL state from the resident surface bracket at now_s
        (fp_pbl_from_surface).  passthrough: optional contiguous float32 CUDA
        tensor of 3*met_nx*met_ny that receives sanitized sshf, ssr, stress."""
        o = PblOptions() if options is None else options
        pp = None
        if passthrough is not None:
            n = self.met_shape[0] * self.met_shape[1]
            if passthrough.numel() != 3 * n or not passthrough.is_contiguous():
                raise ValueError("passthrough must be a contiguous tensor of 3*met_nx*met_ny floats")
            pp = passthrough.data_ptr()
        self._chk(self._lib.fp_pbl_from_surface(self._h, C.byref(o), int(now_s), 1 if strict else 0, pp))

    def download_pbl(self) -> dict:
        """Current ustar, wstar, hmix, oli (fp_pbl_download)."""
        n = self.met_shape[0] * self.met_shape[1]
        out = {k: np.empty(n, np.float32) for k in ("ustar", "wstar", "hmix", "oli")}
        self._chk(self._lib.fp_pbl_download(self._h, *[_fp(out[k]) for k in ("ustar", "wstar", "hmix", "oli")]))
        return out

    # -- met ingest
    def prefetch_bracket(self, u0, v0, w0, u1, v1, w1, level_heights):
        """Start copying the next bracket on the copy stream (fp_met_prefetch_bracket).
        The arrays are kept referenced until commit_prefetch; pinned_array()
        memory makes the copy overlap the steps."""
        n = int(np.prod(self.met_shape))
        arrs = [_f32(a, n) for a in (u0, v0, w0, u1, v1, w1)]
        lh = _f32(level_heights, self.met_shape[2])
        self._chk(self._lib.fp_met_prefetch_bracket(self._h, *[_fp(a) for a in arrs], _fp(lh)))
        self._prefetch_keep = arrs

    def commit_prefetch(self):
        """Switch to the prefetched bracket (fp_met_commit_prefetch)."""
        self._chk(self._lib.fp_met_commit_prefetch(self._h))
        self._prefetch_keep = None

    def hybrid_to_height(self, ps, temperature, ak_pa, bk, humidity=None, fields: bool = False) -> dict:
        """transform_hybrid_sigma_to_height on the device (fp_met_hybrid_to_height).
        Returns mean_heights (the level heights) and, with fields=True, pressure
        and height per column; raises FpError (err attribute: VtError) on a
        transform error."""
        nx, ny, nz = self.met_shape
        p = _f32(ps, nx * ny)
        tt = _f32(temperature, nx * ny * nz)
        q = None if humidity is None else _f32(humidity, nx * ny * nz)
        a, b = _f32(ak_pa, nz), _f32(bk, nz)
        mean = np.empty(nz, np.float32)
        pres = np.empty(nx * ny * nz, np.float32) if fields else None
        hgt = np.empty(nx * ny * nz, np.float32) if fields else None
        err = VtError()
        rc = self._lib.fp_met_hybrid_to_height(self._h, _fp(p), _fp(tt), _fp(q), _fp(a), _fp(b), _fp(mean),
                                               _fp(pres), _fp(hgt), C.byref(err))
        if rc != FP_OK:
            e = FpError(rc, self._lib.fp_last_error(self._h).decode())
            e.err = err
            raise e
        out = {"mean_heights": mean}
        if fields:
            out["pressure"] = pres.reshape(nx, ny, nz)
            out["height"] = hgt.reshape(nx, ny, nz)
        return out

    # -- concentration output intervals
    def output_reset(self):
        self._chk(self._lib.fp_output_reset(self._h))

    def output_sample(self):
        """Add the current concentration grid to the interval sums (fp_output_sample)."""
        self._chk(self._lib.fp_output_sample(self._h))

    def output_finalize(self, geometry: "OutputGeometry | None" = None, concentration: bool = False) -> dict:
        """Interval fields (fp_output_finalize): count (u32), mass_kg (mean),
        and conc_kg_m3 when concentration=True (needs geometry)."""
        n = int(np.prod(self.out_shape))
        count = np.empty(n, np.uint32)
        mass = np.empty(n, np.float32)
        conc = np.empty(n, np.float32) if concentration else None
        ns = C.c_uint32()
        self._chk(self._lib.fp_output_finalize(self._h, None if geometry is None else C.byref(geometry),
                                               count.ctypes.data_as(_u32), _fp(mass), _fp(conc), C.byref(ns)))
        out = {"count": count, "mass_kg": mass, "n_samples": int(ns.value)}
        if conc is not None:
            out["conc_kg_m3"] = conc
        return out

    def upload_forcing(self, dry_vd=None, wet_lambda=None, wet_fraction=None):
        conv = [None if a is None else _f32(a, self.capacity) for a in (dry_vd, wet_lambda, wet_fraction)]
        self._chk(self._lib.fp_forcing_upload(self._h, *[_fp(a) for a in conv]))

    # -- the step
    def step(self, cfg: StepConfig) -> tuple:
        p = cfg.to_c()
        nxt = (C.c_uint32 * 4)()
        self._chk(self._lib.fp_step(self._h, C.byref(p), nxt))
        return tuple(int(v) for v in nxt)

    def step_vjp(self, cfg: StepConfig, lam_out, want_wind: bool = False):
        """J^T lam_out of one step at the current particles (fp_step_vjp).
        lam_out: (7, dispatch) for x, y, z, turb_u, turb_v, turb_w, mass[0].
        Returns lam_in (7, dispatch) and, if asked, the (u, v) wind adjoints
        per met column."""
        n = self.particle_count()
        lo = np.ascontiguousarray(np.asarray(lam_out, np.float32).reshape(7, n))
        li = np.zeros((7, n), np.float32)
        nxy = self.met_shape[0] * self.met_shape[1]
        wu = np.zeros(nxy, np.float64) if want_wind else None
        wv = np.zeros(nxy, np.float64) if want_wind else None
        self._chk(self._lib.fp_step_vjp(self._h, C.byref(cfg.to_c()), _fp(lo), _fp(li),
                                        None if wu is None else wu.ctypes.data_as(_d),
                                        None if wv is None else wv.ctypes.data_as(_d)))
        return (li, wu, wv) if want_wind else li

    # -- source-term injection (ReleaseManager::inject_for_time)
    def inject_releases(self, grid: dict, releases: list, released, now_s: int, slots=None,
                        first_slot: int = 0) -> int:
        """Emit the particles the releases' windows owe by now_s straight into
        the store (fp_release_inject).  grid: GridDomain fields; releases:
        dicts of Release fields; released: uint64 array of emitted counts,
        advanced in place.  Returns the number of particles written."""
        g = GridDomain(**grid)
        arr = (Release * max(len(releases), 1))()
        for i, r in enumerate(releases):
            arr[i] = Release(**r)
        rel = np.ascontiguousarray(np.asarray(released, np.uint64))
        sl = None if slots is None else np.ascontiguousarray(np.asarray(slots, np.uint32))
        n = C.c_uint64()
        rc = self._lib.fp_release_inject(self._h, C.byref(g), arr, len(releases), rel.ctypes.data_as(_u64),
                                         int(now_s), None if sl is None else sl.ctypes.data_as(_u32),
                                         0 if sl is None else len(sl), int(first_slot), C.byref(n))
        if isinstance(released, np.ndarray):
            released[:] = rel
        self.last_injected = int(n.value)
        self._chk(rc)
        return int(n.value)

    # -- concentration grid
    def configure_grid(self, outheights=None, refine=(1, 1), species=0, mass_scale=1.0e6):
        g = _grid_params(outheights, refine, species, mass_scale)
        self._chk(self._lib.fp_grid_configure(self._h, C.byref(g)))

    def clear_grid(self):
        self._chk(self._lib.fp_grid_clear(self._h))

    def accumulate_grid(self):
        self._chk(self._lib.fp_grid_accumulate(self._h))

    def download_grid(self):
        n = int(np.prod(self.out_shape))
        counts = np.zeros(n, np.uint32)
        mass = np.zeros(n, np.uint32)
        self._chk(self._lib.fp_grid_download(self._h, counts.ctypes.data_as(_u32),
                                             mass.ctypes.data_as(_u32)))
        return counts, mass

    # -- compaction / sort
    def compact(self, want_map: bool = True):
        n = self.particle_count()
        smap = np.zeros(max(n, 1), np.uint32) if want_map else None
        a = C.c_uint64()
        self._chk(self._lib.fp_compact(self._h, C.byref(a),
                                       None if smap is None else smap.ctypes.data_as(_u32)))
        return int(a.value), (None if smap is None else smap[: a.value].copy())

    def sort_morton(self, bits_per_axis: int = 10, bounds=None, want_map: bool = True):
        sp = SortParams()
        sp.bits_per_axis = bits_per_axis
        if bounds is not None:
            sp.explicit_bounds = 1
            for i, v in enumerate(bounds):
                sp.bounds[i] = float(v)
        n = self.particle_count()
        m = np.zeros(max(n, 1), np.uint32) if want_map else None
        self._chk(self._lib.fp_sort_morton(self._h, C.byref(sp),
                                           None if m is None else m.ctypes.data_as(_u32)))
        return None if m is None else m[:n].copy()

    # -- RNG KAT path
    def philox_blocks(self, counter, key, n: int) -> np.ndarray:
        c = np.ascontiguousarray(np.asarray(counter, np.uint32))
        k = np.ascontiguousarray(np.asarray(key, np.uint32))
        out = np.zeros((int(n), 4), np.uint32)
        self._chk(self._lib.fp_philox_blocks(self._h, c.ctypes.data_as(_u32),
                                             k.ctypes.data_as(_u32), int(n),
                                             out.ctypes.data_as(_u32)))
        return out

    # -- device arithmetic probe
    PROBES = {"exp": 0, "log": 1, "sin": 2, "cos": 3, "pow": 4, "sqrt": 5, "div": 6, "log_unit": 7,
              "level": 8}

    def math_probe(self, fn: str, x, y=None) -> np.ndarray:
        xs = _f32(x)
        ys = _f32(np.zeros_like(xs) if y is None else y, xs.size)
        out = np.zeros_like(xs)
        self._chk(self._lib.fp_math_probe(self._h, self.PROBES[fn], _fp(xs), _fp(ys), _fp(out),
                                          xs.size))
        return out

    SWEEPS = {"sqrt_step": 0}

    def math_sweep(self, fn: str, first_bits: int, last_bits: int) -> tuple:
        """Exhaustive device check of a step primitive over float bit patterns
        [first_bits, last_bits] (fp_math_sweep): (failures, lowest failing
        bit pattern or 0xffffffff)."""
        fails = np.zeros(1, np.uint64)
        first = np.zeros(1, np.uint32)
        self._chk(self._lib.fp_math_sweep(self._h, self.SWEEPS[fn], int(first_bits), int(last_bits),
                                          fails.ctypes.data_as(_u64), first.ctypes.data_as(_u32)))
        return int(fails[0]), int(first[0])

    # -- timing / tuning
    def profile(self, enable: bool = True):
        self._chk(self._lib.fp_profile_enable(self._h, 1 if enable else 0))

    def profile_read(self, kernel: int = KERNEL_STEP):
        ms = C.c_double()
        n = C.c_uint64()
        self._chk(self._lib.fp_profile_read(self._h, kernel, C.byref(ms), C.byref(n)))
        return float(ms.value), int(n.value)

    def profile_reset(self):
        self._chk(self._lib.fp_profile_reset(self._h))

    def set_launch(self, block_size: int = 256, particles_per_thread: int = 1):
        self._chk(self._lib.fp_set_launch(self._h, block_size, particles_per_thread))

    def autotune(self, cfg: StepConfig, sort_intervals=(0, 6, 12, 36), warmup_steps: int = 2,
                 measured_steps: int = 7, sort_trial_steps: int = 36, sort_bits: int = 10,
                 cache_path: str | None = None) -> dict:
        """Pick block size, particles/thread and re-sort interval (fp_autotune)."""
        p = AutotuneParams()
        p.warmup_steps, p.measured_steps = warmup_steps, measured_steps
        p.sort_bits, p.sort_trial_steps = sort_bits, sort_trial_steps
        p.n_sort_intervals = len(sort_intervals)
        for i, s in enumerate(sort_intervals):
            p.sort_intervals[i] = s
        p.cache_path = cache_path.encode() if cache_path else None
        r = AutotuneResult()
        self._chk(self._lib.fp_autotune(self._h, C.byref(cfg.to_c()), C.byref(p), C.byref(r)))
        return {"block_size": r.block_size, "particles_per_thread": r.particles_per_thread,
                "sort_interval": r.sort_interval, "from_cache": bool(r.from_cache),
                "step_ms": r.step_ms, "step_ms_with_sort": r.step_ms_with_sort}


def pinned_array(shape, dtype=np.float32) -> np.ndarray:
    """A numpy array in pinned host memory from fp_host_alloc (freed when the
    array is garbage-collected)."""
    import weakref
    lib = library()
    dt = np.dtype(dtype)
    n = int(np.prod(shape))
    nbytes = max(n * dt.itemsize, 1)
    ptr = lib.fp_host_alloc(nbytes)
    if not ptr:
        raise MemoryError(f"fp_host_alloc({nbytes}) failed")
    buf = (C.c_char * nbytes).from_address(ptr)
    arr = np.frombuffer(buf, dtype=dt, count=n).reshape(shape)
    weakref.finalize(buf, lib.fp_host_free, C.c_void_p(ptr))
    return arr


def write_netcdf(path: str, shape, count, mass_kg, time_s: int, species_index: int = 0, dry_kg_m2=None,
                 wet_kg_m2=None, conc_kg_m3=None, fmt: int = 5):
    """write_gridded_output_netcdf (io/netcdf_output.rs:141-229) through
    fp_write_netcdf: fmt 5 = CDF-5 (the reference's types), 2 = CDF-2."""
    lib = library()
    nx, ny, nz = (int(v) for v in shape)
    cnt = np.ascontiguousarray(np.asarray(count, np.uint32).reshape(-1))
    arrs = [_f32(mass_kg, nx * ny * nz)]
    if cnt.size != nx * ny * nz:
        raise ValueError("count must have nx*ny*nz values")
    opt = [None if a is None else _f32(a, n) for a, n in ((dry_kg_m2, nx * ny), (wet_kg_m2, nx * ny),
                                                          (conc_kg_m3, nx * ny * nz))]
    s = NetcdfSnapshot(int(time_s), int(species_index), nx, ny, nz, cnt.ctypes.data, arrs[0].ctypes.data,
                       *[None if a is None else a.ctypes.data for a in opt])
    rc = lib.fp_write_netcdf(os.fsencode(path), C.byref(s), int(fmt))
    if rc != FP_OK:
        raise FpError(rc, lib.fp_last_error(None).decode())


def interpolation_alpha(t0_s: int, t1_s: int, target_s: int, strict: bool = False) -> float:
    """Temporal weight of the later met level (fp_interpolation_alpha;
    simulation/timeloop.rs:2086-2120).  Raises FpError outside the bracket
    when strict, and for an empty bracket."""
    lib = library()
    a = C.c_float()
    rc = lib.fp_interpolation_alpha(int(t0_s), int(t1_s), int(target_s), 1 if strict else 0,
                                    C.byref(a))
    if rc != FP_OK:
        raise FpError(rc, lib.fp_last_error(None).decode())
    return float(a.value)


def _grid_params(outheights, refine, species, mass_scale) -> GridParams:
    g = GridParams()
    g.refine_x, g.refine_y = refine
    g.species = species
    g.mass_scale = mass_scale
    if outheights is not None:
        for i, v in enumerate(outheights):
            g.outheights[i] = float(v)
    return g


class GpuEnsemble:
    """Many independent members advanced by one launch per step (fp_ens_*).

    Members share the vertical grid, the PBL state and the step parameters;
    each has its own particles, met bracket, Philox key and counter and
    output grid.  Every member's result equals what ``GpuTimestep.step``
    gives for that member alone with the same key, counter and alpha.
    """

    def __init__(self, capacities, met_shape, n_species: int = 4, out_shape=None,
                 device: int = 0):
        lib = library()
        caps = np.ascontiguousarray(np.asarray(capacities, np.uint64))
        nx, ny, nz = met_shape
        ox, oy, oz = out_shape if out_shape is not None else (0, 0, 0)
        d = EnsDesc(len(caps), nx, ny, nz, n_species, ox, oy, oz, caps.ctypes.data_as(_u64))
        h = _vp()
        rc = lib.fp_ens_create(device, C.byref(d), C.byref(h))
        if rc != FP_OK:
            raise FpError(rc, lib.fp_ens_last_error(None).decode())
        self._lib = lib
        self._h = h
        self.capacities = caps.copy()
        self.n_members = len(caps)
        self.met_shape = (nx, ny, nz)
        self.out_shape = (ox, oy, oz)
        self.n_species = n_species

    def _chk(self, rc):
        if rc != FP_OK:
            raise FpError(rc, self._lib.fp_ens_last_error(self._h).decode())

    def close(self):
        if self._h:
            self._lib.fp_ens_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        self._chk(self._lib.fp_ens_sync(self._h))

    def upload_member(self, member: int, particles: np.ndarray):
        p = np.ascontiguousarray(particles, dtype=PARTICLE_DTYPE)
        self._chk(self._lib.fp_ens_upload_member(self._h, member, p.ctypes.data, len(p)))

    def download_member(self, member: int, n=None) -> np.ndarray:
        n = int(self.capacities[member]) if n is None else int(n)
        out = np.zeros(n, PARTICLE_DTYPE)
        self._chk(self._lib.fp_ens_download_member(self._h, member, out.ctypes.data, n))
        return out

    def upload_met_member(self, member: int, u0, v0, w0, u1, v1, w1, level_heights):
        n = int(np.prod(self.met_shape))
        arrs = [_f32(a, n) for a in (u0, v0, w0, u1, v1, w1)]
        lh = _f32(level_heights, self.met_shape[2])
        self._chk(self._lib.fp_ens_met_upload_member(self._h, member, *[_fp(a) for a in arrs],
                                                     _fp(lh)))

    def upload_met_member_dev(self, member: int, ptrs, level_heights):
        lh = _f32(level_heights, self.met_shape[2])
        self._chk(self._lib.fp_ens_met_upload_member_dev(self._h, member,
                                                         *[_vp(int(p)) for p in ptrs], _fp(lh)))

    def upload_pbl(self, ustar, wstar, hmix, oli, precip=None):
        n = self.met_shape[0] * self.met_shape[1]
        arrs = [_f32(a, n) for a in (ustar, wstar, hmix, oli)]
        pr = None if precip is None else _f32(precip, n)
        self._chk(self._lib.fp_ens_pbl_upload(self._h, *[_fp(a) for a in arrs], _fp(pr)))

    def configure_grid(self, outheights=None, refine=(1, 1), species=0, mass_scale=1.0e6):
        g = _grid_params(outheights, refine, species, mass_scale)
        self._chk(self._lib.fp_ens_grid_configure(self._h, C.byref(g)))

    def clear_grid(self):
        self._chk(self._lib.fp_ens_grid_clear(self._h))

    def download_grid_member(self, member: int):
        n = int(np.prod(self.out_shape))
        counts = np.zeros(n, np.uint32)
        mass = np.zeros(n, np.uint32)
        self._chk(self._lib.fp_ens_grid_download_member(self._h, member, counts.ctypes.data_as(_u32),
                                                        mass.ctypes.data_as(_u32)))
        return counts, mass

    def configure(self, cfg: StepConfig, keys=None, counters=None, dispatch=None,
                  time_in_bracket_s: float = 0.0, bracket_s: float = 10800.0):
        """Step parameters, per-member keys (M x 2) / counters (M x 4) / dispatch."""
        k = None if keys is None else np.ascontiguousarray(np.asarray(keys, np.uint32).reshape(-1))
        c = None if counters is None else np.ascontiguousarray(np.asarray(counters, np.uint32).reshape(-1))
        d = None if dispatch is None else np.ascontiguousarray(np.asarray(dispatch, np.uint64))
        for arr, per in ((k, 2), (c, 4), (d, 1)):
            if arr is not None and arr.size != per * self.n_members:
                raise ValueError(f"expected {per} values per member")
        self._chk(self._lib.fp_ens_configure(
            self._h, C.byref(cfg.to_c()),
            None if k is None else k.ctypes.data_as(_u32),
            None if c is None else c.ctypes.data_as(_u32),
            None if d is None else d.ctypes.data_as(_u64),
            float(time_in_bracket_s), float(bracket_s)))

    def step(self, n_steps: int = 1, use_graph: bool = False) -> np.ndarray:
        """n_steps of every member; returns the (M, 4) counters after them."""
        out = np.zeros((self.n_members, 4), np.uint32)
        self._chk(self._lib.fp_ens_step(self._h, int(n_steps), 1 if use_graph else 0,
                                        out.ctypes.data_as(_u32)))
        return out

    def set_launch(self, block_size: int = 256, particles_per_thread: int = 1):
        self._chk(self._lib.fp_ens_set_launch(self._h, block_size, particles_per_thread))

    def sort_morton(self, bits_per_axis: int = 10, bounds=None):
        """Morton re-sort of every member's dispatch range (fp_ens_sort_morton)."""
        sp = SortParams()
        sp.bits_per_axis = bits_per_axis
        if bounds is not None:
            sp.explicit_bounds = 1
            for i, v in enumerate(bounds):
                sp.bounds[i] = float(v)
        self._chk(self._lib.fp_ens_sort_morton(self._h, C.byref(sp)))

    def autotune(self, sort_intervals=(0, 6, 12, 36), warmup_steps: int = 2, measured_steps: int = 7,
                 sort_trial_steps: int = 36, sort_bits: int = 10) -> dict:
        """Pick block size, particles/thread and re-sort interval of the graph
        path on the current state (fp_ens_autotune); the state is kept."""
        p = AutotuneParams()
        p.warmup_steps, p.measured_steps = warmup_steps, measured_steps
        p.sort_bits, p.sort_trial_steps = sort_bits, sort_trial_steps
        p.n_sort_intervals = len(sort_intervals)
        for i, s in enumerate(sort_intervals):
            p.sort_intervals[i] = s
        r = AutotuneResult()
        self._chk(self._lib.fp_ens_autotune(self._h, C.byref(p), C.byref(r)))
        return {"block_size": r.block_size, "particles_per_thread": r.particles_per_thread,
                "sort_interval": r.sort_interval, "step_ms": r.step_ms, "step_ms_with_sort": r.step_ms_with_sort}

    def last_step_ms(self) -> float:
        ms = C.c_double()
        self._chk(self._lib.fp_ens_last_step_ms(self._h, C.byref(ms)))
        return float(ms.value)


class GpuAdjoint:
    """Differentiable run of point sources (fp_adj_*): forward to receptor
    values and misfit, backward to d(misfit)/d(emissions, release positions,
    wind coefficients).

    ``particles_per_source`` fixes the sources (source s owns a contiguous
    slot range); the met holds ``n_met_levels`` time levels; receptors are
    nodes (i, j, l) of the observation lattice (met columns x ``obs_nz``
    levels ``obs_dz_m`` apart).
    """

    def __init__(self, particles_per_source, met_shape, n_met_levels: int = 2, n_modes: int = 0,
                 obs_nz: int = 11, obs_dz_m: float = 500.0, n_steps: int = 72,
                 obs_first_step: int = 1, checkpoint_every: int = 8, compact: bool = False,
                 emission_only: bool = False, device: int = 0):
        lib = library()
        counts = np.ascontiguousarray(np.asarray(particles_per_source, np.uint32))
        nx, ny, nz = met_shape
        flags = (ADJ_COMPACT if compact else 0) | (ADJ_EMISSION_ONLY if emission_only else 0)
        d = AdjDesc(len(counts), nx, ny, nz, n_met_levels, n_modes, obs_nz, obs_dz_m, n_steps,
                    obs_first_step, checkpoint_every, flags, counts.ctypes.data_as(_u32))
        h = _vp()
        rc = lib.fp_adj_create(device, C.byref(d), C.byref(h))
        if rc != FP_OK:
            raise FpError(rc, lib.fp_adj_last_error(None).decode())
        self._lib = lib
        self._h = h
        self.counts = counts.copy()
        self.n_sources = len(counts)
        self.n_particles = int(counts.sum())
        self.met_shape = (nx, ny, nz)
        self.n_met_levels = n_met_levels
        self.n_modes = n_modes
        self.n_steps = n_steps
        self.n_receptors = 0

    def _chk(self, rc):
        if rc != FP_OK:
            raise FpError(rc, self._lib.fp_adj_last_error(self._h).decode())

    def close(self):
        if self._h:
            self._lib.fp_adj_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_met(self, levels, level_heights, seconds_between_levels: float):
        """levels: n_met_levels tuples (u, v, w) of flat met arrays."""
        if len(levels) != self.n_met_levels:
            raise ValueError(f"expected {self.n_met_levels} met levels")
        n = int(np.prod(self.met_shape))
        u, v, w = (np.ascontiguousarray(np.concatenate([_f32(l[c], n) for l in levels]))
                   for c in range(3))
        lh = _f32(level_heights, self.met_shape[2])
        self._chk(self._lib.fp_adj_met_upload(self._h, _fp(u), _fp(v), _fp(w), _fp(lh),
                                              float(seconds_between_levels)))

    def upload_basis(self, basis_u, basis_v):
        n = self.n_modes * self.met_shape[0] * self.met_shape[1]
        self._chk(self._lib.fp_adj_basis_upload(self._h, _fp(_f32(basis_u, n)), _fp(_f32(basis_v, n))))

    def upload_pbl(self, ustar, wstar, hmix, oli, precip=None):
        n = self.met_shape[0] * self.met_shape[1]
        arrs = [_f32(a, n) for a in (ustar, wstar, hmix, oli)]
        pr = None if precip is None else _f32(precip, n)
        self._chk(self._lib.fp_adj_pbl_upload(self._h, *[_fp(a) for a in arrs], _fp(pr)))

    def set_receptors(self, nodes, observed=None):
        nd = np.ascontiguousarray(np.asarray(nodes, np.uint32).reshape(-1, 3))
        ob = None if observed is None else np.ascontiguousarray(np.asarray(observed, np.float64))
        if ob is not None and ob.size != len(nd):
            raise ValueError("one observation per receptor")
        self._chk(self._lib.fp_adj_receptors(self._h, len(nd), nd.ctypes.data_as(_u32),
                                             None if ob is None else ob.ctypes.data_as(_d)))
        self.n_receptors = len(nd)

    def forward(self, cfg: StepConfig, emission, position, coeff=None):
        """Run forward; returns (receptor values, loss)."""
        q = _f32(emission, self.n_sources)
        r = _f32(position, 3 * self.n_sources)
        c = _f32(np.zeros(0) if coeff is None else coeff, self.n_modes)
        vals = np.zeros(max(self.n_receptors, 1), np.float64)
        loss = C.c_double()
        self._chk(self._lib.fp_adj_forward(self._h, C.byref(cfg.to_c()), _fp(q), _fp(r),
                                           _fp(c) if self.n_modes else None,
                                           vals.ctypes.data_as(_d), C.byref(loss)))
        return vals[: self.n_receptors], float(loss.value)

    def backward(self):
        """Gradient of the last forward's loss: (d/d emission, d/d position (S, 3), d/d coeff)."""
        gq = np.zeros(self.n_sources, np.float64)
        gr = np.zeros((self.n_sources, 3), np.float64)
        gc = np.zeros(max(self.n_modes, 1), np.float64)
        self._chk(self._lib.fp_adj_backward(self._h, gq.ctypes.data_as(_d), gr.ctypes.data_as(_d),
                                            gc.ctypes.data_as(_d)))
        return gq, gr, gc[: self.n_modes]

    def active_counts(self) -> np.ndarray:
        """Active particles at the start of every step and after the last
        (n_steps + 1 values) of the last forward run."""
        out = np.zeros(self.n_steps + 1, np.uint64)
        self._chk(self._lib.fp_adj_active_counts(self._h, out.ctypes.data_as(_u64)))
        return out

    def download_state(self, n=None) -> np.ndarray:
        n = self.n_particles if n is None else int(n)
        out = np.zeros(n, PARTICLE_DTYPE)
        self._chk(self._lib.fp_adj_download_state(self._h, out.ctypes.data, n))
        return out

    def last_ms(self):
        f, b = C.c_double(), C.c_double()
        self._chk(self._lib.fp_adj_last_ms(self._h, C.byref(f), C.byref(b)))
        return float(f.value), float(b.value)


def receptor_misfit_function():
    """torch.autograd.Function over GpuAdjoint: ``apply(adj, cfg, q, r, c)``
    returns the misfit as a 0-d tensor; backward returns its gradient with
    respect to q (S), r (S, 3) and c (K).  Host tensors (float64 grads)."""
    import torch

    class ReceptorMisfit(torch.autograd.Function):
        @staticmethod
        def forward(ctx, adj, cfg, q, r, c):
            _, loss = adj.forward(cfg, q.detach().cpu().numpy(), r.detach().cpu().numpy(),
                                  None if c is None else c.detach().cpu().numpy())
            ctx.adj = adj
            return torch.tensor(loss, dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad_out):
            gq, gr, gc = ctx.adj.backward()
            g = float(grad_out)
            return (None, None, torch.from_numpy(gq * g), torch.from_numpy(gr * g),
                    torch.from_numpy(gc * g))

    return ReceptorMisfit


__all__ = [
    "GpuTimestep", "GpuEnsemble", "GpuAdjoint", "receptor_misfit_function", "StepConfig",
    "interpolation_alpha", "PARTICLE_DTYPE", "SOA_FIELDS", "FpError", "LibraryMissing",
    "library", "ABI", "LIB_PATH",
]
