"""ctypes binding of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Imported by tests/,
``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py``; never
by the product.

Particles are numpy structured arrays with the reference's 96-byte layout
(particles/mod.rs:76-168).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")

PARTICLE_DTYPE = np.dtype(
    [
        ("pos_x", "<f4"), ("pos_y", "<f4"), ("pos_z", "<f4"),
        ("cell_x", "<i4"), ("cell_y", "<i4"), ("flags", "<u4"),
        ("mass", "<f4", (4,)),
        ("vel_u", "<f4"), ("vel_v", "<f4"), ("vel_w", "<f4"),
        ("turb_u", "<f4"), ("turb_v", "<f4"), ("turb_w", "<f4"),
        ("time", "<i4"), ("timestep", "<i4"), ("time_mem", "<i4"), ("time_split", "<i4"),
        ("release_point", "<i4"), ("class_id", "<i4"), ("cbt", "<i4"), ("pad0", "<u4"),
    ]
)
assert PARTICLE_DTYPE.itemsize == 96

_f = C.POINTER(C.c_float)
_u32 = C.POINTER(C.c_uint32)


class Met(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32),
                ("u0", _f), ("v0", _f), ("w0", _f), ("u1", _f), ("v1", _f), ("w1", _f),
                ("level_heights", _f)]


class Pbl(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32),
                ("ustar", _f), ("wstar", _f), ("hmix", _f), ("oli", _f), ("precip", _f)]


class StepParams(C.Structure):
    _fields_ = [
        ("dt_seconds", C.c_float), ("alpha", C.c_float),
        ("x_scale", C.c_float), ("y_scale", C.c_float), ("z_scale", C.c_float),
        ("direction", C.c_int32),
        ("key", C.c_uint32 * 2), ("counter", C.c_uint32 * 4),
        ("n_substeps", C.c_uint32), ("min_height_m", C.c_float), ("rho_grad_over_rho", C.c_float),
        ("dry_mode", C.c_uint32), ("dry_vd", C.c_float), ("dry_href", C.c_float),
        ("dry_vd_slots", _f),
        ("wet_mode", C.c_uint32), ("wet_lambda", C.c_float), ("wet_fraction", C.c_float),
        ("wet_lambda_slots", _f), ("wet_fraction_slots", _f),
        ("wet_a", C.c_float), ("wet_b", C.c_float),
        ("loss_mode", C.c_uint32), ("loss_min_mass", C.c_float),
        ("n_threads", C.c_int32),
    ]


class GridParams(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32),
                ("refine_x", C.c_uint32), ("refine_y", C.c_uint32),
                ("species", C.c_uint32), ("mass_scale", C.c_float),
                ("outheights", C.c_float * 16)]


_LIBS: dict[str, C.CDLL] = {}


def lib(variant: str = "fpm") -> C.CDLL:
    """Load the oracle: ``"fpm"`` (deterministic math) or ``"libm"`` (glibc)."""
    if variant not in _LIBS:
        name = "liboracle.so" if variant == "fpm" else "liboracle_libm.so"
        path = os.path.join(BUILD, name)
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run make -C oracle)")
        L = C.CDLL(path)
        L.orc_step.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(Met), C.POINTER(Pbl),
                               C.POINTER(StepParams), _u32]
        L.orc_step.restype = C.c_int
        L.orc_philox4x32_10.argtypes = [_u32, _u32, _u32]
        L.orc_counter_add.argtypes = [_u32, C.c_uint64, _u32]
        L.orc_philox_stream.argtypes = [_u32, _u32, C.c_uint64, _u32]
        L.orc_u32_to_uniform01.argtypes = [C.c_uint32]
        L.orc_u32_to_uniform01.restype = C.c_float
        L.orc_box_muller.argtypes = [C.c_float, C.c_float, _f]
        L.orc_hanna.argtypes = [C.c_float] * 5 + [_f]
        L.orc_hanna_short_w.argtypes = [C.c_float] * 5 + [_f]
        L.orc_reflect.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_int32),
                                  C.c_float, C.c_float]
        L.orc_update_horizontal.argtypes = [C.c_float] * 5
        L.orc_update_horizontal.restype = C.c_float
        L.orc_update_vertical.argtypes = [C.c_float] * 7
        L.orc_update_vertical.restype = C.c_float
        L.orc_deposit.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(Pbl), C.POINTER(StepParams)]
        L.orc_advect.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(Met), C.POINTER(StepParams)]
        L.orc_grid_accumulate.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(GridParams),
                                          _u32, _u32]
        L.orc_compact.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, _u32]
        L.orc_compact.restype = C.c_uint64
        L.orc_morton_encode_3d_21bit.argtypes = [C.c_uint32] * 3
        L.orc_morton_encode_3d_21bit.restype = C.c_uint64
        L.orc_morton_keys.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_int,
                                      C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
        L.orc_sort_morton.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_int,
                                      C.POINTER(C.c_double), C.c_void_p, _u32]
        L.orc_sort_morton.restype = C.c_int
        L.orc_math_eval.argtypes = [C.c_int, _f, _f, _f, C.c_uint64]
        L.orc_level_eval.argtypes = [_f, C.c_uint32, _f, _f, C.c_uint64]
        _LIBS[variant] = L
    return _LIBS[variant]


def _fp(a):
    return None if a is None else a.ctypes.data_as(_f)


def _u(a):
    return a.ctypes.data_as(_u32)


def _arr(x, dtype):
    return np.ascontiguousarray(np.asarray(x, dtype=dtype))


# ---------------------------------------------------------------- Philox

def philox4x32_10(counter, key, variant="fpm"):
    c = _arr(counter, np.uint32)
    k = _arr(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib(variant).orc_philox4x32_10(_u(c), _u(k), _u(out))
    return out


def counter_add(counter, offset):
    c = _arr(counter, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_counter_add(_u(c), C.c_uint64(int(offset)), _u(out))
    return out


def philox_stream(counter, key, n_blocks):
    c = _arr(counter, np.uint32)
    k = _arr(key, np.uint32)
    out = np.zeros((int(n_blocks), 4), np.uint32)
    lib().orc_philox_stream(_u(c), _u(k), C.c_uint64(int(n_blocks)), _u(out))
    return out


def u32_to_uniform01(v):
    return lib().orc_u32_to_uniform01(C.c_uint32(int(v)))


def box_muller(u0, u1, variant="fpm"):
    out = np.zeros(2, np.float32)
    lib(variant).orc_box_muller(C.c_float(u0), C.c_float(u1), _fp(out))
    return out


HANNA_FIELDS = ("ust", "wst", "ol", "h", "zeta", "sigu", "sigv", "sigw", "dsigwdz",
                "dsigw2dz", "tlu", "tlv", "tlw")


def hanna(ust, wst, ol, h, z, variant="fpm"):
    out = np.zeros(13, np.float32)
    lib(variant).orc_hanna(C.c_float(ust), C.c_float(wst), C.c_float(ol), C.c_float(h),
                           C.c_float(z), _fp(out))
    return dict(zip(HANNA_FIELDS, out.tolist()))


def hanna_short_w(ust, wst, ol, h, z, variant="fpm"):
    """hanna_short_w: (sigma_w, dsigma_w/dz, T_Lw) at height z."""
    out = np.zeros(3, np.float32)
    lib(variant).orc_hanna_short_w(C.c_float(ust), C.c_float(wst), C.c_float(ol), C.c_float(h),
                                   C.c_float(z), _fp(out))
    return dict(zip(("sigw", "dsigwdz", "tlw"), out.tolist()))


def reflect(z, turb_w, cbt, z_min, h, variant="fpm"):
    """PBL reflection of one particle: returns (z, turb_w, cbt)."""
    zz, tw, cb = C.c_float(z), C.c_float(turb_w), C.c_int32(cbt)
    lib(variant).orc_reflect(C.byref(zz), C.byref(tw), C.byref(cb), C.c_float(z_min), C.c_float(h))
    return zz.value, tw.value, cb.value


def update_horizontal(prev, sigma, tl, dt, eta, variant="fpm"):
    return lib(variant).orc_update_horizontal(prev, sigma, tl, dt, eta)


def update_vertical(prev, sigw, dsigwdz, tlw, dt, rho, eta, variant="fpm"):
    return lib(variant).orc_update_vertical(prev, sigw, dsigwdz, tlw, dt, rho, eta)


# ---------------------------------------------------------------- step

class _Keep:
    """Holds numpy arrays alive while ctypes structs point into them."""

    def __init__(self):
        self.refs = []

    def f32(self, a):
        if a is None:
            return None
        a = _arr(a, np.float32)
        self.refs.append(a)
        return a


def make_met(met: dict, keep: _Keep) -> Met:
    nx, ny, nz = met["shape"]
    lh = met.get("level_heights")
    if lh is None:
        lh = np.zeros(nz, np.float32)
    return Met(nx, ny, nz,
               *[_fp(keep.f32(np.asarray(met[k]).reshape(-1)))
                 for k in ("u0", "v0", "w0", "u1", "v1", "w1")],
               _fp(keep.f32(lh)))


def make_pbl(pbl: dict, keep: _Keep) -> Pbl:
    nx, ny = pbl["shape"]
    pre = pbl.get("precip")
    return Pbl(nx, ny, *[_fp(keep.f32(np.asarray(pbl[k]).reshape(-1)))
                         for k in ("ustar", "wstar", "hmix", "oli")],
               _fp(keep.f32(None if pre is None else np.asarray(pre).reshape(-1))))


STEP_DEFAULTS = dict(
    dt_seconds=600.0, alpha=0.0, x_scale=1.0, y_scale=1.0, z_scale=1.0, direction=1,
    key=(0xDECAFBAD, 0x12345678), counter=(0, 0, 0, 0), n_substeps=4, min_height_m=0.01,
    rho_grad_over_rho=0.0, dry_mode=0, dry_vd=0.0, dry_href=15.0, dry_vd_slots=None,
    wet_mode=0, wet_lambda=0.0, wet_fraction=0.0, wet_lambda_slots=None,
    wet_fraction_slots=None, wet_a=0.0, wet_b=0.0, loss_mode=0, loss_min_mass=0.0,
    n_threads=1,
)


def make_step_params(params: dict, keep: _Keep) -> StepParams:
    p = dict(STEP_DEFAULTS)
    unknown = set(params) - set(p)
    if unknown:
        raise KeyError(f"unknown step parameters: {sorted(unknown)}")
    p.update(params)
    sp = StepParams()
    for name, _ in StepParams._fields_:
        v = p[name]
        if name in ("key", "counter"):
            arr = getattr(sp, name)
            for i, x in enumerate(v):
                arr[i] = int(x) & 0xFFFFFFFF
        elif name.endswith("_slots"):
            setattr(sp, name, _fp(keep.f32(v)))
        else:
            setattr(sp, name, v)
    return sp


def step(particles: np.ndarray, met: dict, pbl: dict, params: dict, n_dispatch=None,
         variant="fpm"):
    """One production timestep in place; returns the next Philox counter."""
    assert particles.dtype == PARTICLE_DTYPE and particles.flags.c_contiguous
    n = len(particles) if n_dispatch is None else int(n_dispatch)
    keep = _Keep()
    m, b, sp = make_met(met, keep), make_pbl(pbl, keep), make_step_params(params, keep)
    nxt = np.zeros(4, np.uint32)
    rc = lib(variant).orc_step(particles.ctypes.data, n, C.byref(m), C.byref(b), C.byref(sp),
                               _u(nxt))
    if rc != 0:
        raise RuntimeError(f"orc_step failed: {rc}")
    return nxt


def advect(particles: np.ndarray, met: dict, params: dict, variant="fpm"):
    """Advection stage only (advection_dual_wind.wgsl), in place."""
    keep = _Keep()
    m, sp = make_met(met, keep), make_step_params(params, keep)
    lib(variant).orc_advect(particles.ctypes.data, len(particles), C.byref(m), C.byref(sp))


def deposit(particles: np.ndarray, params: dict, pbl: dict | None = None, variant="fpm"):
    """Dry then wet deposition only (the last two stages of the step), in place."""
    keep = _Keep()
    b = make_pbl(pbl, keep) if pbl is not None else Pbl()
    sp = make_step_params(params, keep)
    lib(variant).orc_deposit(particles.ctypes.data, len(particles), C.byref(b), C.byref(sp))


# ---------------------------------------------------------------- gridding

def make_grid_params(shape, outheights=None, refine=(1, 1), species=0, mass_scale=1.0e6):
    nx, ny, nz = shape
    gp = GridParams()
    gp.nx, gp.ny, gp.nz = nx, ny, nz
    gp.refine_x, gp.refine_y = refine
    gp.species = species
    gp.mass_scale = mass_scale
    oh = np.zeros(16, np.float32)
    if outheights is not None:
        oh[: len(outheights)] = outheights
    for i in range(16):
        gp.outheights[i] = float(oh[i])
    return gp


def grid_accumulate(particles, shape, outheights=None, refine=(1, 1), species=0,
                    mass_scale=1.0e6, counts=None, mass_fp=None):
    nx, ny, nz = shape
    if counts is None:
        counts = np.zeros(nx * ny * nz, np.uint32)
    if mass_fp is None:
        mass_fp = np.zeros(nx * ny * nz, np.uint32)
    gp = make_grid_params(shape, outheights, refine, species, mass_scale)
    lib().orc_grid_accumulate(particles.ctypes.data, len(particles), C.byref(gp),
                              _u(counts), _u(mass_fp))
    return counts, mass_fp


# ---------------------------------------------------------------- compaction / sort

def compact(particles):
    out = np.zeros_like(particles)
    smap = np.zeros(len(particles), np.uint32)
    a = lib().orc_compact(particles.ctypes.data, len(particles), out.ctypes.data, _u(smap))
    return int(a), out, smap[:a].copy()


def morton_encode(x, y, z):
    return int(lib().orc_morton_encode_3d_21bit(int(x), int(y), int(z)))


def _bounds_ptr(bounds):
    if bounds is None:
        return 0, None
    b = (C.c_double * 6)(*[float(v) for v in bounds])
    return 1, b


def morton_keys(particles, bits, bounds=None):
    explicit, b = _bounds_ptr(bounds)
    keys = np.zeros(len(particles), np.uint64)
    lib().orc_morton_keys(particles.ctypes.data, len(particles), bits, explicit, b,
                          keys.ctypes.data_as(C.POINTER(C.c_uint64)))
    return keys


def sort_morton(particles, bits, bounds=None):
    explicit, b = _bounds_ptr(bounds)
    out = np.zeros_like(particles)
    n2o = np.zeros(len(particles), np.uint32)
    rc = lib().orc_sort_morton(particles.ctypes.data, len(particles), bits, explicit, b,
                               out.ctypes.data, _u(n2o))
    if rc != 0:
        raise ValueError(f"orc_sort_morton failed: {rc}")
    return out, n2o


MATH_FN = {"exp": 0, "exp_libm": 1, "log": 2, "log_libm": 3, "sin": 4, "sin_libm": 5,
           "cos": 6, "cos_libm": 7, "pow": 8, "pow_libm": 9, "exp_d": 10, "log_d": 11,
           "sin_d": 12, "cos_d": 13, "pow_d": 14}


def math_eval(fn: str, x, y=None):
    x = _arr(x, np.float32)
    y = _arr(np.zeros_like(x) if y is None else y, np.float32)
    out = np.zeros_like(x)
    lib().orc_math_eval(MATH_FN[fn], _fp(x), _fp(y), _fp(out), len(x))
    return out


def level_eval(level_heights, z):
    """meters_to_grid_level of each height z over the level table
    (advection_dual_wind.wgsl:94-117)."""
    lh = _arr(level_heights, np.float32)
    z = _arr(z, np.float32)
    out = np.zeros_like(z)
    lib().orc_level_eval(_fp(lh), len(lh), _fp(z), _fp(out), len(z))
    return out


# ---------------------------------------------------------------- release

class GridDomain(C.Structure):
    _fields_ = [("xlon0", C.c_double), ("ylat0", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
                ("nx", C.c_uint32), ("ny", C.c_uint32)]


class Release(C.Structure):
    _fields_ = [("lon_min", C.c_double), ("lon_max", C.c_double), ("lat_min", C.c_double),
                ("lat_max", C.c_double), ("z_min", C.c_double), ("z_max", C.c_double),
                ("mass_kg", C.c_double), ("particle_count", C.c_uint64),
                ("start_s", C.c_int64), ("end_s", C.c_int64),
                ("release_point", C.c_int32), ("class_id", C.c_int32)]


def release_inject(grid: dict, releases: list, released, now_s: int, particle_time: int,
                   max_out: int = 1 << 20):
    """ReleaseManager::inject_for_time (release/mod.rs:90-152).  grid: dict of
    GridDomain fields; releases: dicts of Release fields; released: per-release
    emitted counts (advanced in place).  Returns (particles, status)."""
    g = GridDomain(**grid)
    arr = (Release * max(len(releases), 1))()
    for i, r in enumerate(releases):
        arr[i] = Release(**r)
    rel = np.ascontiguousarray(np.asarray(released, np.uint64))
    out = np.zeros(max_out, PARTICLE_DTYPE)
    st = C.c_int()
    L = lib()
    fn = L.orc_release_inject
    fn.restype = C.c_uint64
    fn.argtypes = [C.POINTER(GridDomain), C.POINTER(Release), C.c_uint32, C.POINTER(C.c_uint64), C.c_int64,
                   C.c_int32, C.c_void_p, C.c_uint64, C.POINTER(C.c_int)]
    n = fn(C.byref(g), arr, len(releases), rel.ctypes.data_as(C.POINTER(C.c_uint64)), now_s, particle_time,
           out.ctypes.data, max_out, C.byref(st))
    released[:] = rel
    return out[:n].copy(), int(st.value)


class PblOptions(C.Structure):
    _fields_ = [("roughness_length_m", C.c_float), ("wind_reference_height_m", C.c_float),
                ("heat_flux_neutral_threshold_w_m2", C.c_float), ("fallback_mixing_height_m", C.c_float),
                ("hmix_min_m", C.c_float), ("hmix_max_m", C.c_float)]


PBL_DEFAULTS = dict(roughness_length_m=0.1, wind_reference_height_m=10.0, heat_flux_neutral_threshold_w_m2=1.0,
                    fallback_mixing_height_m=800.0, hmix_min_m=100.0, hmix_max_m=4500.0)
PBL_OUTPUTS = ("ustar", "wstar", "hmix", "oli", "sshf", "ssr", "surfstr")


def pbl_from_surface(s0, s1, alpha: float, options: dict | None = None, variant="fpm") -> dict:
    """interpolate_surface_fields_linear (io/temporal.rs:122-214) + the GPU PBL
    diagnostics (pbl_diagnostics.wgsl:220-305).  s0, s1: sequences of the ten
    surface fields (pressure, T2m, u10, v10, stress, sshf, ssr, hmix, ustar,
    1/L) of equal length.  Returns the seven outputs by name."""
    a0 = [np.ascontiguousarray(np.asarray(x, np.float32).reshape(-1)) for x in s0]
    a1 = [np.ascontiguousarray(np.asarray(x, np.float32).reshape(-1)) for x in s1]
    if len(a0) != 10 or len(a1) != 10:
        raise ValueError("ten surface fields per level")
    n = a0[0].size
    if any(x.size != n for x in a0 + a1):
        raise ValueError("surface fields differ in size")
    o = PblOptions(**{**PBL_DEFAULTS, **(options or {})})
    out = {k: np.empty(n, np.float32) for k in PBL_OUTPUTS}
    p0 = (C.c_void_p * 10)(*[x.ctypes.data for x in a0])
    p1 = (C.c_void_p * 10)(*[x.ctypes.data for x in a1])
    po = (C.c_void_p * 7)(*[out[k].ctypes.data for k in PBL_OUTPUTS])
    L = lib(variant)
    fn = L.orc_pbl_from_surface
    fn.restype = None
    fn.argtypes = [C.c_uint64, C.c_void_p, C.c_void_p, C.c_float, C.POINTER(PblOptions), C.c_void_p]
    fn(n, p0, p1, alpha, C.byref(o), po)
    return out


def output_volumes(ylat0: float, dx: float, dy: float, outheights, ny: int) -> np.ndarray:
    """Cell volumes [m^3] per output row and level (orc_output_volumes): (ny, nz)."""
    oh = np.ascontiguousarray(np.asarray(outheights, np.float32).reshape(-1))
    vol = np.empty(ny * oh.size, np.float32)
    fn = lib().orc_output_volumes
    fn.restype = None
    fn.argtypes = [C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]
    fn(ylat0, dx, dy, oh.ctypes.data, ny, oh.size, vol.ctypes.data)
    return vol.reshape(ny, oh.size)


def output_finalize(acc_count, acc_mass, n_samples: int, mass_scale: float, shape, vol=None) -> dict:
    """Interval fields from the 64-bit sums (orc_output_finalize)."""
    nx, ny, nz = shape
    ac = np.ascontiguousarray(np.asarray(acc_count, np.uint64).reshape(-1))
    am = np.ascontiguousarray(np.asarray(acc_mass, np.uint64).reshape(-1))
    n = ac.size
    count = np.empty(n, np.uint32)
    mass = np.empty(n, np.float32)
    conc = np.empty(n, np.float32) if vol is not None else None
    v = None if vol is None else np.ascontiguousarray(np.asarray(vol, np.float32).reshape(-1))
    fn = lib().orc_output_finalize
    fn.restype = None
    fn.argtypes = [C.c_uint64, C.c_void_p, C.c_void_p, C.c_uint32, C.c_float, C.c_void_p, C.c_uint32, C.c_uint32,
                   C.c_void_p, C.c_void_p, C.c_void_p]
    fn(n, ac.ctypes.data, am.ctypes.data, n_samples, mass_scale, None if v is None else v.ctypes.data, ny, nz,
       count.ctypes.data, mass.ctypes.data, None if conc is None else conc.ctypes.data)
    out = {"count": count, "mass_kg": mass}
    if conc is not None:
        out["conc_kg_m3"] = conc
    return out


def hybrid_to_height(ps, t, ak, bk, q=None, variant="fpm"):
    """transform_hybrid_sigma_to_height (orc_hybrid_to_height): ps (nx, ny),
    t and q (nx, ny, nz), ak and bk (nz).  Returns (code, (i, j, k), pressure,
    height, mean)."""
    ps = np.ascontiguousarray(np.asarray(ps, np.float32))
    t = np.ascontiguousarray(np.asarray(t, np.float32))
    nx, ny, nz = t.shape
    qa = None if q is None else np.ascontiguousarray(np.asarray(q, np.float32))
    a = np.ascontiguousarray(np.asarray(ak, np.float32))
    b = np.ascontiguousarray(np.asarray(bk, np.float32))
    pressure = np.zeros((nx, ny, nz), np.float32)
    height = np.zeros((nx, ny, nz), np.float32)
    mean = np.zeros(nz, np.float32)
    ijk = (C.c_uint32 * 3)()
    fn = lib(variant).orc_hybrid_to_height
    fn.restype = C.c_int
    fn.argtypes = [C.c_uint32] * 3 + [C.c_void_p] * 8 + [C.c_void_p]
    code = fn(nx, ny, nz, ps.ctypes.data, t.ctypes.data, None if qa is None else qa.ctypes.data, a.ctypes.data,
              b.ctypes.data, pressure.ctypes.data, height.ctypes.data, mean.ctypes.data, ijk)
    return code, tuple(ijk), pressure, height, mean
