"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

The oracle is a plain-C restatement of the reference estimator
(crates/sky-spectral-path-tracer/src/sky_trace.wgsl) over the analytic
atmosphere; see sky_oracle.h.  Only tests/, the smoke check and the
CPU-baseline leg of bench.py import this module, as a checker.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field, fields
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libsky_oracle.so")

FLAG_PHASE_IS = 1
FLAG_RATIO_TRACKING = 2
FLAG_CARRIED_TAU = 4
PROJ_PANORAMA = 0
PROJ_SKY_VIEW_LUT = 1


class _Params(C.Structure):
    _fields_ = [
        ("ground_radius_km", C.c_double),
        ("atmosphere_radius_km", C.c_double),
        ("rayleigh_sca_550_m_inv", C.c_double),
        ("rayleigh_h_km", C.c_double),
        ("ozone_peak_m_inv", C.c_double),
        ("ozone_center_km", C.c_double),
        ("ozone_half_width_km", C.c_double),
        ("ozone_lambda0_nm", C.c_double),
        ("ozone_sigma_nm", C.c_double),
        ("aerosol_sca_550_m_inv", C.c_double),
        ("aerosol_abs_550_m_inv", C.c_double),
        ("aerosol_h_km", C.c_double),
        ("aerosol_g", C.c_double),
        ("angstrom", C.c_double),
        ("ground_albedo", C.c_double),
        ("sun_temperature_k", C.c_double),
        ("sun_angular_radius_rad", C.c_double),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("spp", C.c_uint32),
        ("sample_offset", C.c_uint32),
        ("seed", C.c_uint64),
        ("sun_elevation_deg", C.c_double),
        ("sun_azimuth_deg", C.c_double),
        ("observer_altitude_km", C.c_double),
        ("direct_light_samples", C.c_uint32),
        ("projection", C.c_uint32),
        ("max_depth", C.c_uint32),
        ("rr_start_depth", C.c_uint32),
        ("watchdog_limit", C.c_uint32),
        ("max_segments", C.c_uint32),
        ("majorant_layers", C.c_uint32),
        ("flags", C.c_uint32),
        ("majorant_margin", C.c_double),
        ("n_lambda", C.c_uint32),
        ("lambda_nm", C.POINTER(C.c_double)),
        ("weight_nm", C.POINTER(C.c_double)),
        ("shell_growth", C.c_double),
        ("tables", C.c_void_p),
    ]


class _Tables(C.Structure):
    """orc_tables: the reference's tabulated SceneData (data.rs:53-152)."""
    _fields_ = [
        ("atmosphere_points", C.c_uint32),
        ("aerosol_points", C.c_uint32),
        ("atmosphere", C.POINTER(C.c_double)),
        ("aerosol", C.POINTER(C.c_double)),
        ("aerosol_optics", C.POINTER(C.c_double)),
        ("ozone_cross_section_cm2", C.POINTER(C.c_double)),
        ("rayleigh_cross_section_m2", C.POINTER(C.c_double)),
        ("solar_irradiance_w_m2", C.POINTER(C.c_double)),
    ]


class Counters(C.Structure):
    _fields_ = [
        (n, C.c_uint64)
        for n in (
            "paths",
            "boundary_queries",
            "segments",
            "free_flights",
            "density_evals",
            "scatter_vertices",
            "ground_vertices",
            "nee_samples",
            "shadow_rays",
            "rng_draws",
            "watchdog",
            "graze64",
        )
    ]

    def as_dict(self) -> dict:
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class _Lambda(C.Structure):
    _fields_ = [
        (n, C.c_double)
        for n in ("beta_r", "beta_o", "beta_as", "beta_aa", "sun_radiance", "wx", "wy", "wz")
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library not built: {_LIB_PATH} (run make -C oracle)")
        L = C.CDLL(_LIB_PATH)
        P = C.POINTER(_Params)
        dp = C.POINTER(C.c_double)
        L.orc_pcg4d.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_rng_u32.argtypes = [C.c_uint32] * 4 + [C.c_uint64]
        L.orc_rng_u32.restype = C.c_uint32
        L.orc_rng_f32.argtypes = [C.c_uint32] * 4 + [C.c_uint64]
        L.orc_rng_f32.restype = C.c_float
        L.orc_cmf_xyz.argtypes = [C.c_double, dp]
        L.orc_planck_radiance.argtypes = [C.c_double, C.c_double]
        L.orc_planck_radiance.restype = C.c_double
        L.orc_default_params.argtypes = [P]
        L.orc_lambda_table.argtypes = [P, C.POINTER(_Lambda)]
        L.orc_sun_direction.argtypes = [P, dp]
        L.orc_majorant_grid.argtypes = [P, dp]
        L.orc_coefficients.argtypes = [P, C.POINTER(_Lambda), C.c_double, dp, dp, dp, dp]
        L.orc_next_boundary_reference.argtypes = [P, dp, dp, dp, C.POINTER(C.c_int)]
        for fn in (L.orc_segments_reference, L.orc_segments_walk):
            fn.argtypes = [P, dp, dp, C.c_double, dp, C.POINTER(C.c_int), C.c_int]
        L.orc_optical_depth.argtypes = [P, C.POINTER(_Lambda), dp, dp, C.c_double]
        L.orc_optical_depth.restype = C.c_double
        L.orc_optical_depth_band.argtypes = [P, C.c_uint32, dp, dp, C.c_double]
        L.orc_optical_depth_band.restype = C.c_double
        L.orc_optical_depth_batch.argtypes = [P, C.c_uint32, C.c_uint32, dp, dp, dp, dp]
        L.orc_coefficients_band.argtypes = [P, C.c_uint32, C.c_double, dp, dp, dp, dp]
        L.orc_camera_dir.argtypes = [P, C.c_double, C.c_double, dp]
        for s in ("f32", "f64"):
            fn = getattr(L, f"orc_render_{s}")
            fn.argtypes = [P, dp, dp, dp, C.POINTER(Counters), C.c_int]
            fn = getattr(L, f"orc_render_pixels_{s}")
            fn.argtypes = [P, C.POINTER(C.c_uint32), C.c_uint32, dp, dp, C.POINTER(Counters), C.c_int]
            fn = getattr(L, f"orc_trace_path_{s}")
            fn.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(Counters)]
            fn.restype = C.c_double
        L.orc_ray_setup_f32.argtypes = [P, dp, dp, C.c_double, dp]
        L.orc_ray_setup_f32.restype = C.c_int
        L.orc_transmittance_mc_f32.argtypes = [P, C.c_uint32, dp, dp, C.c_uint32, C.c_int]
        L.orc_transmittance_mc_f32.restype = C.c_double
        for name, args in (("orc_sample_rayleigh_mu", [C.c_double]), ("orc_sample_hg_mu", [C.c_double] * 2),
                           ("orc_phase_rayleigh", [C.c_double]), ("orc_phase_hg", [C.c_double] * 2)):
            getattr(L, name).argtypes = args
            getattr(L, name).restype = C.c_double
        L.orc_frame_dir.argtypes = [dp, C.c_double, C.c_double, C.c_double, dp]
        L.orc_sun_cone_dir.argtypes = [P, C.c_double, C.c_double, dp]
        L.orc_sun_cone_dir.restype = C.c_double
        L.orc_cosine_dir.argtypes = [dp, C.c_double, C.c_double, dp]
        L.orc_cosine_dir.restype = C.c_double
        fp = C.POINTER(C.c_float)
        u8 = C.c_void_p
        L.orc_band_widths.argtypes = [fp, C.c_uint32, dp]
        L.orc_quadrature_weights.argtypes = [fp, C.c_uint32, dp, u8, C.c_uint32, C.c_uint32, dp]
        L.orc_quadrature_weights.restype = C.c_int
        L.orc_xyz_to_oklab.argtypes = [dp, dp]
        L.orc_percentile.argtypes = [dp, C.c_uint64, C.c_double]
        L.orc_percentile.restype = C.c_double
        L.orc_reference_oklab.argtypes = [fp, C.c_uint32, C.c_uint64, dp, dp]
        L.orc_reference_oklab.restype = C.c_double
        L.orc_evaluate_combos.argtypes = [fp, C.c_uint32, C.c_uint64, dp, C.c_double, fp, dp, u8, C.c_uint32,
                                          C.c_uint64, dp, C.c_void_p, C.c_uint64, C.c_int, C.c_int, dp]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def midpoint_grid(n: int, lo: float = 380.0, hi: float = 780.0) -> np.ndarray:
    """Midpoint-rule wavelength grid: 16 -> 392.5 + 25 i, 64 -> 383.125 + 6.25 i."""
    step = (hi - lo) / n
    return lo + step * (np.arange(n) + 0.5)


@dataclass
class Params:
    """Analytic atmosphere + render configuration (defaults = the 128x64 case)."""

    ground_radius_km: float = 6360.0
    atmosphere_radius_km: float = 6460.0
    rayleigh_sca_550_m_inv: float = 1.3558e-5
    rayleigh_h_km: float = 8.0
    ozone_peak_m_inv: float = 1.9e-6
    ozone_center_km: float = 25.0
    ozone_half_width_km: float = 15.0
    ozone_lambda0_nm: float = 600.0
    ozone_sigma_nm: float = 80.0
    aerosol_sca_550_m_inv: float = 3.996e-6
    aerosol_abs_550_m_inv: float = 4.4e-6
    aerosol_h_km: float = 1.2
    aerosol_g: float = 0.8
    angstrom: float = 1.3
    ground_albedo: float = 0.3
    sun_temperature_k: float = 5778.0
    sun_angular_radius_rad: float = 0.2675 * np.pi / 180.0
    width: int = 128
    height: int = 64
    spp: int = 64
    sample_offset: int = 0
    seed: int = 1
    sun_elevation_deg: float = 10.0
    sun_azimuth_deg: float = 0.0
    observer_altitude_km: float = 0.2
    direct_light_samples: int = 1
    projection: int = PROJ_PANORAMA
    max_depth: int = 8
    rr_start_depth: int = 4
    watchdog_limit: int = 1024
    max_segments: int = 100000
    majorant_layers: int = 64
    flags: int = 0
    majorant_margin: float = 1.05
    lambda_nm: Sequence[float] = field(default_factory=lambda: list(midpoint_grid(16)))
    weight_nm: Optional[Sequence[float]] = None
    shell_growth: float = 1.0
    # Tabulated medium: any object with the arrays atmosphere [n, 3] (altitude,
    # air_cm3, ozone_cm3), aerosol [n, 5], aerosol_optics [n_lambda, 4, 2],
    # ozone_cross_section_cm2 [n_lambda] and optional rayleigh_cross_section_m2,
    # solar_irradiance_w_m2 [n_lambda].  None: the analytic profiles.
    tables: Optional[object] = None

    def to_c(self):
        p = _Params()
        for f in fields(self):
            if f.name in ("lambda_nm", "weight_nm", "tables"):
                continue
            setattr(p, f.name, getattr(self, f.name))
        keep = []
        if self.tables is not None:
            t = self.tables
            arrs = {k: np.ascontiguousarray(getattr(t, k), dtype=np.float64)
                    for k in ("atmosphere", "aerosol", "aerosol_optics", "ozone_cross_section_cm2")}
            opt = {k: (None if getattr(t, k, None) is None
                       else np.ascontiguousarray(getattr(t, k), dtype=np.float64))
                   for k in ("rayleigh_cross_section_m2", "solar_irradiance_w_m2")}
            tc = _Tables(atmosphere_points=arrs["atmosphere"].shape[0], aerosol_points=arrs["aerosol"].shape[0])
            for k, a in {**arrs, **opt}.items():
                setattr(tc, k, None if a is None else _dp(a))
            p.tables = C.cast(C.pointer(tc), C.c_void_p)
            keep += [tc, arrs, opt]
        lam = np.ascontiguousarray(self.lambda_nm, dtype=np.float64)
        if self.weight_nm is None:
            n = len(lam)
            w = np.full(n, (lam[-1] - lam[0]) / (n - 1) if n > 1 else 1.0)
        else:
            w = np.ascontiguousarray(self.weight_nm, dtype=np.float64)
        p.n_lambda = len(lam)
        p.lambda_nm = _dp(lam)
        p.weight_nm = _dp(w)
        # keep the arrays alive with the struct
        p._keep = (lam, w, keep)
        return p

    @property
    def n_lambda(self) -> int:
        return len(self.lambda_nm)

    def weights(self) -> np.ndarray:
        lam = np.asarray(self.lambda_nm, dtype=np.float64)
        if self.weight_nm is not None:
            return np.asarray(self.weight_nm, dtype=np.float64)
        n = len(lam)
        return np.full(n, (lam[-1] - lam[0]) / (n - 1) if n > 1 else 1.0)


def pcg4d(v) -> tuple:
    a = (C.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in v])
    o = (C.c_uint32 * 4)()
    lib().orc_pcg4d(a, o)
    return tuple(int(x) for x in o)


def rng_u32(pixel, sample, band, dim, seed) -> int:
    return int(lib().orc_rng_u32(pixel, sample, band, dim, seed))


def rng_f32(pixel, sample, band, dim, seed) -> float:
    return float(lib().orc_rng_f32(pixel, sample, band, dim, seed))


def cmf_xyz(lam: float) -> np.ndarray:
    out = np.zeros(3)
    lib().orc_cmf_xyz(lam, _dp(out))
    return out


def planck(lam_nm: float, temperature_k: float) -> float:
    return float(lib().orc_planck_radiance(lam_nm, temperature_k))


def lambda_table(p: Params) -> np.ndarray:
    """Per-wavelength [beta_r, beta_o, beta_as, beta_aa, L_sun, wx, wy, wz]."""
    cp = p.to_c()
    arr = (_Lambda * p.n_lambda)()
    if lib().orc_lambda_table(C.byref(cp), arr):
        raise ValueError("invalid wavelength set")
    return np.array([[getattr(a, n) for n, _ in _Lambda._fields_] for a in arr])


def _lambda_struct(p: Params, band: int):
    cp = p.to_c()
    arr = (_Lambda * p.n_lambda)()
    lib().orc_lambda_table(C.byref(cp), arr)
    return cp, arr[band]


def coefficients(p: Params, band: int, h: float) -> np.ndarray:
    """[sigma_R, sigma_As, sigma_O, sigma_Aa] (km^-1) of band at altitude h."""
    cp = p.to_c()
    out = [C.c_double() for _ in range(4)]
    lib().orc_coefficients_band(C.byref(cp), band, h, *[C.byref(o) for o in out])
    return np.array([o.value for o in out])


def shell_heights(p: Params) -> np.ndarray:
    """Shell boundary heights h_0..h_L (km)."""
    cp = p.to_c()
    f = lib().orc_shell_height
    f.restype = C.c_double
    f.argtypes = [C.POINTER(_Params), C.c_int]
    return np.array([f(C.byref(cp), k) for k in range(p.majorant_layers + 1)])


def majorant_grid(p: Params) -> np.ndarray:
    cp = p.to_c()
    out = np.zeros((p.n_lambda, p.majorant_layers))
    if lib().orc_majorant_grid(C.byref(cp), _dp(out)):
        raise ValueError("invalid parameters")
    return out


def sun_direction(p: Params) -> np.ndarray:
    cp = p.to_c()
    out = np.zeros(3)
    lib().orc_sun_direction(C.byref(cp), _dp(out))
    return out


def next_boundary_reference(p: Params, o, d):
    cp = p.to_c()
    o = np.ascontiguousarray(o, dtype=np.float64)
    d = np.ascontiguousarray(d, dtype=np.float64)
    t = C.c_double()
    k = C.c_int()
    hit = lib().orc_next_boundary_reference(C.byref(cp), _dp(o), _dp(d), C.byref(t), C.byref(k))
    return bool(hit), t.value, k.value


def segments(p: Params, o, d, t_max: float, reference: bool):
    cp = p.to_c()
    o = np.ascontiguousarray(o, dtype=np.float64)
    d = np.ascontiguousarray(d, dtype=np.float64)
    cap = 4096
    ends = np.zeros(cap)
    layers = (C.c_int * cap)()
    fn = lib().orc_segments_reference if reference else lib().orc_segments_walk
    n = fn(C.byref(cp), _dp(o), _dp(d), t_max, _dp(ends), layers, cap)
    return ends[:n].copy(), np.array(layers[:n])


def optical_depth(p: Params, band: int, o, d, t_max: float) -> float:
    cp = p.to_c()
    o = np.ascontiguousarray(o, dtype=np.float64)
    d = np.ascontiguousarray(d, dtype=np.float64)
    return float(lib().orc_optical_depth_band(C.byref(cp), band, _dp(o), _dp(d), t_max))


def optical_depth_batch(p: Params, band: int, o, d, t_max) -> np.ndarray:
    """optical_depth for many rays: o, d [n, 3], t_max [n] -> tau [n]."""
    cp = p.to_c()
    o = np.ascontiguousarray(np.reshape(o, (-1, 3)), dtype=np.float64)
    d = np.ascontiguousarray(np.reshape(d, (-1, 3)), dtype=np.float64)
    t = np.ascontiguousarray(np.ravel(t_max), dtype=np.float64)
    out = np.zeros(t.size)
    if lib().orc_optical_depth_batch(C.byref(cp), band, t.size, _dp(o), _dp(d), _dp(t), _dp(out)):
        raise ValueError("invalid band or parameters")
    return out


def camera_dir(p: Params, u: float, v: float) -> np.ndarray:
    cp = p.to_c()
    out = np.zeros(3)
    lib().orc_camera_dir(C.byref(cp), u, v, _dp(out))
    return out


def render(p: Params, precision: str = "f32", threads: int = 0, want_m2: bool = False):
    """Returns (film [H,W,nl], film_m2 or None, xyz [H,W,3], counters dict, status)."""
    if threads <= 0:
        threads = os.cpu_count() or 1
    cp = p.to_c()
    nl = p.n_lambda
    film = np.zeros((p.height, p.width, nl))
    m2 = np.zeros((p.height, p.width, nl)) if want_m2 else None
    xyz = np.zeros((p.height, p.width, 3))
    cnt = Counters()
    fn = getattr(lib(), f"orc_render_{precision}")
    rc = fn(C.byref(cp), _dp(film), _dp(m2) if m2 is not None else None, _dp(xyz), C.byref(cnt), threads)
    return film, m2, xyz, cnt.as_dict(), rc


def render_pixels(p: Params, pixels, precision: str = "f32", threads: int = 0):
    """Renders listed pixel indices; returns (film [n,nl], film_m2 [n,nl], counters, status)."""
    if threads <= 0:
        threads = os.cpu_count() or 1
    cp = p.to_c()
    px = np.ascontiguousarray(pixels, dtype=np.uint32)
    nl = p.n_lambda
    film = np.zeros((len(px), nl))
    m2 = np.zeros((len(px), nl))
    cnt = Counters()
    fn = getattr(lib(), f"orc_render_pixels_{precision}")
    rc = fn(C.byref(cp), px.ctypes.data_as(C.POINTER(C.c_uint32)), len(px), _dp(film), _dp(m2), C.byref(cnt), threads)
    return film, m2, cnt.as_dict(), rc


def trace_path(p: Params, x: int, y: int, sample: int, band: int, precision: str = "f32"):
    cp = p.to_c()
    cnt = Counters()
    fn = getattr(lib(), f"orc_trace_path_{precision}")
    L = fn(C.byref(cp), x, y, sample, band, C.byref(cnt))
    return float(L), cnt.as_dict()


def ray_setup_f32(p: Params, o, d, e: float = 0.0):
    """The FP32 tracer's ray set-up: (dict of hit, kind, t_max, layer, b, q,
    w_o; whether the FP64 grazing branch ran)."""
    cp = p.to_c()
    o = np.ascontiguousarray(o, dtype=np.float64)
    d = np.ascontiguousarray(d, dtype=np.float64)
    out = np.zeros(7)
    g = lib().orc_ray_setup_f32(C.byref(cp), _dp(o), _dp(d), float(e), _dp(out))
    if g < 0:
        raise ValueError("invalid configuration")
    keys = ("hit", "kind", "t_max", "layer", "b", "q", "wo")
    r = dict(zip(keys, out.tolist()))
    for k in ("hit", "kind", "layer"):
        r[k] = int(r[k])
    return r, bool(g)


def transmittance_mc(p: Params, band: int, o, d, n: int, ratio: bool) -> float:
    cp = p.to_c()
    o = np.ascontiguousarray(o, dtype=np.float64)
    d = np.ascontiguousarray(d, dtype=np.float64)
    return float(lib().orc_transmittance_mc_f32(C.byref(cp), band, _dp(o), _dp(d), n, int(ratio)))


def sample_rayleigh_mu(xi: float) -> float:
    return float(lib().orc_sample_rayleigh_mu(xi))


def sample_hg_mu(g: float, xi: float) -> float:
    return float(lib().orc_sample_hg_mu(g, xi))


def phase_rayleigh(mu: float) -> float:
    return float(lib().orc_phase_rayleigh(mu))


def phase_hg(g: float, mu: float) -> float:
    return float(lib().orc_phase_hg(g, mu))


def frame_dir(axis, mu: float, st: float, phi: float) -> np.ndarray:
    a = np.ascontiguousarray(axis, dtype=np.float64)
    out = np.zeros(3)
    lib().orc_frame_dir(_dp(a), mu, st, phi, _dp(out))
    return out


def sun_cone_dir(p: Params, xi0: float, xi1: float):
    """Returns (direction, pdf, inside_cone)."""
    cp = p.to_c()
    out = np.zeros(3)
    pdf = float(lib().orc_sun_cone_dir(C.byref(cp), xi0, xi1, _dp(out)))
    return out, abs(pdf), pdf > 0


def cosine_dir(n, u: float, xi: float):
    a = np.ascontiguousarray(n, dtype=np.float64)
    out = np.zeros(3)
    pdf = float(lib().orc_cosine_dir(_dp(a), u, xi, _dp(out)))
    return out, pdf


# ------------------------------------------------ sparse-wavelength combo scan
# (crates/sky-wavelength-optimizer/src/main.rs, restated in scan_oracle.c)

SCAN_VORONOI, SCAN_CMF_MASS, SCAN_CMF_MOMENTS = 0, 1, 2
SCAN_FIELDS = ("mean", "rmse", "p50", "p90", "p95", "p99", "max", "score")


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def band_widths(wl) -> np.ndarray:
    w = np.ascontiguousarray(wl, dtype=np.float32)
    out = np.zeros(w.size)
    lib().orc_band_widths(_fp(w), w.size, _dp(out))
    return out


def quadrature_weights(wl, combo, strategy, cie=None) -> np.ndarray:
    """main.rs:735-1070 for one combo -> [count] weights (nm)."""
    w = np.ascontiguousarray(wl, dtype=np.float32)
    c = np.ascontiguousarray(combo, dtype=np.uint8)
    ci = np.ascontiguousarray(np.zeros((w.size, 3)) if cie is None else cie, dtype=np.float64)
    out = np.zeros(8)
    if lib().orc_quadrature_weights(_fp(w), w.size, _dp(ci), c.ctypes.data, c.size, strategy, _dp(out)):
        raise ValueError("invalid combo")
    return out[:c.size]


def xyz_to_oklab(xyz) -> np.ndarray:
    x = np.ascontiguousarray(xyz, dtype=np.float64)
    out = np.zeros(3)
    lib().orc_xyz_to_oklab(_dp(x), _dp(out))
    return out


def percentile(values, q: float) -> float:
    v = np.ascontiguousarray(values, dtype=np.float64)
    return float(lib().orc_percentile(_dp(v), v.size, q))


def reference_oklab(values, cie):
    """main.rs:276-341 -> (reference Oklab [n, 3], p99 Y scale)."""
    v = np.ascontiguousarray(values, dtype=np.float32)
    ci = np.ascontiguousarray(cie, dtype=np.float64)
    out = np.zeros((v.shape[0], 3))
    scale = lib().orc_reference_oklab(_fp(v), v.shape[1], v.shape[0], _dp(ci), _dp(out))
    return out, float(scale)


def evaluate_combos(values, ref_oklab, scale, wl, cie, combos, weights, idx=None, exact=False,
                    threads=0) -> np.ndarray:
    """main.rs:414-488 per combo -> [n_combos, 8] (SCAN_FIELDS), on `threads`
    host threads (0: OpenMP's default)."""
    v = np.ascontiguousarray(values, dtype=np.float32)
    r = np.ascontiguousarray(ref_oklab, dtype=np.float64)
    w = np.ascontiguousarray(wl, dtype=np.float32)
    ci = np.ascontiguousarray(cie, dtype=np.float64)
    cb = np.ascontiguousarray(np.atleast_2d(combos), dtype=np.uint8)
    wt = np.ascontiguousarray(np.atleast_2d(weights), dtype=np.float64)
    ix = None if idx is None else np.ascontiguousarray(idx, dtype=np.uint32)
    out = np.zeros((cb.shape[0], 8))
    lib().orc_evaluate_combos(_fp(v), v.shape[1], v.shape[0], _dp(r), scale, _fp(w), _dp(ci), cb.ctypes.data,
                              cb.shape[1], cb.shape[0], _dp(wt), None if ix is None else ix.ctypes.data,
                              0 if ix is None else ix.size, int(exact), int(threads), _dp(out))
    return out
