"""ctypes binding of libsky_hip.so (C ABI declared in include/sky_hip.h).

The library is built in-tree (sky-tracer_amd/lib/libsky_hip.so).  There is no
fallback: if the library is missing every call raises, so the HIP path is the
only path.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_PKG_DIR, "..", "lib", "libsky_hip.so"))
HEADER_PATH = os.path.normpath(os.path.join(_PKG_DIR, "..", "..", "include", "sky_hip.h"))

SKY_OK = 0
SKY_ERR_INVALID_CONFIG = 1
SKY_ERR_NO_DEVICE = 2
SKY_ERR_DEVICE_INIT = 3
SKY_ERR_COPY = 4
SKY_ERR_SYNC = 5
SKY_ERR_WATCHDOG = 6
SKY_ERR_OOM = 7

FLAG_PHASE_IS = 1
FLAG_RATIO_TRACKING = 2
FLAG_CARRIED_TAU = 4
FLAG_SIMPLE_KERNEL = 256


class Planet(C.Structure):
    _fields_ = [("ground_radius_km", C.c_double), ("atmosphere_radius_km", C.c_double)]


class Atmosphere(C.Structure):
    _fields_ = [
        (n, C.c_double)
        for n in (
            "rayleigh_sca_550_m_inv",
            "rayleigh_h_km",
            "ozone_peak_m_inv",
            "ozone_center_km",
            "ozone_half_width_km",
            "ozone_lambda0_nm",
            "ozone_sigma_nm",
            "aerosol_sca_550_m_inv",
            "aerosol_abs_550_m_inv",
            "aerosol_h_km",
            "aerosol_g",
            "angstrom",
            "ground_albedo",
            "sun_temperature_k",
            "sun_angular_radius_rad",
        )
    ] + [("tables", C.c_void_p)]


class MediumTablesC(C.Structure):
    """sky_medium_tables (include/sky_hip.h)."""
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


class Wavelengths(C.Structure):
    _fields_ = [
        ("count", C.c_uint32),
        ("center_nm", C.POINTER(C.c_double)),
        ("weight_nm", C.POINTER(C.c_double)),
    ]


class RenderConfigC(C.Structure):
    _fields_ = [
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
        ("majorant_table", C.POINTER(C.c_float)),
        ("shell_growth", C.c_double),
    ]


class RenderStats(C.Structure):
    _fields_ = [
        ("kernel_ms", C.c_double),
        ("total_ms", C.c_double),
        ("path_samples", C.c_uint64),
        ("watchdog", C.c_uint32),
        ("workgroups", C.c_uint32),
        ("block", C.c_uint32),
        ("items", C.c_uint32),
        ("adjoint_value", C.c_double),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("waves_per_cu", C.c_uint32),
        ("pixels_per_item", C.c_uint32),
        ("event_threshold", C.c_uint32),
        ("track_steps", C.c_uint32),
        ("occupancy", C.c_uint32),
    ]


class BatchItem(C.Structure):
    _fields_ = [
        ("planet", C.POINTER(Planet)),
        ("atm", C.POINTER(Atmosphere)),
        ("wl", C.POINTER(Wavelengths)),
        ("cfg", C.POINTER(RenderConfigC)),
        ("xyz", C.POINTER(C.c_float)),
        ("film", C.POINTER(C.c_float)),
    ]


class BatchStats(C.Structure):
    _fields_ = [
        ("kernel_ms", C.c_double),
        ("total_ms", C.c_double),
        ("path_samples", C.c_uint64),
        ("renders", C.c_uint32),
        ("items", C.c_uint32),
        ("watchdog", C.c_uint32),
        ("graph_reused", C.c_uint32),
        ("latency_p50_ms", C.c_double),
        ("latency_p99_ms", C.c_double),
        ("latency_max_ms", C.c_double),
    ]


BATCH_NO_GRAPH = 1

class ScanStats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("total_ms", C.c_double), ("evaluations", C.c_uint64)]


class ComboStats(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("mean", "rmse", "p50", "p90", "p95", "p99", "max")] + [
        ("sample_count", C.c_uint64)]


_lib = None


def lib() -> C.CDLL:
    """Loads libsky_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libsky_hip.so not found at {LIB_PATH}: build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` (or make -C sky-tracer_amd)"
        )
    L = C.CDLL(LIB_PATH)
    ctxp = C.c_void_p
    P, A, W, R = C.POINTER(Planet), C.POINTER(Atmosphere), C.POINTER(Wavelengths), C.POINTER(RenderConfigC)
    fp = C.POINTER(C.c_float)
    dp = C.POINTER(C.c_double)
    sig = {
        "sky_context_create": ([C.c_int, C.POINTER(ctxp)], C.c_int),
        "sky_context_destroy": ([ctxp], None),
        "sky_context_set_tuning": ([ctxp, C.POINTER(Tuning)], C.c_int),
        "sky_context_get_tuning": ([ctxp, C.POINTER(Tuning)], C.c_int),
        "sky_default_planet": ([P], None),
        "sky_default_atmosphere": ([A], None),
        "sky_default_config": ([R], None),
        "sky_wavelength_table": ([A, W, fp, C.c_uint32], C.c_int),
        "sky_majorant_grid": ([P, A, W, C.c_uint32, C.c_double, C.c_double, fp, C.c_uint32], C.c_int),
        "sky_validate_config": ([P, A, W, R], C.c_int),
        "sky_render": ([ctxp, P, A, W, R, fp, fp, fp, C.POINTER(RenderStats)], C.c_int),
        "sky_render_device": ([ctxp, P, A, W, R, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.POINTER(RenderStats)], C.c_int),
        "sky_render_device_noalloc": ([ctxp, P, A, W, R, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.POINTER(RenderStats)], C.c_int),
        "sky_optical_depth": ([ctxp, P, A, W, dp, C.c_uint32, dp], C.c_int),
        "sky_render_adjoint": ([ctxp, P, A, W, R, fp, dp, C.POINTER(RenderStats)], C.c_int),
        "sky_render_adjoint_device": ([ctxp, P, A, W, R, C.c_void_p, dp, C.c_void_p, C.POINTER(RenderStats)],
                                      C.c_int),
        "sky_render_batch": ([ctxp, C.POINTER(BatchItem), C.c_uint32, C.c_uint32, dp, C.POINTER(BatchStats)],
                             C.c_int),
        "sky_scan_dataset_create": ([C.c_int, fp, C.c_uint32, C.c_uint64, fp, C.POINTER(ctxp)], C.c_int),
        "sky_scan_dataset_destroy": ([ctxp], None),
        "sky_scan_combos": ([ctxp, C.c_void_p, fp, C.c_uint32, C.c_uint64, C.c_void_p, C.c_uint64, dp, dp,
                             C.POINTER(ScanStats)], C.c_int),
        "sky_scan_combos_device": ([ctxp, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint64, C.c_void_p, C.c_uint64,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(ScanStats)], C.c_int),
        "sky_refine_combos": ([ctxp, C.c_void_p, fp, C.c_uint32, C.c_uint64, C.c_void_p, C.c_uint64,
                               C.POINTER(ComboStats), C.POINTER(ScanStats)], C.c_int),
        "sky_band_widths": ([fp, C.c_uint32, dp], C.c_int),
        "sky_quadrature_weights": ([fp, C.c_uint32, dp, C.c_void_p, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32,
                                    dp], C.c_int),
        "sky_cie_1931_2deg": ([fp, C.c_uint32, dp], C.c_int),
        "sky_last_error": ([], C.c_char_p),
        "sky_version": ([], C.c_char_p),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def last_error() -> str:
    msg = lib().sky_last_error()
    return msg.decode() if msg else ""
