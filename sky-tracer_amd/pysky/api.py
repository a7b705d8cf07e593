"""Python mirror of the reference's integrator interface.

Mirrors crates/sky-spectral-path-tracer/src/{config.rs, film.rs, integrator.rs}
and the scene constructor of crates/sky-core/src/data.rs:15-51:

    scene = load_scene(sun_elevation_deg, sun_azimuth_deg)     # load_scene_data
    film  = render(scene, RenderConfig(width=..., spp=...))     # integrator::render

`render` raises a RenderError subclass with the reference's variants
(integrator.rs:19-26).  Everything is computed by libsky_hip.so on the GPU.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _lib as L

# ----------------------------------------------------------------- errors


class RenderError(Exception):
    """integrator.rs:19-26."""

    kind = "RenderError"

    def __init__(self, message: str = ""):
        super().__init__(message)
        self.message = message


class InvalidConfig(RenderError):
    kind = "InvalidConfig"


class NoAdapter(RenderError):
    kind = "NoAdapter"


class RequestDevice(RenderError):
    kind = "RequestDevice"


class BufferMap(RenderError):
    kind = "BufferMap"


class Poll(RenderError):
    kind = "Poll"


class Watchdog(RenderError):
    kind = "Watchdog"


class OutOfMemory(RenderError):
    kind = "OutOfMemory"


_ERRORS = {
    L.SKY_ERR_INVALID_CONFIG: InvalidConfig,
    L.SKY_ERR_NO_DEVICE: NoAdapter,
    L.SKY_ERR_DEVICE_INIT: RequestDevice,
    L.SKY_ERR_COPY: BufferMap,
    L.SKY_ERR_SYNC: Poll,
    L.SKY_ERR_WATCHDOG: Watchdog,
    L.SKY_ERR_OOM: OutOfMemory,
}


def check(status: int) -> None:
    if status != L.SKY_OK:
        raise _ERRORS.get(status, RenderError)(L.last_error())


# ----------------------------------------------------------------- config


class OutputProjection(enum.IntEnum):
    """config.rs:14-36."""

    Panorama = 0
    SkyViewLut = 1

    def label(self) -> str:
        return "panorama" if self is OutputProjection.Panorama else "sky-view LUT"


@dataclass
class RenderConfig:
    """config.rs:2-12 with RenderConfig::default (config.rs:38-52), plus the
    extensions of this integrator (defaults reproduce the reference)."""

    width: int = 1024
    height: int = 512
    spp: int = 256
    seed: int = 0x5EC7_2026_0430
    sun_elevation_deg: float = 0.0
    sun_azimuth_deg: float = 0.0
    observer_altitude_km: float = 0.2
    direct_light_samples: int = 1
    output_projection: OutputProjection = OutputProjection.Panorama
    # extensions
    sample_offset: int = 0
    max_depth: int = 0  # 0 = uncapped (reference)
    rr_start_depth: int = 4
    watchdog_limit: int = 1024
    max_segments: int = 100000
    majorant_layers: int = 64
    majorant_margin: float = 1.05
    phase_importance_sampling: bool = False
    ratio_tracking: bool = False
    carried_tau: bool = False
    simple_kernel: bool = False
    # optional fixed majorant grid [bands, layers] (km^-1)
    majorant_table: Optional[np.ndarray] = None
    # Majorant shell thickness growth from the ground up (1: uniform shells,
    # the reference's); any value samples the same free-flight law.
    shell_growth: float = 1.0

    @property
    def flags(self) -> int:
        f = 0
        if self.phase_importance_sampling:
            f |= L.FLAG_PHASE_IS
        if self.ratio_tracking:
            f |= L.FLAG_RATIO_TRACKING
        if self.carried_tau:
            f |= L.FLAG_CARRIED_TAU
        if self.simple_kernel:
            f |= L.FLAG_SIMPLE_KERNEL
        return f


# ------------------------------------------------------------------ scene


@dataclass
class AtmosphereParams:
    """Analytic medium (replaces the tabulated profiles of atmosphere.rs:121-133)."""

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
    # Tabulated medium (pysky.tables.MediumTables, e.g. from
    # pysky.tables.load_reference_scene); None: the analytic profiles above.
    tables: Optional[object] = None

    def to_c(self) -> L.Atmosphere:
        a = L.Atmosphere()
        for name, _ in L.Atmosphere._fields_:
            if name != "tables":
                setattr(a, name, float(getattr(self, name)))
        if self.tables is not None:
            tc = self.tables.to_c()
            a.tables = C.cast(C.pointer(tc), C.c_void_p)
            a._keep = tc
        return a


# The ten parameters the adjoint differentiates, in gradient order.
PARAM_NAMES = (
    "rayleigh_sca_550_m_inv",
    "rayleigh_h_km",
    "ozone_peak_m_inv",
    "ozone_center_km",
    "aerosol_sca_550_m_inv",
    "aerosol_abs_550_m_inv",
    "aerosol_h_km",
    "aerosol_g",
    "angstrom",
    "ground_albedo",
)


def independent_seed(k: int) -> int:
    """Seed of the k-th of a set of statistically independent renders.

    A path's random numbers are keyed by (pixel ^ seed_lo, sample ^ seed_hi,
    band, dimension) (sky_trace.wgsl:799-800), so seeds that differ only in
    their low bits hand the same random streams to nearby pixels or samples
    (seed 2 gives pixel p the stream seed 1 gives pixel p ^ 3).  k << 48 puts
    seed_hi = k * 65536: every render gets sample keys of its own for up to
    65536 samples per pixel."""
    return int(k) << 48


def midpoint_grid(n: int, lo: float = 380.0, hi: float = 780.0) -> np.ndarray:
    """Midpoint wavelengths: 16 -> 392.5 + 25 i nm, 64 -> 383.125 + 6.25 i nm."""
    step = (hi - lo) / n
    return lo + step * (np.arange(n) + 0.5)


@dataclass
class SceneData:
    """Scene of one render (mirrors sky_core::atmosphere::SceneData):
    planet, analytic atmosphere, sun direction and the wavelength set."""

    ground_radius_km: float = 6360.0
    atmosphere_radius_km: float = 6460.0
    atmosphere: AtmosphereParams = field(default_factory=AtmosphereParams)
    sun_elevation_deg: float = 0.0
    sun_azimuth_deg: float = 0.0
    wavelengths_nm: Sequence[float] = field(default_factory=lambda: list(midpoint_grid(16)))
    weights_nm: Optional[Sequence[float]] = None

    @property
    def band_count(self) -> int:
        return len(self.wavelengths_nm)

    def planet_c(self) -> L.Planet:
        p = L.Planet()
        p.ground_radius_km = self.ground_radius_km
        p.atmosphere_radius_km = self.atmosphere_radius_km
        return p

    def wavelengths_c(self):
        lam = np.ascontiguousarray(self.wavelengths_nm, dtype=np.float64)
        w = None if self.weights_nm is None else np.ascontiguousarray(self.weights_nm, dtype=np.float64)
        wl = L.Wavelengths()
        wl.count = len(lam)
        wl.center_nm = lam.ctypes.data_as(C.POINTER(C.c_double))
        wl.weight_nm = w.ctypes.data_as(C.POINTER(C.c_double)) if w is not None else None
        wl._keep = (lam, w)
        return wl


def load_scene(sun_elevation_deg: float = 0.0, sun_azimuth_deg: float = 0.0,
               wavelengths_nm: Optional[Sequence[float]] = None,
               atmosphere: Optional[AtmosphereParams] = None, **kw) -> SceneData:
    """Counterpart of sky_core::data::load_scene_data (data.rs:15-51) for the
    analytic atmosphere; `wavelengths_nm` defaults to 16 midpoint bands."""
    return SceneData(
        atmosphere=atmosphere or AtmosphereParams(),
        sun_elevation_deg=sun_elevation_deg,
        sun_azimuth_deg=sun_azimuth_deg,
        wavelengths_nm=list(midpoint_grid(16)) if wavelengths_nm is None else list(wavelengths_nm),
        **kw,
    )


def config_c(scene: SceneData, cfg: RenderConfig) -> L.RenderConfigC:
    c = L.RenderConfigC()
    c.width, c.height, c.spp = cfg.width, cfg.height, cfg.spp
    c.sample_offset = cfg.sample_offset
    c.seed = cfg.seed & 0xFFFFFFFFFFFFFFFF
    # the sun comes from the scene, as in PackedScene::new (integrator.rs:343-345)
    c.sun_elevation_deg = scene.sun_elevation_deg
    c.sun_azimuth_deg = scene.sun_azimuth_deg
    c.observer_altitude_km = cfg.observer_altitude_km
    c.direct_light_samples = cfg.direct_light_samples
    c.projection = int(cfg.output_projection)
    c.max_depth = cfg.max_depth
    c.rr_start_depth = cfg.rr_start_depth
    c.watchdog_limit = cfg.watchdog_limit
    c.max_segments = cfg.max_segments
    c.majorant_layers = cfg.majorant_layers
    c.flags = cfg.flags
    c.majorant_margin = cfg.majorant_margin
    c.shell_growth = cfg.shell_growth
    if cfg.majorant_table is not None:
        t = np.ascontiguousarray(cfg.majorant_table, dtype=np.float32).ravel()
        if t.size != scene.band_count * cfg.majorant_layers:
            raise InvalidConfig(f"majorant table has {t.size} entries, expected "
                                f"{scene.band_count * cfg.majorant_layers}")
        c.majorant_table = t.ctypes.data_as(C.POINTER(C.c_float))
        c._keep = t
    return c


# ------------------------------------------------------------------- film


class Film:
    """film.rs:2-62: width x height x band_count radiance values, pixel-major
    ([(y * width + x) * band_count + band]).  `xyz` holds the in-kernel CIE
    XYZ image ([(y * width + x) * 3 + c]) and `m2` the mean squared path
    radiance when they were requested."""

    def __init__(self, width: int, height: int, band_count: int, values: np.ndarray,
                 xyz: Optional[np.ndarray] = None, m2: Optional[np.ndarray] = None, stats=None):
        assert values is None or values.size == width * height * band_count
        self._w, self._h, self._b = width, height, band_count
        self._values = values
        self.xyz = xyz
        self.m2 = m2
        self.stats = stats

    def width(self) -> int:
        return self._w

    def height(self) -> int:
        return self._h

    def band_count(self) -> int:
        return self._b

    def values(self) -> np.ndarray:
        return self._values

    def pixel_spectrum(self, pixel: int) -> np.ndarray:
        return self._values[pixel * self._b:(pixel + 1) * self._b]

    def set_pixel_spectrum(self, pixel: int, spectrum) -> None:
        assert len(spectrum) == self._b
        self._values[pixel * self._b:(pixel + 1) * self._b] = spectrum

    def as_array(self) -> np.ndarray:
        return self._values.reshape(self._h, self._w, self._b)


# ---------------------------------------------------------------- context


class Context:
    """Owns one libsky_hip context (device, stream, cached device buffers)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        check(L.lib().sky_context_create(device, C.byref(self._h)))

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def set_tuning(self, waves_per_cu=0, pixels_per_item=0, event_threshold=0, track_steps=0, occupancy=0):
        t = L.Tuning(waves_per_cu, pixels_per_item, event_threshold, track_steps, occupancy)
        check(L.lib().sky_context_set_tuning(self._h, C.byref(t)))

    def close(self):
        if self._h:
            L.lib().sky_context_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: dict = {}


def default_context(device: int = 0) -> Context:
    if device not in _default_ctx:
        _default_ctx[device] = Context(device)
    return _default_ctx[device]


def _f32p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def render(scene: SceneData, config: RenderConfig, *, want_film: bool = True, want_xyz: bool = True,
           want_m2: bool = False, context: Optional[Context] = None) -> Film:
    """integrator::render (integrator.rs:125-128).  Returns a Film whose values
    are normalised by 1/spp and clamped >= 0 (integrator.rs:307-310)."""
    nb = scene.band_count
    npx = config.width * config.height
    planet, atm, wl = scene.planet_c(), scene.atmosphere.to_c(), scene.wavelengths_c()
    cfg = config_c(scene, config)
    lib = L.lib()
    check(lib.sky_validate_config(C.byref(planet), C.byref(atm), C.byref(wl), C.byref(cfg)))
    film = np.zeros(npx * nb, dtype=np.float32) if want_film else None
    m2 = np.zeros(npx * nb, dtype=np.float32) if want_m2 else None
    xyz = np.zeros(npx * 3, dtype=np.float32) if want_xyz else None
    stats = L.RenderStats()
    if context is not None:
        ctx = context.handle
    else:
        # The library validates before it looks at the device, so a bad
        # configuration reports InvalidConfig even without a GPU, as
        # validate_config runs before the adapter request (integrator.rs:126-127).
        try:
            ctx = default_context().handle
        except NoAdapter:
            ctx = None
    check(lib.sky_render(ctx, C.byref(planet), C.byref(atm), C.byref(wl), C.byref(cfg),
                         _f32p(film), _f32p(m2), _f32p(xyz), C.byref(stats)))
    return Film(config.width, config.height, nb, film, xyz=xyz, m2=m2, stats=stats)


def optical_depth(scene: SceneData, rays: np.ndarray, context: Optional[Context] = None) -> np.ndarray:
    """Deterministic optical depth (GPU Gauss-Legendre) for rays [n, 7] =
    (origin xyz km, direction xyz, t_max km); returns [n, band_count]."""
    rays = np.ascontiguousarray(rays, dtype=np.float64).reshape(-1, 7)
    out = np.zeros((rays.shape[0], scene.band_count), dtype=np.float64)
    planet, atm, wl = scene.planet_c(), scene.atmosphere.to_c(), scene.wavelengths_c()
    ctx = (context or default_context()).handle
    check(L.lib().sky_optical_depth(ctx, C.byref(planet), C.byref(atm), C.byref(wl),
                                    rays.ctypes.data_as(C.POINTER(C.c_double)), rays.shape[0],
                                    out.ctypes.data_as(C.POINTER(C.c_double))))
    return out


def wavelength_table(scene: SceneData) -> np.ndarray:
    """Per-band [beta_r, beta_o, beta_as, beta_aa, L_sun, wx, wy, wz] (host)."""
    nb = scene.band_count
    out = np.zeros(8 * nb, dtype=np.float32)
    atm, wl = scene.atmosphere.to_c(), scene.wavelengths_c()
    check(L.lib().sky_wavelength_table(C.byref(atm), C.byref(wl), _f32p(out), out.size))
    return out.reshape(nb, 8)


def majorant_grid(scene: SceneData, layers: int = 64, margin: float = 1.05, shell_growth: float = 1.0) -> np.ndarray:
    """MajorantGrid values (medium.rs:63-83), [band, layer] in km^-1 (host),
    over uniform shells or shells growing by `shell_growth`."""
    nb = scene.band_count
    out = np.zeros(nb * layers, dtype=np.float32)
    planet, atm, wl = scene.planet_c(), scene.atmosphere.to_c(), scene.wavelengths_c()
    check(L.lib().sky_majorant_grid(C.byref(planet), C.byref(atm), C.byref(wl), layers, margin, shell_growth,
                                    _f32p(out), out.size))
    return out.reshape(nb, layers)
