"""pysky -- Python host of the MI355X spectral sky path tracer (libsky_hip.so).

Mirrors the reference's integrator interface (render(scene, config) -> Film,
RenderError variants) over the C ABI of include/sky_hip.h.
"""
from ._lib import LIB_PATH, HEADER_PATH, lib  # noqa: F401
from .api import (  # noqa: F401
    PARAM_NAMES,
    AtmosphereParams,
    BufferMap,
    Context,
    Film,
    InvalidConfig,
    NoAdapter,
    OutOfMemory,
    OutputProjection,
    Poll,
    RenderConfig,
    RenderError,
    RequestDevice,
    SceneData,
    Watchdog,
    default_context,
    load_scene,
    majorant_grid,
    independent_seed,
    midpoint_grid,
    optical_depth,
    render,
    wavelength_table,
)
from .batch import BatchResult, ragged_batch, render_batch  # noqa: F401
from .adjoint import fixed_majorant, params_vector, render_adjoint, with_params  # noqa: F401
