"""Device-buffer entry points for torch tensors (sky_render_device).

Torch only supplies device memory and the stream; the integrator is the HIP
library.  Outputs are float32 CUDA tensors on the context's device.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import torch

from . import _lib as L
from .api import Context, RenderConfig, SceneData, check, config_c, default_context


def _ptr(t: Optional[torch.Tensor]):
    if t is None:
        return None
    assert t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()
    return C.c_void_p(t.data_ptr())


def render_device(scene: SceneData, config: RenderConfig, xyz: Optional[torch.Tensor] = None,
                  film: Optional[torch.Tensor] = None, film_m2: Optional[torch.Tensor] = None,
                  context: Optional[Context] = None) -> L.RenderStats:
    """Renders into caller-owned device tensors (xyz [H, W, 3], film and
    film_m2 [H, W, bands]); returns the launch statistics."""
    npx = config.width * config.height
    for t, n in ((xyz, 3 * npx), (film, scene.band_count * npx), (film_m2, scene.band_count * npx)):
        if t is not None and t.numel() != n:
            raise ValueError(f"output tensor has {t.numel()} elements, expected {n}")
    ctx = (context or default_context(torch.cuda.current_device())).handle
    planet, atm, wl = scene.planet_c(), scene.atmosphere.to_c(), scene.wavelengths_c()
    cfg = config_c(scene, config)
    stats = L.RenderStats()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(L.lib().sky_render_device(ctx, C.byref(planet), C.byref(atm), C.byref(wl), C.byref(cfg),
                                    _ptr(film), _ptr(film_m2), _ptr(xyz), stream, C.byref(stats)))
    return stats
